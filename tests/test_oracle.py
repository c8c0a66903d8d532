"""Pins the FP64 oracle: against the reference's golden fixtures (generated by the reference's own
Python specification, rust/crates/optflow-core/tests/fixtures), against numpy/scipy where a step is
a library routine, and against closed forms. Tolerances are those of the reference's Rust parity
tests (rust/crates/optflow-core/tests/parity_*.rs) unless a tighter one is stated."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from conftest import load, meta, rms
from optflow_soil_amd.config import Config


# ---------------------------------------------------------------- FFT (core/fft.rs:159-226)
@pytest.mark.parametrize("n", [1, 2, 7, 12, 64, 97, 320, 352, 1088])
def test_fft1_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert np.abs(O.fft1(x) - np.fft.fft(x)).max() < 1e-11 * max(1, n)
    assert np.abs(O.fft1(x, inverse=True) - np.fft.ifft(x)).max() < 1e-12 * max(1, n)


def test_fft_small_matches_naive_dft():
    rng = np.random.default_rng(1)
    for n in (3, 5, 6, 11, 13, 17):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        k = np.arange(n)
        naive = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
        assert np.abs(O.fft1(x) - naive).max() < 1e-12


def test_fft2_matches_numpy():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(66, 90)) + 1j * rng.normal(size=(66, 90))
    assert np.abs(O.fft2(a) - np.fft.fft2(a)).max() < 1e-10
    assert np.abs(O.fft2(a, inverse=True) - np.fft.ifft2(a)).max() < 1e-13


# ---------------------------------------------------------------- interp (parity_interp.rs:13-64)
def test_interp_fixtures():
    img = load("interp_img")
    x, y = load("interp_x"), load("interp_y")
    assert rms(O.sample(O.bspline_coeffs(img), x, y), load("interp_sample")) < 1e-6
    assert rms(O.warp(img, load("interp_warp_ux"), load("interp_warp_uy")), load("interp_warp")) < 1e-6
    for s in (1, 2, 4):
        assert np.abs(O.gauss_blur(img, s) - load(f"interp_blur_{s}")).max() < 1e-6


def test_spline_reproduces_grid_values():
    # pyt/test_interp.py:7: the spline through the samples passes through them
    rng = np.random.default_rng(3)
    img = rng.normal(size=(23, 31))
    yy, xx = np.mgrid[0:23, 0:31].astype(float)
    assert np.abs(O.sample(O.bspline_coeffs(img), xx, yy) - img).max() < 1e-10  # 21-sample horizon: z^21 ~ 1e-12


def test_blur_matches_scipy_correlate():
    rng = np.random.default_rng(4)
    a = rng.normal(size=(30, 41))
    r = int(4 * 2.5 + 0.5)
    t = np.arange(-r, r + 1)
    k = np.exp(-0.5 * (t / 2.5) ** 2)
    k /= k.sum()
    ref = ndimage.correlate1d(ndimage.correlate1d(a, k, axis=0, mode="nearest"), k, axis=1, mode="nearest")
    assert np.abs(O.gauss_blur(a, 2.5) - ref).max() < 1e-13


# ---------------------------------------------------------------- gabor (parity_gabor.rs:18-69)
def test_gabor_fixture():
    img = load("gabor_img")
    rr, ri, gx, gy, amp = O.reference_phase(img, Config())
    ref_re, ref_im = load("gabor_r_re"), load("gabor_r_im")
    gx_py, gy_py = load("gabor_gx"), load("gabor_gy")
    for k in range(12):
        a_py = np.sqrt(ref_re[k] ** 2 + ref_im[k] ** 2 + 1e-12)
        err = np.hypot(rr[k] - ref_re[k], ri[k] - ref_im[k]).max() / a_py.max()
        assert err < 1e-4
        strong = a_py > 0.05 * a_py.max()
        assert strong.sum() > a_py.size / 20
        assert np.abs(gx[k][strong] - gx_py[k][strong]).max() < 1e-3
        assert np.abs(gy[k][strong] - gy_py[k][strong]).max() < 1e-3
        assert np.abs(amp[k][strong] - a_py[strong]).max() < 1e-4 * a_py.max()


def test_plane_wave_phase_gradient_is_wave_vector():
    # parity_gabor.rs:133-165: a plane wave's local phase gradient is its wave vector
    # on a non-square image with sides that are not powers of two; the one-sided filter leaks
    # exp(-8) of the negative frequency, so the interior error is ~5e-4 rad/px (bound 1e-2 as there)
    cfg = Config()
    h, w = 45, 70
    yy, xx = np.mgrid[0:h, 0:w].astype(float)
    wave = 2 * np.pi / (cfg.lam * cfg.scales[1])
    for k in range(cfg.n_orient):
        th = np.pi * k / cfg.n_orient
        kx, ky = wave * np.cos(th), wave * np.sin(th)
        img = np.cos(kx * xx + ky * yy + 0.3)
        _, _, gx, gy, _ = O.reference_phase(img, cfg)
        c = np.s_[12:-12, 12:-12]
        f = cfg.n_orient + k  # the middle scale's filter of orientation k
        assert np.abs(gx[f][c] - kx).max() < 1e-2
        assert np.abs(gy[f][c] - ky).max() < 1e-2


def test_constant_image_has_no_response():
    rr, ri, gx, gy, amp = O.reference_phase(np.full((70, 70), 0.4), Config())
    assert np.abs(rr).max() < 1e-12 and np.abs(ri).max() < 1e-12


# ---------------------------------------------------------------- POC (parity_poc.rs:39-172)
def test_poc_shift_fixture():
    m = meta()["poc_shift"]
    a, b = load("poc_a"), load("poc_b")
    dx, dy, pk = O.poc_shift(a[32:96, 32:96], b[32:96, 32:96])
    assert abs(dx - m["dx"]) < 1e-5 and abs(dy - m["dy"]) < 1e-5 and abs(pk - m["peak"]) < 1e-5


def test_poc_shift_pure_translation():
    rng = np.random.default_rng(5)
    base = ndimage.gaussian_filter(rng.normal(size=(96, 96)), 1.5)
    a = base[10:74, 10:74]
    b = base[7:71, 15:79]  # b(x) = a(x - d) with d = (-5, 3)
    dx, dy, pk = O.poc_shift(a, b)
    # the Hann window leaves the parabolic sub-pixel step a small bias; the integer peak is exact
    assert round(dx) == -5 and round(dy) == 3 and abs(dx + 5) < 0.1 and abs(dy - 3) < 0.1 and pk > 0.5


@pytest.mark.parametrize("case", ["grid", "strip", "init", "fallback", "far"])
def test_coarse_flow_fixtures(case):
    m = meta()
    a, b = load("poc_a"), load("poc_b")
    cfg, init = Config(), None
    if case == "strip":
        a, b = load("poc_strip_a"), load("poc_strip_b")
        cfg = Config(max_disp=100)
        assert O.n_levels(100, 64, a.shape) == m["poc_strip"]["n_levels"]
    if case in ("init", "fallback"):
        init = (load("poc_init_guess_ux"), load("poc_init_guess_uy"))
    if case == "fallback":
        b = np.full(a.shape, 0.5)
    if case == "far":
        f = m["poc_far"]
        a, b = load("poc_strip_a"), load("poc_strip_b")
        yy, xx = np.mgrid[0:96, 0:512].astype(float)
        init = ((f["guess_x"][0] + f["guess_x"][1] * xx).astype(np.float32).astype(np.float64),
                (f["guess_y"][0] + f["guess_y"][1] * yy).astype(np.float32).astype(np.float64))
        cfg = Config(max_disp=f["max_disp"])
    ux, uy, peak = O.coarse_flow(a, b, cfg, init=init)
    key = {"grid": "poc_grid", "strip": "poc_strip", "init": "poc_init", "fallback": "poc_fallback",
           "far": "poc_far"}[case]
    assert ux.shape == load(key + "_ux").shape
    assert np.abs(ux - load(key + "_ux")).max() < 1e-3
    assert np.abs(uy - load(key + "_uy")).max() < 1e-3
    if case != "strip":
        assert np.abs(peak - load(key + "_peak")).max() < 1e-3


def test_level_window_and_n_levels_tables():
    m = meta()
    for h, w, win in m["level_window"]:
        assert O.level_window((h, w), 64) == (tuple(win) if win else None)
    for d, h, w, lev in m["n_levels"]:
        assert O.n_levels(d, 64, (h, w)) == lev


# ---------------------------------------------------------------- refine (parity_refine.rs:58-177)
@pytest.mark.parametrize("order", [1, 2])
def test_refine_fixture(order):
    ux, uy, conf, resid, _ = O.refine(load("refine_a"), load("refine_b"), load("refine_init_ux"),
                                      load("refine_init_uy"), Config(shape_order=order))
    # the reference's bound is 1e-3 RMS; the oracle reproduces the float32 fixtures to rounding
    for got, key in ((ux, "ux"), (uy, "uy"), (conf, "conf"), (resid, "resid")):
        assert rms(got, load(f"refine_o{order}_{key}")) < 1e-6


@pytest.mark.parametrize("order", [1, 2])
def test_refine_stripe_is_degenerate(order):
    init = meta()["refine_stripe"]["init"]
    a, b = load("refine_stripe_a"), load("refine_stripe_b")
    ux, uy, conf, resid, _ = O.refine(a, b, np.full(a.shape, init[0]), np.full(a.shape, init[1]),
                                      Config(shape_order=order))
    assert conf.max() < Config().conf_min
    assert rms(ux, load(f"refine_stripe_o{order}_ux")) < 1e-5
    assert rms(uy, load(f"refine_stripe_o{order}_uy")) < 1e-5
    assert rms(conf, load(f"refine_stripe_o{order}_conf")) < 1e-5


def test_refine_border_fixture():
    f = meta()["refine_border"]
    a, b = load("refine_border_a"), load("refine_border_b")
    # the initial field: the fixture's truth plus the stated offset (make_parity_fixtures.py:318-330)
    n = a.shape[0]
    yy, xx = np.mgrid[0:n, 0:n].astype(float)
    cx, cy = f["center"]
    gx = f["exx"] * (xx - cx) + f["t"][0]
    gy = f["eyy"] * (yy - cy) + f["t"][1]
    ux, uy, conf, _, _ = O.refine(a, b, gx + f["init_offset"][0], gy + f["init_offset"][1], Config(edge_guard=0.0))
    assert np.abs(ux - load("refine_border_ux")).max() < 1e-3
    assert np.abs(uy - load("refine_border_uy")).max() < 1e-3
    assert np.abs(conf - load("refine_border_conf")).max() < 1e-3


def test_local_solve_recovers_affine_field():
    # the LS with exact data (c = g . u) returns the field and zero residual
    rng = np.random.default_rng(6)
    h, w = 40, 50
    gx, gy = rng.normal(size=(2, 12, h, w))
    wk = rng.uniform(0.5, 1.5, size=(12, h, w))
    yy, xx = np.mgrid[0:h, 0:w].astype(float)
    ux = 0.3 + 0.01 * xx - 0.02 * yy
    uy = -0.7 + 0.015 * xx + 0.005 * yy
    c = gx * ux + gy * uy
    sums = np.stack([(wk * gx * gx).sum(0), (wk * gx * gy).sum(0), (wk * gy * gy).sum(0), (wk * gx * c).sum(0),
                     (wk * gy * c).sum(0), (wk * c * c).sum(0), wk.sum(0)])
    oux, ouy, en, resid, slope = O.local_solve(sums, np.zeros((h, w)), np.zeros((h, w)), 4.0, 1)
    c = np.s_[16:-16, 16:-16]  # beyond the window radius the replicated edges do not enter
    assert np.abs(oux - ux)[c].max() < 1e-9 and np.abs(ouy - uy)[c].max() < 1e-9
    assert np.abs(slope[..., 0] - 0.01)[c].max() < 1e-9 and np.abs(slope[..., 3] - 0.005)[c].max() < 1e-9
    assert resid[c].max() < 1e-5


# ---------------------------------------------------------------- EDT and components
def test_edt_fixtures_and_ties():
    v = load("edt_valid") > 0.5
    assert np.abs(O.edt(v) - load("edt_dist")).max() < 1e-5
    src, idx = load("edt_ft_src") > 0.5, load("edt_ft_idx").astype(np.int64)
    for k in range(4):
        _, got = O.edt(~src[k], return_indices=True)
        assert (got == idx[k]).all()


def test_edt_matches_scipy_random():
    rng = np.random.default_rng(7)
    for p in (0.01, 0.1, 0.5):
        b = rng.uniform(size=(45, 67)) > p
        d, idx = O.edt(b, return_indices=True)
        dr, (iy, ix) = ndimage.distance_transform_edt(b, return_indices=True)
        assert np.abs(d - dr).max() < 1e-12
        assert (idx == iy * 67 + ix).all()


def test_large_components_matches_scipy_label():
    rng = np.random.default_rng(8)
    b = ndimage.gaussian_filter(rng.normal(size=(80, 90)), 2) > 0.1
    lab, n = ndimage.label(b)
    sizes = np.bincount(lab.ravel())
    keep = sizes >= 40
    keep[0] = False
    assert (O.large_components(b, 40) == keep[lab]).all()


# ---------------------------------------------------------------- strain (parity_strain.rs:12-100)
def test_strain_fixture():
    g = O.displacement_gradient(load("strain_ux"), load("strain_uy"), 15)
    for k, a in zip(("dudx", "dudy", "dvdx", "dvdy"), g):
        ref = load("strain_" + k)
        assert (np.isfinite(a) == np.isfinite(ref)).all()
        f = np.isfinite(ref)
        assert np.abs(a[f] - ref[f]).max() < 2e-7
    sf = O.strain_fields(*g)
    for k in ("Exx", "Eyy", "Exy", "gmax", "vol"):
        ref = load("strain_" + k)
        f = np.isfinite(ref)
        assert np.abs(sf[k][f] - ref[f]).max() < 2e-7


def test_quadratic_field_gives_exact_gradient():
    # pyt/test_strain.py:6: a quadratic field is fitted exactly away from the border
    yy, xx = np.mgrid[0:40, 0:50].astype(float)
    ux = 0.01 * xx + 0.002 * yy + 1e-4 * xx * yy + 3e-4 * xx ** 2
    uy = -0.003 * xx + 1e-4 * yy ** 2
    dudx, dudy, dvdx, dvdy = O.displacement_gradient(ux, uy, 15)
    c = np.s_[8:-8, 8:-8]
    assert np.abs(dudx[c] - (0.01 + 1e-4 * yy + 6e-4 * xx)[c]).max() < 1e-10
    assert np.abs(dvdy[c] - (2e-4 * yy)[c]).max() < 1e-10


# ---------------------------------------------------------------- pair and sequences (parity_track.rs)
def test_pair_fixture():
    a, b, roi = load("pair_a"), load("pair_b"), load("pair_roi") > 0.5
    r = O.track_pair(a, b, Config(), mask=roi)
    ux_py = load("pair_ux")
    assert (np.isfinite(r["ux"]) == np.isfinite(ux_py)).mean() >= 0.999
    assert rms(r["ux"], ux_py) < 1e-5 and rms(r["uy"], load("pair_uy")) < 1e-5
    assert rms(r["Exx"], load("pair_Exx")) < 1e-6 and rms(r["gmax"], load("pair_gmax")) < 1e-6
    assert rms(r["conf"], load("pair_conf"), np.ones(a.shape, bool)) < 1e-5
    assert np.isnan(r["ux"][~roi]).all()


def test_sequence_chain_fixtures():
    m = meta()
    out = O.track_sequence(list(load("seq_frames")), Config(ref_update_conf=2.0))
    assert [o["timing"]["reference"] for o in out] == m["seq_reference"]
    for k, o in enumerate(out):
        assert (np.isfinite(o["ux"]) == np.isfinite(load("seq_ux")[k])).mean() >= 0.999
        assert rms(o["ux"], load("seq_ux")[k]) < 1e-5 and rms(o["uy"], load("seq_uy")[k]) < 1e-5
    out = O.track_sequence(list(load("seqc1_frames")), Config(ref_update_conf=2.0))[1:]
    for k, o in enumerate(out):
        assert (np.isfinite(o["ux"]) == np.isfinite(load("seqc1_ux")[k])).mean() >= 0.999
        assert rms(o["ux"], load("seqc1_ux")[k]) < 1e-5
        assert rms(o["conf"], load("seqc1_conf")[k], np.ones(o["conf"].shape, bool)) < 1e-5


@pytest.mark.parametrize("name,forced", [("seq", False), ("seqc1", False), ("seq", True), ("seqc1", True)])
def test_sequence_mean_conf_and_reference_moves(name, forced):
    m = meta()["seq_mean_conf_forced" if forced else "seq_mean_conf"][name]
    out = O.track_sequence(list(load(name + "_frames")), Config(ref_update_conf=2.0 if forced else 0.5))
    assert [o["timing"]["reference"] for o in out] == m["reference"]
    assert np.abs(np.array([o["timing"]["mean_conf"] for o in out]) - m["mean_conf"]).max() < 1e-6


def test_refine_shifted_window_matches_full_frame():
    """refine_raw with image 2 a displaced window of the second frame
    (rust/crates/optflow-core/src/phaseflow.rs:1441-1456): cropping frame 2 to a window that holds
    every sampled position and passing the crop's offset as the shift gives the interior field of
    the whole-frame refine (the spline of the crop differs only within about 15 px of the crop's
    border, z^15 < 3e-9). A wrong sign or axis of the shift moves every sample by twice the offset
    and fails by pixels."""
    from optflow_soil_amd import synth
    pat = synth.make_pattern(96 + 2 * 64, 112 + 2 * 64, seed=3)
    f = synth.stretch_translate(0.01, (1.3, -0.8), (56.0, 48.0))
    A = synth.as_float(synth.render(pat)).astype(np.float64)
    B = synth.as_float(synth.render(pat, f, noise=0.005, seed=1)).astype(np.float64)
    h, w = A.shape
    ux0, uy0 = np.full((h, w), 1.0), np.full((h, w), -1.0)
    cfg = Config(outer_iters=3)
    full = O.refine(A, B, ux0, uy0, cfg)
    # window of frame 2 from (x0, y0) = (3, 5) to the far edge minus 2: pixel (x, y) of A sits at
    # (x - 3, y - 5) in it
    x0, y0 = 3, 5
    win = np.ascontiguousarray(B[y0:h - 2, x0:w - 2])
    part = O.refine(A, win, ux0, uy0, cfg, shift=(-x0, -y0))
    inner = np.s_[40:-40, 40:-40]
    assert np.abs(part[0][inner] - full[0][inner]).max() < 1e-4
    assert np.abs(part[1][inner] - full[1][inner]).max() < 1e-4
    # the same window with the shift's sign flipped lands elsewhere
    bad = O.refine(A, win, ux0, uy0, cfg, shift=(x0, y0))
    assert np.abs(bad[0][inner] - full[0][inner]).max() > 0.1


def test_box_sum_is_the_zero_padded_window_sum():
    """rust/crates/optflow-core/src/track/hybrid.rs:548-568."""
    h, w = 7, 9
    a = (((np.arange(h * w) * 37) % 11) - 4.0).reshape(h, w)
    for r in (1, 2, 3, 20):
        got = O.box_sum(a, r)
        for y in range(h):
            for x in range(w):
                want = a[max(0, y - r):y + r + 1, max(0, x - r):x + r + 1].sum()
                assert abs(got[y, x] - want) < 1e-9


def test_local_affine_gate_compares_with_the_model_of_the_support():
    """A plane of displacement on the support with a hole in the middle
    (rust/crates/optflow-core/src/track/hybrid.rs:570-620): the model predicts the hole; a target
    off the plane by more than tol fails, one within passes; too few support pixels fail."""
    h = w = 40
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    px, py = 0.5 + 0.01 * xx - 0.02 * yy, -1.0 + 0.015 * yy
    hole = (xx >= 15) & (xx < 25) & (yy >= 15) & (yy < 25)
    sup = ~hole
    ux, uy = np.where(sup, px, np.nan), np.where(sup, py, np.nan)

    def gate(dx, radius, min_support):
        return O.local_affine_agreement(ux, uy, sup, px + dx, py, radius, min_support, 0.5)[20, 20]

    assert gate(0.0, 12.0, 20)
    assert gate(0.4, 12.0, 20)
    assert not gate(0.6, 12.0, 20)
    assert not gate(0.0, 2.0, 20)
    assert gate(0.0, 8.0, 20)
    assert not gate(0.0, 8.0, 100000)


@pytest.mark.parametrize("name,fb_max", [("A", 10), ("B", 2)])
def test_sequence_hybrid_fixtures(name, fb_max):
    """seq_mode "hybrid" against the reference's fixtures and timing shares
    (rust/crates/optflow-core/tests/parity_track.rs:531-637): reference, n_fb_max, mean confidence,
    the fallback / gate-rejected / re-acquired shares, the carried-frames map fb, ux (and conf for
    variant A). Variant B (fb_max 2) uses the re-acquisition without a fallback value."""
    frames = load("hyb_frames")
    cfg = Config(seq_mode="hybrid", ref_update_conf=0.8, fb_max=fb_max)
    out = O.track_sequence(list(frames), cfg)
    want = meta()["hybrid"][name]
    ux_py, fb_py = load(f"hyb{name}_ux"), load(f"hyb{name}_fb")
    assert len(out) == len(want)
    for k, (r, t) in enumerate(zip(out, want)):
        tm = r["timing"]
        assert tm["reference"] == t["reference"]
        assert tm["n_fb_max"] == t["n_fb_max"]
        assert abs(tm["mean_conf"] - t["mean_conf"]) < 1e-4
        for key in ("fallback_frac", "fb_gate_rejected", "reacquired"):
            assert abs(tm[key] - t[key]) < 2e-3, key
        assert (r["fb"] == fb_py[k]).mean() >= 0.999
        assert (np.isfinite(r["ux"]) == np.isfinite(ux_py[k])).mean() >= 0.995
        assert rms(r["ux"], ux_py[k]) < 1e-3
        if name == "A":
            assert rms(r["conf"], load("hybA_conf")[k]) < 1e-3


def test_tiled_fine_stage_matches_the_whole_frame():
    """The tiled fine stage (tile.rs:225-500; cores of `tile` px, each refined on its extension
    against the window of image 2 its start reaches) agrees with the whole-frame refine to the
    reference's own bar for tiling, 1e-3 px (coret/tile_equivalence.rs:118-136): 2 x 3 tiles of
    128 px with the 64 px margin."""
    from optflow_soil_amd import synth
    h, w = 224, 288
    pat = synth.make_pattern(h, w, seed=4)
    f = synth.stretch_translate(0.01, (1.7, -0.9), (w / 2, h / 2))
    a = synth.as_float(synth.render(pat)).astype(np.float64)
    b = synth.as_float(synth.render(pat, f, noise=0.005, seed=2)).astype(np.float64)
    whole = O.track_pair(a, b, Config())
    tiled = O.track_pair(a, b, Config(tile=128))
    assert (np.isfinite(whole["ux"]) == np.isfinite(tiled["ux"])).mean() >= 0.999
    fin = np.isfinite(whole["ux"]) & np.isfinite(tiled["ux"])
    assert np.abs(whole["ux"] - tiled["ux"])[fin].max() < 1e-3
    assert np.abs(whole["uy"] - tiled["uy"])[fin].max() < 1e-3
    # a frame no wider than the tile is refined whole: identical bits
    small = O.track_pair(a[:128, :128], b[:128, :128], Config(tile=128))
    ref = O.track_pair(a[:128, :128], b[:128, :128], Config())
    assert np.array_equal(small["raw_ux"], ref["raw_ux"])
    with pytest.raises(ValueError, match="tile must be 0"):
        Config(tile=32).validate()
