"""Parity of the HIP engine (through the C ABI) with the FP64 oracle on the same inputs, at sizes
the oracle finishes in seconds, and against the reference's golden fixtures.

Bar (the reference's Rust-vs-Python contract, rust/README.md:3-5, and the parity tests
rust/crates/optflow-core/tests/parity_*.rs): displacement RMS <= 1e-3 px, strain RMS <= 1e-5,
confidence RMS <= 1e-3, valid masks identical on >= 99.9 % of pixels."""
import numpy as np
import pytest

import oracle as O
from conftest import load, meta, rms
from optflow_soil_amd.config import Config

pytestmark = pytest.mark.gpu

DISP_TOL = 1e-3
STRAIN_TOL = 1e-5
CONF_TOL = 1e-3
VALID_AGREE = 0.999


@pytest.fixture(scope="module")
def eng():
    import optflow_soil_amd as ofx
    e = ofx.Engine(0)
    yield e
    e.close()


def f32(a):
    return np.asarray(a, dtype=np.float32)


def f64(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


STRAIN_PLANES = ("Exx", "Eyy", "Exy", "gmax", "vol")
# ux / uy of pixels valid on both sides: RMS and the 99.99th percentile of |GPU - oracle| within
# the displacement bar, and no pixel further off than DISP_MAX_TOL (a pixel whose solve is
# near-degenerate amplifies the float-vs-double difference of its sums).
DISP_P9999_TOL = 1e-3
DISP_MAX_TOL = 1e-2


def compare_pair(r, o, label, roi=None):
    """Every output plane of a pair against the oracle. Valid-mask agreement is taken over ROI
    pixels when an ROI is given (outside it both sides are NaN by construction)."""
    valid_g, valid_o = np.isfinite(r.ux), np.isfinite(o["ux"])
    region = np.ones(valid_g.shape, bool) if roi is None else np.asarray(roi, bool)
    agree = float((valid_g == valid_o)[region].mean())
    both = valid_g & valid_o
    d = {"agree": agree, "n_valid": int(both.sum()),
         "ux": rms(r.ux.astype(np.float64), o["ux"], both), "uy": rms(r.uy.astype(np.float64), o["uy"], both),
         "conf": rms(r.conf.astype(np.float64), o["conf"], np.ones(both.shape, bool))}
    if both.any():
        dev = np.maximum(np.abs(r.ux.astype(np.float64) - o["ux"])[both], np.abs(r.uy.astype(np.float64) - o["uy"])[both])
        d["disp_p9999"] = float(np.quantile(dev, 0.9999))
        d["disp_max"] = float(dev.max())
    else:
        d["disp_p9999"] = d["disp_max"] = 0.0
    # strain is NaN wherever fewer than half of the window's pixels are valid; a frame with
    # too few valid pixels has no strain to compare, and then both sides must say so
    fin = np.isfinite(r.strain["Exx"]) & np.isfinite(o["Exx"])
    d["strain_agree"] = float((np.isfinite(r.strain["Exx"]) == np.isfinite(o["Exx"]))[region].mean())
    for p in STRAIN_PLANES:
        d[p] = rms(r.strain[p].astype(np.float64), o[p], fin) if fin.any() else 0.0
    print(label, " ".join(f"{k}={v:.3g}" for k, v in d.items()), flush=True)
    assert agree >= VALID_AGREE and d["strain_agree"] >= VALID_AGREE
    assert d["ux"] <= DISP_TOL and d["uy"] <= DISP_TOL
    assert d["disp_p9999"] <= DISP_P9999_TOL and d["disp_max"] <= DISP_MAX_TOL
    for p in STRAIN_PLANES:
        assert d[p] <= STRAIN_TOL, p
    assert d["conf"] <= CONF_TOL
    return d


@pytest.mark.parametrize("order", [1, 2])
def test_refine_fixture(eng, order):
    a, b = load("refine_a"), load("refine_b")
    ix, iy = load("refine_init_ux"), load("refine_init_uy")
    cfg = Config(shape_order=order)
    ux, uy, en, res = eng.refine_raw(f32(a), f32(b), f32(ix), f32(iy), cfg)
    oux, ouy, oconf, ores, oen = O.refine(f64(a), f64(b), f64(ix), f64(iy), cfg)
    e = {"ux": rms(ux, oux), "uy": rms(uy, ouy), "resid": rms(res, ores), "energy_rel": rms(en, oen) / np.abs(oen).mean(),
         "fixture_ux": rms(ux, load(f"refine_o{order}_ux"))}
    print("refine", " ".join(f"{k}={v:.3g}" for k, v in e.items()), flush=True)
    assert e["ux"] <= DISP_TOL and e["uy"] <= DISP_TOL and e["fixture_ux"] <= DISP_TOL
    assert e["resid"] <= 1e-3 and e["energy_rel"] <= 1e-4


def test_refine_shifted_window(eng):
    """FineEngine::refine_raw with image 2 a displaced window of the second frame (the tiled
    driver's call, rust/crates/optflow-core/src/tile.rs:114-146, phaseflow.rs:1441-1456): GPU vs
    oracle on the same window and shift. The comparison leaves out the border band of 4 sigma_w
    (16 px) that the border rule never reports (phaseflow.rs:1623-1629): there the top rows and
    left columns sample outside the window, the clamped spline has no texture along one axis, and
    the ill-posed solves amplify float-vs-double differences."""
    from optflow_soil_amd import synth
    pat = synth.make_pattern(96 + 2 * 64, 112 + 2 * 64, seed=3)
    f = synth.stretch_translate(0.01, (1.3, -0.8), (56.0, 48.0))
    A = synth.as_float(synth.render(pat))
    B = synth.as_float(synth.render(pat, f, noise=0.005, seed=1))
    h, w = A.shape
    win = np.ascontiguousarray(B[5:h - 2, 3:w - 2])
    ix, iy = np.full((h, w), 1.0, np.float32), np.full((h, w), -1.0, np.float32)
    cfg = Config()
    ux, uy, en, res = eng.refine_raw(A, win, ix, iy, cfg, shift=(-3, -5))
    oux, ouy, _, ores, oen = O.refine(f64(A), f64(win), f64(ix), f64(iy), cfg, shift=(-3, -5))
    b = int(4 * cfg.sigma_w)
    inner = np.zeros((h, w), bool)
    inner[b:-b, b:-b] = True
    e = {"ux": rms(ux, oux, inner), "uy": rms(uy, ouy, inner), "resid": rms(res, ores, inner),
         "energy_rel": rms(en, oen, inner) / np.abs(oen[inner]).mean(),
         "ux_full_frame": rms(ux, oux), "uy_full_frame": rms(uy, ouy)}
    print("refine shifted", " ".join(f"{k}={v:.3g}" for k, v in e.items()), flush=True)
    assert e["ux"] <= DISP_TOL and e["uy"] <= DISP_TOL
    assert e["resid"] <= 1e-3 and e["energy_rel"] <= 1e-4


@pytest.mark.parametrize("case", ["pair", "hyb23"])
def test_coarse_flow(eng, case):
    """The coarse stage and dense start (poc.py:216-239, track.py:131-137) against the oracle: the
    POC grid is integer peak + parabolic sub-pixel fit, so a float-vs-double difference stays far
    below a pixel; a different peak or clean decision would show as a jump of a pixel or more."""
    if case == "pair":
        a, b = load("pair_a"), load("pair_b")
    else:
        fr = load("hyb_frames")
        a, b = fr[2], fr[3]
    gx, gy = eng.coarse_flow(f32(a), f32(b), Config())
    ox, oy = O.coarse_dense(f64(a), f64(b), Config())
    e = {"ux": rms(gx, ox), "uy": rms(gy, oy), "max": float(max(np.abs(gx - ox).max(), np.abs(gy - oy).max()))}
    print(f"coarse {case} " + " ".join(f"{k}={v:.3g}" for k, v in e.items()), flush=True)
    assert e["ux"] <= DISP_TOL and e["uy"] <= DISP_TOL and e["max"] <= 1e-2


def test_pair_where_a_block_appears(eng):
    """Frames 1 -> 2 of the hybrid fixture: a foreign block appears, so a patch of the frame has
    confidence near 0 and its pixels take the pre-warp's first-order fill from the nearest trusted
    pixel, chosen with scipy's tie rule (smallest squared distance, then smallest column, then
    smallest row; phaseflow.py:344-355, edt.rs:175-183). The raw (unmasked) field there feeds
    later hybrid frames, so it is compared too, away from the 28 px border band."""
    fr = load("hyb_frames")
    r = eng.track_pair(f32(fr[1]), f32(fr[2]), Config())
    o = O.track_pair(f64(fr[1]), f64(fr[2]), Config())
    inner = np.zeros(r.ux.shape, bool)
    inner[28:-28, 28:-28] = True
    e = {"raw_ux": rms(r.raw_ux.astype(np.float64), o["raw_ux"], inner),
         "raw_uy": rms(r.raw_uy.astype(np.float64), o["raw_uy"], inner),
         "raw_max": float(max(np.abs(r.raw_ux - o["raw_ux"])[inner].max(), np.abs(r.raw_uy - o["raw_uy"])[inner].max()))}
    print("block appears " + " ".join(f"{k}={v:.3g}" for k, v in e.items()), flush=True)
    compare_pair(r, o, "block appears")
    assert e["raw_ux"] <= 1e-4 and e["raw_uy"] <= 1e-4 and e["raw_max"] <= 1e-2


@pytest.mark.parametrize("h,w", [(131, 163), (99, 70)])
def test_pair_with_chirp_z_lengths(eng, h, w):
    """Frame sizes whose padded lengths (side + 64) have a prime factor of 67 or more go through the
    chirp-z form of the line transform: 163 + 64 = 227 and 99 + 64 = 163 are prime, 70 + 64 =
    2 x 67. GPU against the oracle (whose transforms are direct double-precision DFTs)."""
    from optflow_soil_amd import synth
    pat = synth.make_pattern(h, w, seed=h + w)
    f = synth.affine_field(0.004, -0.003, 0.002, 0.3, (1.7, -0.9), ((w - 1) / 2.0, (h - 1) / 2.0))
    a = synth.as_float(synth.render(pat, None, noise=0.005, seed=1))
    b = synth.as_float(synth.render(pat, f, noise=0.005, seed=2))
    print(f"chirp-z pair {w}x{h}", flush=True)
    r = eng.track_pair(a, b, Config())
    o = O.track_pair(f64(a), f64(b), Config())
    compare_pair(r, o, f"chirp-z pair {w}x{h}")


def test_pair_fixture(eng):
    a, b, roi = load("pair_a"), load("pair_b"), load("pair_roi") > 0.5
    r = eng.track_pair(f32(a), f32(b), Config(), mask=roi)
    o = O.track_pair(f64(a), f64(b), Config(), mask=roi)
    compare_pair(r, o, "pair", roi=roi)
    # and against the reference's own fixture
    assert (np.isfinite(r.ux) == np.isfinite(load("pair_ux"))).mean() >= VALID_AGREE
    assert rms(r.ux.astype(np.float64), load("pair_ux")) <= DISP_TOL
    assert np.isnan(r.ux[~roi]).all()


def test_sequence_chain_fixture(eng):
    frames = [f32(f) for f in load("seq_frames")]
    cfg = Config(ref_update_conf=2.0)
    out = eng.track_sequence(frames, cfg)
    assert [res.timing["reference"] for _, res in out] == meta()["seq_reference"]
    for k, (_, res) in enumerate(out):
        ref = load("seq_ux")[k]
        assert (np.isfinite(res.ux) == np.isfinite(ref)).mean() >= VALID_AGREE
        assert rms(res.ux.astype(np.float64), ref) <= DISP_TOL
        assert rms(res.uy.astype(np.float64), load("seq_uy")[k]) <= DISP_TOL


def test_c1_chain_workload(eng):
    """256 x 256 speckle, 5 frames, seq_mode chain, uint8 input read as v / 255."""
    from optflow_soil_amd import synth
    frames, truth = synth.workload_c1()
    ff = [synth.as_float(f) for f in frames]
    cfg = Config()
    out = eng.track_sequence(ff, cfg)
    ref = O.track_sequence([f.astype(np.float64) for f in ff], cfg)
    for (k, r), o in zip(out, ref):
        assert r.timing["reference"] == o["timing"]["reference"]
        d = compare_pair(r, o, f"c1 frame {k}")
        gx, gy = truth(k)
        fin = np.isfinite(r.ux)
        epe = float(np.sqrt(np.mean((r.ux[fin] - gx[fin]) ** 2 + (r.uy[fin] - gy[fin]) ** 2)))
        print(f"c1 frame {k}: EPE vs truth {epe:.4f} px over {int(fin.sum())} px", flush=True)
        assert epe < 0.2


@pytest.mark.parametrize("name,fb_max", [
    ("A", 10),
    ("B", 2),
])
def test_sequence_hybrid_fixtures(eng, name, fb_max):
    """seq_mode "hybrid" on the GPU against the reference's fixtures, with the bar of the
    reference's own test (rust/crates/optflow-core/tests/parity_track.rs:531-637): reference and
    n_fb_max equal, mean confidence within 1e-4, the fallback / rejected / re-acquired shares within
    2e-3, fb equal on >= 99.9 % of pixels, NaN agreement >= 99.5 %, conf RMS < 1e-3 for variant A.
    ux RMS < 1e-3 against the fixture and the oracle is taken over the pixels whose carried-frames
    count fb agrees: a pixel whose gate decision flips between float and double legs carries a
    different (direct or adjacent-leg) value, and the >= 99.9 % fb bar bounds those."""
    frames = load("hyb_frames")
    cfg = Config(seq_mode="hybrid", ref_update_conf=0.8, fb_max=fb_max)
    print(f"hybrid {name}: GPU sequence of {len(frames)} frames {frames.shape[1:]}", flush=True)
    out = eng.track_sequence([f32(f) for f in frames], cfg)
    print(f"hybrid {name}: oracle sequence", flush=True)
    orc = O.track_sequence([f64(f) for f in frames], cfg)
    want = meta()["hybrid"][name]
    ux_py, fb_py = load(f"hyb{name}_ux"), load(f"hyb{name}_fb")
    assert len(out) == len(want)
    for (k, r), t, o in zip(out, want, orc):
        tm = r.timing
        same = r.fb == fb_py[k - 1]
        fb_same = float(same.mean())
        agree = float((np.isfinite(r.ux) == np.isfinite(ux_py[k - 1])).mean())
        both = same & np.isfinite(r.ux) & np.isfinite(ux_py[k - 1])
        e_fix = rms(r.ux.astype(np.float64), ux_py[k - 1], both)
        both_o = (r.fb == o["fb"]) & np.isfinite(r.ux) & np.isfinite(o["ux"])
        e_orc = rms(r.ux.astype(np.float64), o["ux"], both_o)
        e_all = rms(r.ux.astype(np.float64), ux_py[k - 1])
        print(f"hybrid {name} frame {k}: ref {tm['reference']}/{t['reference']} n_fb_max {tm['n_fb_max']}/"
              f"{t['n_fb_max']} fallback {tm['fallback_frac']:.4f}/{t['fallback_frac']:.4f} rejected "
              f"{tm['fb_gate_rejected']:.4f}/{t['fb_gate_rejected']:.4f} reacquired {tm['reacquired']:.4f}/"
              f"{t['reacquired']:.4f} fb_same {fb_same:.5f} agree {agree:.5f} ux_fixture {e_fix:.2e} "
              f"ux_oracle {e_orc:.2e} ux_fixture_all_pixels {e_all:.2e}", flush=True)
        assert tm["reference"] == t["reference"]
        assert tm["n_fb_max"] == t["n_fb_max"]
        assert abs(tm["mean_conf"] - t["mean_conf"]) < 1e-4
        for key in ("fallback_frac", "fb_gate_rejected", "reacquired"):
            assert abs(tm[key] - t[key]) < 2e-3, key
        assert fb_same >= 0.999
        assert agree >= 0.995
        assert e_fix < DISP_TOL and e_orc < DISP_TOL
        if name == "A":
            assert rms(r.conf.astype(np.float64), load("hybA_conf")[k - 1]) < CONF_TOL


@pytest.mark.parametrize("mode", ["chain", "hybrid"])
def test_streaming_sequence_matches_the_frame_stack(eng, mode):
    """ofx_seq_begin / ofx_seq_step (one frame per call, the Sequence iterator of
    track.rs:1112-1225) run the same kernels in the same order as ofx_track_sequence over the
    whole stack, so every output plane and timing share is bit-identical."""
    frames = [f32(f) for f in (load("seq_frames") if mode == "chain" else load("hyb_frames"))]
    cfg = Config(ref_update_conf=2.0) if mode == "chain" else Config(seq_mode="hybrid", ref_update_conf=0.8)
    whole = eng.track_sequence(frames, cfg)
    print(f"streaming {mode}: {len(frames)} frames", flush=True)
    streamed = list(eng.iter_sequence(iter(frames), cfg))
    assert [k for k, _ in streamed] == [k for k, _ in whole]
    for (k, a), (_, b) in zip(whole, streamed):
        for name in ("ux", "uy", "conf", "raw_ux", "raw_uy"):
            np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=f"frame {k} {name}")
        for name in a.strain:
            np.testing.assert_array_equal(a.strain[name], b.strain[name], err_msg=f"frame {k} {name}")
        for key in ("reference", "mean_conf", "n_fb_max", "fallback_frac", "fb_gate_rejected", "reacquired"):
            assert a.timing.get(key) == b.timing.get(key), (k, key)
        if mode == "hybrid":
            np.testing.assert_array_equal(a.fb, b.fb)
    # the sequence released its lane: a pair call on the same engine runs
    eng.track_pair(frames[0], frames[1], Config())


def test_batched_pairs_match_single_pairs(eng):
    """ofx_track_pairs spreads uneven pairs over three lanes (threads and streams); each result is
    bit-identical to ofx_track_pair on the same pair."""
    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth
    sizes = [(96, 128), (128, 128), (160, 96), (112, 112), (128, 96)]
    pairs = []
    for k, (h, w) in enumerate(sizes):
        pat = synth.make_pattern(h, w, seed=20 + k)
        f = synth.stretch_translate(0.004 * (k + 1), (0.7 * k - 1.0, 0.4), (w / 2, h / 2))
        a = synth.as_float(synth.render(pat))
        b = synth.as_float(synth.render(pat, f, noise=0.005, seed=k))
        mask = np.zeros((h, w), bool)
        mask[8:-8, 10:-6] = True
        pairs.append((a, b, mask if k % 2 else None))
    e3 = ofx.Engine(0, lanes=3)
    try:
        print("batched: 5 pairs on 3 lanes", flush=True)
        got = e3.track_pairs(pairs, Config())
    finally:
        e3.close()
    for k, (p, r) in enumerate(zip(pairs, got)):
        one = eng.track_pair(p[0], p[1], Config(), mask=p[2])
        for name in ("ux", "uy", "conf", "raw_ux", "raw_uy"):
            np.testing.assert_array_equal(getattr(one, name), getattr(r, name), err_msg=f"pair {k} {name}")
        for name in one.strain:
            np.testing.assert_array_equal(one.strain[name], r.strain[name], err_msg=f"pair {k} {name}")


def test_tiled_pair(eng):
    """Config.tile: the reference's tiled fine stage (tile.rs:225-500) on the GPU, 2 x 3 tiles of
    128 px in three groups that run side by side: against the oracle's tiled pair at the pair bar,
    against the GPU's whole-frame pair to the reference's tiling bar, 1e-3 px, and bit for bit
    against a second run (each tile is refined alone, whatever runs beside it)."""
    from optflow_soil_amd import synth
    h, w = 224, 288
    pat = synth.make_pattern(h, w, seed=4)
    f = synth.stretch_translate(0.01, (1.7, -0.9), (w / 2, h / 2))
    a = synth.as_float(synth.render(pat))
    b = synth.as_float(synth.render(pat, f, noise=0.005, seed=2))
    print("tiled pair 224x288, tile 128", flush=True)
    r = eng.track_pair(a, b, Config(tile=128))
    o = O.track_pair(f64(a), f64(b), Config(tile=128))
    compare_pair(r, o, "tiled pair")
    again = eng.track_pair(a, b, Config(tile=128))
    for name in ("ux", "uy", "conf", "raw_ux", "raw_uy"):
        np.testing.assert_array_equal(getattr(r, name), getattr(again, name), err_msg=name)
    whole = eng.track_pair(a, b, Config())
    fin = np.isfinite(whole.ux) & np.isfinite(r.ux)
    print(f"tiled vs whole frame: max ux {np.abs(whole.ux - r.ux)[fin].max():.2e}", flush=True)
    assert np.abs(whole.ux - r.ux)[fin].max() < 1e-3 and np.abs(whole.uy - r.uy)[fin].max() < 1e-3


def test_tiled_pair_wider_than_one_line_transform(eng):
    """A frame with a side over 4032 px runs when tiled: the line limit applies to the largest
    window the fine stage refines, a tile's extension (the reference's check_window(min(w, side),
    min(h, side)), track.rs:754-762), not to the whole frame. 4160 x 192 in tiles of 1024 against
    the oracle's tiled pair; untiled, the same frame is refused as too large before any work."""
    from optflow_soil_amd import synth
    h, w = 192, 4160
    pat = synth.make_pattern(h, w, seed=11)
    f = synth.stretch_translate(0.004, (1.2, -0.6), (w / 2, h / 2))
    a = synth.as_float(synth.render(pat))
    b = synth.as_float(synth.render(pat, f, noise=0.005, seed=5))
    assert eng.check_window(w, h, Config(tile=1024)) is None
    with pytest.raises(ValueError, match="exceeds 4032 px"):
        eng.check_window(w, h, Config())
    with pytest.raises(ValueError, match="exceeds 4032 px"):
        eng.track_pair(a, b, Config())
    print(f"tiled pair {w}x{h}, tile 1024", flush=True)
    r = eng.track_pair(a, b, Config(tile=1024))
    o = O.track_pair(f64(a), f64(b), Config(tile=1024))
    compare_pair(r, o, f"tiled pair {w}x{h}")


@pytest.mark.parametrize("order", [1, 2])
def test_refine_stripe_runs_the_degenerate_paths(eng, order):
    """The reference's stripe fixture (texture along one axis only): the aperture test marks the
    windows degenerate, which keep their input displacement and get resid = pi, and the energy
    reset runs (phaseflow.rs:598-859; the reference's GPU test degenerate_stripes_run_the_reset_paths)."""
    init = meta()["refine_stripe"]["init"]
    a, b = load("refine_stripe_a"), load("refine_stripe_b")
    ix, iy = np.full(a.shape, init[0], np.float32), np.full(a.shape, init[1], np.float32)
    cfg = Config(shape_order=order)
    ux, uy, en, res = eng.refine_raw(f32(a), f32(b), ix, iy, cfg)
    oux, ouy, _, ores, oen = O.refine(f64(a), f64(b), f64(ix), f64(iy), cfg)
    degenerate = ores == np.pi
    e = {"ux": rms(ux, load(f"refine_stripe_o{order}_ux")), "uy": rms(uy, load(f"refine_stripe_o{order}_uy")),
         "degenerate_agree": float(((res == np.pi) == degenerate).mean()), "n_degenerate": int(degenerate.sum())}
    print(f"stripe o{order} " + " ".join(f"{k}={v:.3g}" for k, v in e.items()), flush=True)
    assert e["ux"] <= DISP_TOL and e["uy"] <= DISP_TOL and e["degenerate_agree"] >= VALID_AGREE
    assert np.array_equal(ux[degenerate & (res == np.pi)], ix[degenerate & (res == np.pi)].astype(np.float64))


def test_refine_border_fixture(eng):
    """The reference's border fixture (guard off, the field reaching the image border): ux, uy within
    1e-3 px of the fixture everywhere (parity_refine.rs:126-177)."""
    f = meta()["refine_border"]
    a, b = load("refine_border_a"), load("refine_border_b")
    n = a.shape[0]
    yy, xx = np.mgrid[0:n, 0:n].astype(float)
    cx, cy = f["center"]
    gx = f["exx"] * (xx - cx) + f["t"][0] + f["init_offset"][0]
    gy = f["eyy"] * (yy - cy) + f["t"][1] + f["init_offset"][1]
    ux, uy, _, _ = eng.refine_raw(f32(a), f32(b), f32(gx), f32(gy), Config(edge_guard=0.0))
    e = {"ux_max": float(np.abs(ux - load("refine_border_ux")).max()),
         "uy_max": float(np.abs(uy - load("refine_border_uy")).max())}
    print("border " + " ".join(f"{k}={v:.3g}" for k, v in e.items()), flush=True)
    assert e["ux_max"] < 1e-3 and e["uy_max"] < 1e-3


PAIR_VARIANTS = {
    "freeze3": dict(prewarp_freeze_at=3),
    "order2": dict(shape_order=2),
    "bank6x2": dict(n_orient=6, scales=(0.8, 1.3)),
    "windows3": dict(sigma_w=3.0, sigma_s=2.5),
    "guard_off": dict(edge_guard=0.0),
    "guard20": dict(edge_guard=20.0, edge_guard_min_area=100.0),
    "strain21": dict(strain_win=21),
    "iters4x3": dict(outer_iters=4, inner_iters=3),
    "strict": dict(conf_min=0.5, phase_tol=0.3),
    "two_levels": dict(max_disp=48.0),
}


@pytest.mark.parametrize("variant", sorted(PAIR_VARIANTS))
def test_pair_settings(eng, variant):
    """Whole pairs with non-default settings, GPU against the oracle at the pair bar: the freeze of
    the low mask (prewarp.rs LowFreeze), the quadratic shape function, another filter bank, other
    window widths (the generic moment kernels), the guard off or wider, a wider strain window, other
    iteration counts, stricter confidence, and max_disp 48 (two POC pyramid levels and non-square
    level windows on the generic POC kernel)."""
    from optflow_soil_amd import synth
    h, w = (256, 256) if variant == "two_levels" else (192, 224)
    pat = synth.make_pattern(h, w, seed=7)
    f = synth.stretch_translate(0.008, (2.3, -1.1), (w / 2, h / 2))
    a = synth.as_float(synth.render(pat))
    b = synth.as_float(synth.render(pat, f, noise=0.005, seed=3))
    roi = synth.convex_polygon_mask(h, w, seed=7, coverage=0.8)
    cfg = Config(**PAIR_VARIANTS[variant])
    print(f"settings {variant}: {h}x{w}", flush=True)
    r = eng.track_pair(a, b, cfg, mask=roi)
    o = O.track_pair(f64(a), f64(b), cfg, mask=roi)
    compare_pair(r, o, f"settings {variant}", roi=roi)


def test_engine_errors_are_the_reference_messages(eng):
    """Argument errors come back as the reference's messages, before any device work."""
    a = np.zeros((64, 80), np.float32)
    with pytest.raises(ValueError, match="ROI mask is empty"):
        eng.track_pair(a, a, Config(), mask=np.zeros((64, 80), bool))
    with pytest.raises(ValueError, match="image smaller than poc_win"):
        eng.track_pair(a[:40], a[:40], Config())
    with pytest.raises(ValueError, match="ROI mask shape"):
        eng.track_pair(a, a, Config(), mask=np.ones((10, 10), bool))
    with pytest.raises(ValueError, match="poc_win must be a power of two"):
        eng.track_pair(a, a, Config(poc_win=48))
    with pytest.raises(ValueError, match="no frames"):
        list(eng.iter_sequence(iter([]), Config()))


# the oracle's side of the 4000 x 3000 pair takes about a minute on 16 host cores: room for fewer
@pytest.mark.timeout(400)
@pytest.mark.parametrize("workload", ["bowl_and_band_2048", "large_motion_4000x3000"])
def test_benchmark_pairs_at_full_size(eng, workload):
    """The benchmark's pairs at their full sizes, GPU against the oracle over the whole frame (all
    host cores), to the same bar, and the end-point error of both against the synthetic truth:
    the 2048 x 2048 settlement bowl and shear band with its 8-gon ROI, and the 4000 x 3000 pair
    with six Gaussian bumps up to 40 px (max_disp 48, two POC levels, chirp-z transforms)."""
    from optflow_soil_amd import synth
    if workload == "bowl_and_band_2048":
        a, b, roi, (tx, ty) = synth.workload_c2(2048)
        cfg = Config()
    else:
        a, b, (tx, ty) = synth.workload_c5()
        roi, cfg = None, Config(max_disp=48.0)
    A, Bm = synth.as_float(a), synth.as_float(b)
    print(f"{workload}: rendered {A.shape}", flush=True)
    r = eng.track_pair(A, Bm, cfg, mask=roi)
    print(f"{workload}: GPU done; oracle", flush=True)
    o = O.track_pair(f64(A), f64(Bm), cfg, mask=roi)
    compare_pair(r, o, workload, roi=roi)
    for name, ux, uy in (("gpu", r.ux.astype(np.float64), r.uy.astype(np.float64)), ("oracle", o["ux"], o["uy"])):
        fin = np.isfinite(ux)
        epe = float(np.sqrt(np.mean((ux[fin] - tx[fin]) ** 2 + (uy[fin] - ty[fin]) ** 2)))
        print(f"{workload}: {name} EPE vs truth {epe:.5f} px over {fin.mean():.4f} of the frame", flush=True)
        assert epe < 0.06


def test_hybrid_sequence_with_temporal_averaging(eng):
    """The settings of the 200-frame hybrid sequence workload on a reduced case: 12 frames of
    384 x 384 whose field grows linearly to a bowl-and-band field at the last frame, noise 0.01 and
    a brightness drift of -3 % .. +3 %; temporal averaging over +-3 frames (edge windows aligned),
    seq_mode hybrid, prewarp_freeze_at 3, ref_update_conf 0.3. GPU (averaging and tracking)
    against the oracle (averaging and tracking) end to end, to the bar of the hybrid fixtures; and
    the drift of the last frame's cumulative field against the truth."""
    from optflow_soil_amd import synth
    from optflow_soil_amd.frames import averaged, gpu_average
    n, nf = 384, 12
    pat = synth.make_pattern(n, n, seed=21)
    full = synth.bowl_and_band(n, n, depth=6.0, radius=300.0 * n / 2048.0, slip=3.0, band_width=40.0 * n / 2048.0)

    def field_at(k):
        s = k / (nf - 1)
        return lambda X, Y: tuple(s * c for c in full(X, Y))
    frames = []
    for k in range(nf):
        v = synth.render(pat, field_at(k) if k else None, noise=0.0, quantise=False)
        v = v * (0.97 + 0.06 * k / (nf - 1)) + np.random.default_rng(100 + k).normal(0.0, 0.01, v.shape)
        frames.append(synth.as_float(np.clip(np.rint(v * 255.0), 0, 255).astype(np.uint8)))
    cfg = Config(seq_mode="hybrid", prewarp_freeze_at=3, ref_update_conf=0.3)

    def oracle_average(fr, centre, model, c, mask, keep):
        return O.average_window(fr, centre, model, c, mask, keep)[0]
    print("c3-like: averaging", flush=True)
    avg_g = list(averaged(iter(frames), 0, None, 1, 3, cfg=cfg, average=gpu_average(eng)))
    avg_o = list(averaged(iter(frames), 0, None, 1, 3, cfg=cfg, average=oracle_average))
    print("c3-like: GPU sequence", flush=True)
    out = eng.track_sequence([f32(f) for f in avg_g], cfg)
    print("c3-like: oracle sequence", flush=True)
    orc = O.track_sequence([f64(f) for f in avg_o], cfg)
    assert len(out) == len(orc) == nf - 1
    for (k, r), o in zip(out, orc):
        tm, to = r.timing, o["timing"]
        same = r.fb == o["fb"]
        fb_same = float(same.mean())
        agree = float((np.isfinite(r.ux) == np.isfinite(o["ux"])).mean())
        both = same & np.isfinite(r.ux) & np.isfinite(o["ux"])
        e = rms(r.ux.astype(np.float64), o["ux"], both)
        ey = rms(r.uy.astype(np.float64), o["uy"], both)
        print(f"c3-like frame {k}: ref {tm['reference']}/{to['reference']} fb_same {fb_same:.5f} agree {agree:.5f} "
              f"ux {e:.2e} uy {ey:.2e} fallback {tm['fallback_frac']:.4f}/{to['fallback_frac']:.4f}", flush=True)
        assert tm["reference"] == to["reference"]
        assert tm["n_fb_max"] == to["n_fb_max"]
        for key in ("fallback_frac", "fb_gate_rejected", "reacquired"):
            assert abs(tm[key] - to[key]) < 2e-3, key
        assert fb_same >= 0.999 and agree >= 0.995
        assert e < DISP_TOL and ey < DISP_TOL
    k, r = out[-1]
    gx, gy = field_at(k)(*np.mgrid[0:n, 0:n][::-1].astype(np.float64))
    fin = np.isfinite(r.ux)
    drift = float(np.sqrt(np.mean((r.ux[fin] - gx[fin]) ** 2 + (r.uy[fin] - gy[fin]) ** 2)))
    print(f"c3-like: cumulative drift vs truth at frame {k}: {drift:.4f} px over {fin.mean():.3f} of the frame",
          flush=True)
    # the oracle's own drift here is 0.145 px: the last frame's mean is a one-sided window, which
    # the dense alignment brings back only to within its own error
    assert drift < 0.3


@pytest.mark.timeout(170)
def test_batched_uneven_pairs_at_workload_sizes():
    """The batched workload's five size classes (512^2, 768 x 1024, 1024^2, 1536^2, 1024 x 2048), one
    pair each with its affine field and polygon ROI, through ofx_track_pairs on 4 lanes, against
    the oracle pair by pair at the pair bar."""
    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth
    rng = np.random.default_rng(5)
    pairs = []
    for k, (h, w) in enumerate(synth.C4_SIZES):
        pat = synth.make_pattern(h, w, seed=300 + k)
        f = synth.affine_field(*rng.uniform(-0.02, 0.02, 3), rng.uniform(-1.0, 1.0), rng.uniform(-8.0, 8.0, 2),
                               ((w - 1) / 2.0, (h - 1) / 2.0))
        a = synth.as_float(synth.render(pat, None, noise=0.005, seed=2 * k))
        b = synth.as_float(synth.render(pat, f, noise=0.005, seed=2 * k + 1))
        roi = synth.convex_polygon_mask(h, w, seed=k, coverage=float(rng.uniform(0.4, 0.9)), search_step=8)
        pairs.append((a, b, roi))
    e4 = ofx.Engine(0, lanes=4)
    try:
        print("batched: the five size classes on 4 lanes", flush=True)
        got = e4.track_pairs(pairs, Config())
    finally:
        e4.close()
    for (a, b, roi), r in zip(pairs, got):
        print(f"batched pair {a.shape}: oracle", flush=True)
        o = O.track_pair(f64(a), f64(b), Config(), mask=roi)
        compare_pair(r, o, f"batched {a.shape[1]}x{a.shape[0]}", roi=roi)


def test_batched_tiled_pairs_match_single_tiled_pairs(eng):
    """Tiles inside batched pairs: three pairs of 260 x 300 in tiles of 128 on two lanes, so that
    two lanes each run three tile groups on streams of their own at once; every result is
    bit-identical to the same pair tracked alone."""
    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth
    pairs = []
    for k in range(3):
        h, w = 260, 300
        pat = synth.make_pattern(h, w, seed=40 + k)
        f = synth.affine_field(0.003 * (k + 1), -0.002, 0.001, 0.2 * k, (1.2 - k, 0.6), ((w - 1) / 2.0, (h - 1) / 2.0))
        pairs.append((synth.as_float(synth.render(pat)), synth.as_float(synth.render(pat, f, noise=0.005, seed=k))))
    cfg = Config(tile=128)
    e2 = ofx.Engine(0, lanes=2)
    try:
        print("batched tiled: 3 pairs on 2 lanes", flush=True)
        got = e2.track_pairs(pairs, cfg)
    finally:
        e2.close()
    for k, (p, r) in enumerate(zip(pairs, got)):
        one = eng.track_pair(p[0], p[1], cfg)
        for name in ("ux", "uy", "conf", "raw_ux", "raw_uy"):
            np.testing.assert_array_equal(getattr(one, name), getattr(r, name), err_msg=f"pair {k} {name}")


@pytest.mark.parametrize("flat", [False, True])
def test_bounded_energy_reset_matches_the_exact_median(eng, flat, monkeypatch):
    """The first inner iteration of each outer one resets low-energy pixels from bounds on the
    median (two digit passes) and takes the exact median only for pixels those bounds leave open;
    OFX_EXACT_ENERGY_RESET forces the exact median. Both give bit-identical results, on a textured
    pair and on one with flat blocks (energy near zero: reset pixels next to kept ones)."""
    from optflow_soil_amd import synth
    h, w = 300, 340
    pat = synth.make_pattern(h, w, seed=77)
    f = synth.affine_field(0.003, -0.002, 0.001, 0.2, (1.3, -0.7), ((w - 1) / 2.0, (h - 1) / 2.0))
    a = synth.as_float(synth.render(pat))
    b = synth.as_float(synth.render(pat, f, noise=0.005, seed=3))
    if flat:
        for arr in (a, b):
            arr[40:120, 60:150] = 0.5
            arr[200:260, 220:330] = 0.25
    cfg = Config()
    monkeypatch.delenv("OFX_EXACT_ENERGY_RESET", raising=False)
    print(f"bounded energy reset: flat={flat}", flush=True)
    got = eng.track_pair(a, b, cfg)
    monkeypatch.setenv("OFX_EXACT_ENERGY_RESET", "1")
    ref = eng.track_pair(a, b, cfg)
    for name in ("ux", "uy", "conf", "raw_ux", "raw_uy"):
        np.testing.assert_array_equal(getattr(got, name), getattr(ref, name), err_msg=name)
