"""ctypes binding of the FP64 CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, always as the checker or the timed CPU baseline and never as part of the
product path. The oracle restates the reference's algorithm in double precision; every C
function cites the reference file and line it follows.

Parity pinning: the `not gpu` tests check every function here against the reference's own golden
fixtures (tests/golden, from rust/crates/optflow-core/tests/fixtures), against scipy/numpy where a
function reduces to a library routine (FFT, EDT, labels, correlate1d), and against closed forms
(plane waves, translations, quadratic fields). `track_sequence` in hybrid mode is pinned by the
reference's two hybrid fixtures (fb maps and timing shares identical, ux within 1e-7), its local
affine gate by a brute-force window sum and by planted outliers. No function here is unpinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def build():
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _Cfg(ctypes.Structure):
    _fields_ = [("grid_step", ctypes.c_int), ("poc_win", ctypes.c_int), ("max_disp", ctypes.c_double),
                ("poc_peak_min", ctypes.c_double), ("median_thresh", ctypes.c_double), ("lam", ctypes.c_double),
                ("n_scales", ctypes.c_int), ("scales", ctypes.c_double * 16), ("n_orient", ctypes.c_int),
                ("bandwidth", ctypes.c_double), ("dc_cut", ctypes.c_double), ("sigma_w", ctypes.c_double),
                ("shape_order", ctypes.c_int), ("sigma_s", ctypes.c_double), ("outer_iters", ctypes.c_int),
                ("inner_iters", ctypes.c_int), ("conf_min", ctypes.c_double), ("edge_guard", ctypes.c_double),
                ("edge_guard_min_area", ctypes.c_double), ("phase_tol", ctypes.c_double),
                ("strain_win", ctypes.c_int), ("ref_update_conf", ctypes.c_double), ("seq_hybrid", ctypes.c_int),
                ("fb_max", ctypes.c_int), ("fb_radius", ctypes.c_double), ("fb_min_support", ctypes.c_int),
                ("fb_tol", ctypes.c_double), ("prewarp_freeze_at", ctypes.c_int),
                ("tile", ctypes.c_int)]


class _Result(ctypes.Structure):
    _fields_ = [(k, ctypes.POINTER(ctypes.c_double)) for k in
                ("ux", "uy", "conf", "raw_ux", "raw_uy", "exx", "eyy", "exy", "gmax", "vol")] + [
        ("fb", ctypes.POINTER(ctypes.c_int32)), ("t_coarse", ctypes.c_double), ("t_fine", ctypes.c_double),
        ("t_strain", ctypes.c_double), ("reference", ctypes.c_int), ("mean_conf", ctypes.c_double),
        ("n_fb_max", ctypes.c_int), ("fallback_frac", ctypes.c_double), ("fb_gate_rejected", ctypes.c_double),
        ("reacquired", ctypes.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_median.restype = ctypes.c_double
        _lib.orc_last_error.restype = ctypes.c_char_p
    return _lib


def _cfg(cfg):
    c = _Cfg()
    for name, _ in _Cfg._fields_:
        if name in ("n_scales", "scales", "seq_hybrid"):
            continue
        setattr(c, name, getattr(cfg, name))
    c.n_scales = len(cfg.scales)
    for i, s in enumerate(cfg.scales):
        c.scales[i] = s
    c.seq_hybrid = 1 if cfg.seq_mode == "hybrid" else 0
    return c


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise ValueError(lib().orc_last_error().decode())


def set_threads(n):
    lib().orc_set_threads(int(n))


def fft2(a, inverse=False):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.empty_like(a)
    lib().orc_fft2(_p(a), a.shape[0], a.shape[1], int(inverse), _p(out))
    return out


def fft1(a, inverse=False):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.empty_like(a)
    lib().orc_fft1(_p(a), a.shape[0], int(inverse), _p(out))
    return out


def bspline_coeffs(img):
    if np.iscomplexobj(img):
        a = np.ascontiguousarray(img, dtype=np.complex128)
        out = np.empty_like(a)
        lib().orc_bspline_coeffs_c(_p(a), a.shape[0], a.shape[1], _p(out))
        return out
    a = _d(img)
    out = np.empty_like(a)
    lib().orc_bspline_coeffs(_p(a), a.shape[0], a.shape[1], _p(out))
    return out


def sample(coeffs, x, y):
    c = _d(coeffs)
    x, y = _d(x), _d(y)
    out = np.empty(x.shape)
    lib().orc_sample(_p(c), c.shape[0], c.shape[1], _p(x), _p(y), ctypes.c_long(x.size), _p(out))
    return out


def warp(img, ux, uy):
    img, ux, uy = _d(img), _d(ux), _d(uy)
    out = np.empty_like(img)
    lib().orc_warp(_p(img), img.shape[0], img.shape[1], _p(ux), _p(uy), _p(out))
    return out


def gauss_blur(a, sigma):
    a = _d(a)
    out = np.empty_like(a)
    lib().orc_gauss_blur(_p(a), a.shape[0], a.shape[1], ctypes.c_double(sigma), _p(out))
    return out


def median(v):
    v = _d(v).ravel()
    return lib().orc_median(_p(v), ctypes.c_long(v.size))


def gabor_bank(shape, cfg):
    F = cfg.n_orient * len(cfg.scales)
    out = np.empty((F,) + tuple(shape))
    c = _cfg(cfg)
    lib().orc_gabor_bank(shape[0], shape[1], ctypes.byref(c), _p(out))
    return out


def reference_phase(i1, cfg):
    i1 = _d(i1)
    F = cfg.n_orient * len(cfg.scales)
    outs = [np.empty((F,) + i1.shape) for _ in range(5)]
    c = _cfg(cfg)
    lib().orc_reference_phase(_p(i1), i1.shape[0], i1.shape[1], ctypes.byref(c), *[_p(o) for o in outs])
    return outs  # r_re, r_im, gx, gy, amp


def local_solve(sums, ux, uy, sigma, order):
    sums = _d(sums)
    ux, uy = _d(ux), _d(uy)
    h, w = ux.shape
    o = [np.empty((h, w)) for _ in range(4)]
    slope = np.empty((h, w, 4))
    lib().orc_local_solve(_p(sums), _p(ux), _p(uy), h, w, ctypes.c_double(sigma), order,
                          *[_p(a) for a in o], _p(slope))
    return o[0], o[1], o[2], o[3], slope  # ux, uy, energy, resid, slope


def refine(i1, i2, ux, uy, cfg, shift=(0, 0)):
    """Returns (ux, uy, conf, resid, energy) of phaseflow.refine. i2 may be any window of the second
    frame, pixel (x, y) of i1 sitting at (x + shift[0], y + shift[1]) in it
    (rust/crates/optflow-core/src/phaseflow.rs:1441-1456)."""
    i1, i2, ux, uy = _d(i1), _d(i2), _d(ux), _d(uy)
    h, w = i1.shape
    o = [np.empty((h, w)) for _ in range(5)]
    c = _cfg(cfg)
    _check(lib().orc_refine_shifted(_p(i1), h, w, _p(i2), i2.shape[0], i2.shape[1], ctypes.c_long(shift[0]),
                                    ctypes.c_long(shift[1]), _p(ux), _p(uy), ctypes.byref(c),
                                    *[_p(a) for a in o]))
    return tuple(o)


def poc_shift(a, b):
    a, b = _d(a), _d(b)
    dx, dy, pk = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_poc_shift(_p(a), _p(b), a.shape[0], a.shape[1], ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(pk))
    return dx.value, dy.value, pk.value


def level_window(shape, win):
    wy, wx = ctypes.c_int(), ctypes.c_int()
    ok = lib().orc_level_window(shape[0], shape[1], win, ctypes.byref(wy), ctypes.byref(wx))
    return (wy.value, wx.value) if ok else None


def n_levels(max_disp, win, shape):
    return lib().orc_n_levels(ctypes.c_double(max_disp), win, shape[0], shape[1])


def coarse_flow(i1, i2, cfg, init=None):
    """Grid (ux, uy, peak) of poc.coarse_flow at the level-0 grid points."""
    i1, i2 = _d(i1), _d(i2)
    h, w = i1.shape
    c = _cfg(cfg)
    nx, ny = ctypes.c_int(), ctypes.c_int()
    lib().orc_grid_dims(h, w, ctypes.byref(c), ctypes.byref(nx), ctypes.byref(ny))
    o = [np.empty((ny.value, nx.value)) for _ in range(3)]
    ix = iy = None
    if init is not None:
        ix, iy = _d(init[0]), _d(init[1])
    lib().orc_coarse_flow(_p(i1), _p(i2), h, w, ctypes.byref(c), _p(ix) if ix is not None else None,
                          _p(iy) if iy is not None else None, *[_p(a) for a in o])
    return tuple(o)


def coarse_dense(i1, i2, cfg, init=None):
    """The dense start field (ux, uy) of track.py:131-137: coarse_flow's grid at every pixel."""
    i1, i2 = _d(i1), _d(i2)
    h, w = i1.shape
    ix = iy = None
    if init is not None:
        ix, iy = _d(init[0]), _d(init[1])
    ox, oy = np.empty((h, w)), np.empty((h, w))
    lib().orc_coarse_dense(_p(i1), _p(i2), h, w, ctypes.byref(_cfg(cfg)), _p(ix) if ix is not None else None,
                           _p(iy) if iy is not None else None, _p(ox), _p(oy))
    return ox, oy


def fit_affine(xs, ys, ux, uy):
    """tavg.rs:87-110: ((c0, c1, c2) for ux, (c0, c1, c2) for uy) of the robust affine fit, or None."""
    xs, ys, ux, uy = _d(xs), _d(ys), _d(ux), _d(uy)
    cx, cy = np.empty(3), np.empty(3)
    ok = lib().orc_fit_affine(_p(xs), _p(ys), _p(ux), _p(uy), ctypes.c_long(len(xs)), _p(cx), _p(cy))
    return (cx, cy) if ok else None


TAVG_MODES = {None: 0, "affine": 1, "dense": 2}


def align_to(centre, other, cfg, mask=None, model="affine"):
    """tavg.rs:118-163: `other` resampled onto the grid of `centre`, or None."""
    c, o = _d(centre), _d(other)
    out = np.empty_like(c)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    ok = lib().orc_align_to(_p(c), _p(o), c.shape[0], c.shape[1], TAVG_MODES[model], _p(m) if m is not None else None,
                            ctypes.byref(_cfg(cfg)), _p(out))
    return out if ok else None


def average_window(frames, centre, model=None, cfg=None, mask=None, keep_unaligned=False):
    """tavg.rs:168-207: the mean of the window (model None: plain; "affine" / "dense": every other
    frame aligned to frames[centre] first). Returns (mean, number of frames in it)."""
    fr = _d(np.stack(frames))
    nf, h, w = fr.shape
    out = np.empty((h, w))
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    from optflow_soil_amd.config import Config
    used = lib().orc_tavg_window(_p(fr), nf, centre, h, w, TAVG_MODES[model], _p(m) if m is not None else None,
                                 int(keep_unaligned), ctypes.byref(_cfg(cfg or Config())), _p(out))
    return out, used


def to_gray(pixels, factor=1):
    """io.rs:44-52, :259-281: uint8 / uint16 pixels (h, w) or (h, w, channels) -> f64 grey levels,
    box downscaled by factor."""
    a = np.ascontiguousarray(pixels)
    bits = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16}[a.dtype]
    h, w = a.shape[:2]
    ch = 1 if a.ndim == 2 else a.shape[2]
    out = np.empty((h // factor, w // factor)) if factor > 1 else np.empty((h, w))
    lib().orc_to_gray(a.ctypes.data_as(ctypes.c_void_p), bits, ch, h, w, factor, _p(out))
    return out


def edt(inp, return_indices=False):
    """scipy.ndimage.distance_transform_edt(inp) (and the flat nearest-zero index)."""
    b = np.ascontiguousarray(inp, dtype=np.uint8)
    d = np.empty(b.shape)
    idx = np.empty(b.shape, dtype=np.int32) if return_indices else None
    lib().orc_edt(_p(b), b.shape[0], b.shape[1], _p(d), _p(idx) if idx is not None else None)
    return (d, idx) if return_indices else d


def large_components(b, min_area):
    b = np.ascontiguousarray(b, dtype=np.uint8)
    out = np.empty_like(b)
    lib().orc_large_components(_p(b), b.shape[0], b.shape[1], ctypes.c_double(min_area), _p(out))
    return out.astype(bool)


def displacement_gradient(ux, uy, win):
    ux, uy = _d(ux), _d(uy)
    h, w = ux.shape
    o = [np.empty((h, w)) for _ in range(4)]
    lib().orc_displacement_gradient(_p(ux), _p(uy), h, w, win, *[_p(a) for a in o])
    return tuple(o)


def strain_fields(dudx, dudy, dvdx, dvdy):
    exx = dudx + 0.5 * (dudx ** 2 + dvdx ** 2)
    eyy = dvdy + 0.5 * (dudy ** 2 + dvdy ** 2)
    exy = 0.5 * (dudy + dvdx) + 0.5 * (dudx * dudy + dvdx * dvdy)
    return {"Exx": exx, "Eyy": eyy, "Exy": exy, "gmax": 2.0 * np.sqrt(((exx - eyy) / 2) ** 2 + exy ** 2),
            "vol": (1 + dudx) * (1 + dvdy) - dudy * dvdx - 1.0}


_RES_KEYS = ("ux", "uy", "conf", "raw_ux", "raw_uy", "exx", "eyy", "exy", "gmax", "vol")


def _new_result(h, w):
    arrs = {k: np.empty((h, w)) for k in _RES_KEYS}
    r = _Result()
    for k in _RES_KEYS:
        setattr(r, k, arrs[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    arrs["fb"] = np.zeros((h, w), np.int32)
    r.fb = arrs["fb"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    return r, arrs


def _as_dict(r, arrs):
    out = dict(arrs)
    out["Exx"], out["Eyy"], out["Exy"] = out.pop("exx"), out.pop("eyy"), out.pop("exy")
    out["timing"] = {"coarse": r.t_coarse, "fine": r.t_fine, "strain": r.t_strain, "reference": r.reference,
                     "mean_conf": r.mean_conf, "n_fb_max": r.n_fb_max, "fallback_frac": r.fallback_frac,
                     "fb_gate_rejected": r.fb_gate_rejected, "reacquired": r.reacquired}
    return out


def track_pair(i1, i2, cfg, mask=None, init=None):
    i1, i2 = _d(i1), _d(i2)
    h, w = i1.shape
    r, arrs = _new_result(h, w)
    c = _cfg(cfg)
    m = np.ascontiguousarray(mask, dtype=np.uint8) if mask is not None else None
    ix = iy = None
    if init is not None:
        ix, iy = _d(init[0]), _d(init[1])
    _check(lib().orc_track_pair(_p(i1), _p(i2), h, w, _p(m) if m is not None else None,
                                _p(ix) if ix is not None else None, _p(iy) if iy is not None else None,
                                ctypes.byref(c), ctypes.byref(r)))
    return _as_dict(r, arrs)


def box_sum(a, r):
    """Sum over the (2r + 1)^2 window, zero outside (rust/crates/optflow-core/src/track/hybrid.rs:67-112)."""
    a = _d(a)
    out = np.empty_like(a)
    lib().orc_box_sum(_p(a), a.shape[0], a.shape[1], int(r), _p(out))
    return out


def local_affine_agreement(ux, uy, support, tx, ty, radius, min_support, tol):
    """The hybrid gate (rust/crates/optflow-core/src/track/hybrid.rs:147-260)."""
    ux, uy, tx, ty = _d(ux), _d(uy), _d(tx), _d(ty)
    sup = np.ascontiguousarray(support, dtype=np.uint8)
    ok = np.empty(ux.shape, np.uint8)
    lib().orc_local_affine_agreement(_p(ux), _p(uy), _p(sup), _p(tx), _p(ty), ux.shape[0], ux.shape[1],
                                     ctypes.c_double(radius), ctypes.c_long(int(min_support)), ctypes.c_double(tol),
                                     _p(ok))
    return ok.astype(bool)


def track_sequence(frames, cfg, mask=None):
    fr = _d(np.stack(frames))
    n, h, w = fr.shape
    rs = (_Result * (n - 1))()
    keep = []
    for k in range(n - 1):
        r, arrs = _new_result(h, w)
        rs[k] = r
        keep.append(arrs)
    c = _cfg(cfg)
    m = np.ascontiguousarray(mask, dtype=np.uint8) if mask is not None else None
    _check(lib().orc_track_sequence(_p(fr), n, h, w, _p(m) if m is not None else None, ctypes.byref(c), rs))
    return [_as_dict(rs[k], keep[k]) for k in range(n - 1)]
