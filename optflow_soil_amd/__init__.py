"""MI355X dense flow engine: Python binding of libofx.so (include/ofx.h).

The functions mirror the reference's Python API (python/optflow_ref/track.py:115 `track_pair`,
:260 `track_sequence`, phaseflow.py:427 `refine`) with the same argument meaning and the same
error behaviour (ValueError for bad input or settings, RuntimeError for device failures; no CPU
fallback). Every computation runs on the GPU through the C ABI; this module only moves arrays.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from .config import Config

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libofx.so")

OFX_OK, OFX_EINVAL, OFX_ETOOLARGE, OFX_EDEVICE, OFX_ENODEVICE, OFX_EBUSY, OFX_EUNSUPPORTED = range(7)

# every symbol include/ofx.h declares
SYMBOLS = ("ofx_config_default", "ofx_create", "ofx_destroy", "ofx_last_error", "ofx_lanes", "ofx_returns_slope",
           "ofx_check_window", "ofx_release_caches", "ofx_refine_raw", "ofx_coarse_flow", "ofx_track_pair",
           "ofx_track_pair_device", "ofx_track_pairs", "ofx_seq_begin", "ofx_seq_step", "ofx_seq_step_device", "ofx_seq_end",
           "ofx_tavg_window", "ofx_frame_to_gray", "ofx_median", "ofx_median_device_noalloc",
           "ofx_track_sequence", "ofx_profile", "ofx_profile_read", "ofx_profile_stage", "ofx_crop_plan",
           "ofx_write_crop_json", "ofx_write_result", "ofx_pyramid_levels", "ofx_poc_level_window",
           "ofx_format_json_number", "ofx_coarse_net_params", "ofx_coarse_net_load", "ofx_coarse_net_propose",
           "ofx_coarse_net_gradient", "ofx_coarse_net_train", "ofx_graph_stats", "ofx_set_schedule",
           "ofx_get_schedule")


class OfxConfig(ctypes.Structure):
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
                ("tile", ctypes.c_int), ("coarse_learned", ctypes.c_int), ("coarse_gate", ctypes.c_double)]


_FP = ctypes.POINTER(ctypes.c_float)
_RES_F32 = ("ux", "uy", "conf", "raw_ux", "raw_uy", "exx", "eyy", "exy", "gmax", "vol")


class OfxResult(ctypes.Structure):
    _fields_ = [(k, _FP) for k in _RES_F32] + [
        ("valid", ctypes.POINTER(ctypes.c_uint8)), ("fb", ctypes.POINTER(ctypes.c_int32)),
        ("t_coarse", ctypes.c_double), ("t_fine", ctypes.c_double), ("t_strain", ctypes.c_double),
        ("reference", ctypes.c_int), ("mean_conf", ctypes.c_double), ("n_fb_max", ctypes.c_int),
        ("fallback_frac", ctypes.c_double), ("fb_gate_rejected", ctypes.c_double), ("reacquired", ctypes.c_double),
        ("coarse_confirmed", ctypes.c_double), ("coarse_fallback", ctypes.c_int)]


_lib = None


def lib():
    """The loaded libofx.so; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m optflow_soil_amd.build` "
                               "(or __graft_entry__.build()); this engine has no CPU fallback")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.ofx_last_error.restype = ctypes.c_char_p
        _lib.ofx_profile_stage.restype = ctypes.c_char_p
        _lib.ofx_coarse_net_params.restype = ctypes.c_size_t
    return _lib


def to_c(cfg: Config) -> OfxConfig:
    cfg.validate()
    c = OfxConfig()
    for name, _ in OfxConfig._fields_:
        if name in ("n_scales", "scales", "seq_hybrid", "coarse_learned"):
            continue
        setattr(c, name, getattr(cfg, name))
    c.n_scales = len(cfg.scales)
    for i, s in enumerate(cfg.scales):
        c.scales[i] = s
    c.seq_hybrid = 1 if cfg.seq_mode == "hybrid" else 0
    c.coarse_learned = 1 if cfg.coarse_model == "learned" else 0
    return c


def _check(rc):
    if rc == OFX_OK:
        return
    msg = lib().ofx_last_error().decode()
    if rc in (OFX_EINVAL, OFX_ETOOLARGE):
        raise ValueError(msg)
    if rc == OFX_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Result:
    """One pair or sequence frame (track.py:15-42): ux, uy NaN where invalid, conf before masking,
    strain dict (Exx, Eyy, Exy, gmax, vol), raw (unmasked) field, valid mask, timing."""
    ux: np.ndarray
    uy: np.ndarray
    conf: np.ndarray
    strain: dict
    raw_ux: np.ndarray
    raw_uy: np.ndarray
    valid: np.ndarray
    timing: dict = field(default_factory=dict)
    fb: np.ndarray = None  # hybrid sequences: consecutive frames carried by the adjacent leg


class OfxSchedule(ctypes.Structure):
    _fields_ = [("tile_groups", ctypes.c_int), ("graphs", ctypes.c_int)]


class OfxNetTrain(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("batch", ctypes.c_int), ("crop", ctypes.c_int), ("n_pairs", ctypes.c_int),
                ("lr", ctypes.c_double), ("lr_final", ctypes.c_double), ("max_disp", ctypes.c_double),
                ("noise", ctypes.c_double), ("seed", ctypes.c_uint64)]


COARSE_NET_FILE = os.path.join(_HERE, "weights", "coarse_net.npy")


def coarse_net_params():
    """Number of parameters of the learned coarse network (flat layout of include/ofx.h)."""
    return int(lib().ofx_coarse_net_params())


class OfxPair(ctypes.Structure):
    _fields_ = [("i1", ctypes.c_void_p), ("i2", ctypes.c_void_p), ("w", ctypes.c_int), ("h", ctypes.c_int),
                ("mask", ctypes.c_void_p), ("init_ux", ctypes.c_void_p), ("init_uy", ctypes.c_void_p)]


def _new_result(h, w, with_fb=False):
    arrs = {k: np.empty((h, w), np.float32) for k in _RES_F32}
    valid = np.empty((h, w), np.uint8)
    r = OfxResult()
    for k in _RES_F32:
        setattr(r, k, arrs[k].ctypes.data_as(_FP))
    r.valid = valid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    r.fb = None
    if with_fb:
        arrs["fb"] = np.zeros((h, w), np.int32)
        r.fb = arrs["fb"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    return r, arrs, valid


def _to_result(r, arrs, valid):
    strain = {"Exx": arrs["exx"], "Eyy": arrs["eyy"], "Exy": arrs["exy"], "gmax": arrs["gmax"], "vol": arrs["vol"]}
    timing = {"coarse": r.t_coarse, "fine": r.t_fine, "strain": r.t_strain, "reference": r.reference,
              "mean_conf": r.mean_conf, "device": "gpu"}
    if r.coarse_confirmed >= 0:
        timing.update(coarse_confirmed=r.coarse_confirmed, coarse_fallback=bool(r.coarse_fallback))
    fb = arrs.get("fb")
    if fb is not None:
        timing.update(n_fb_max=r.n_fb_max, fallback_frac=r.fallback_frac, fb_gate_rejected=r.fb_gate_rejected,
                      reacquired=r.reacquired)
    return Result(arrs["ux"], arrs["uy"], arrs["conf"], strain, arrs["raw_ux"], arrs["raw_uy"], valid.astype(bool),
                  timing, fb)


class Engine:
    """A GPU engine (one HIP stream per lane). Mirrors FineEngine (core/engine.rs:31-125)."""

    def __init__(self, device=0, lanes=1):
        self._h = ctypes.c_void_p()
        _check(lib().ofx_create(int(device), int(lanes), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().ofx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def lanes(self):
        return lib().ofx_lanes(self._h)

    def profile(self, enable=True):
        """Start (and zero) or stop the per-stage GPU timer."""
        _check(lib().ofx_profile(self._h, int(enable)))

    def profile_read(self):
        """{stage: (milliseconds, calls)} since profile(True)."""
        names = []
        while True:
            nm = lib().ofx_profile_stage(len(names))
            if nm is None:
                break
            names.append(nm.decode())
        ms = (ctypes.c_double * len(names))()
        calls = (ctypes.c_int64 * len(names))()
        _check(lib().ofx_profile_read(self._h, ms, calls, len(names)))
        return {n: (ms[i], calls[i]) for i, n in enumerate(names)}

    def schedule(self):
        """The engine's schedule: {"tile_groups": n, "graphs": bool}."""
        sc = OfxSchedule()
        _check(lib().ofx_get_schedule(self._h, ctypes.byref(sc)))
        return {"tile_groups": sc.tile_groups, "graphs": bool(sc.graphs)}

    def set_schedule(self, tile_groups=None, graphs=None):
        """Changes the schedule knobs given (ofx_set_schedule); results do not depend on them."""
        cur = self.schedule()
        sc = OfxSchedule(int(cur["tile_groups"] if tile_groups is None else tile_groups),
                         int(cur["graphs"] if graphs is None else graphs))
        _check(lib().ofx_set_schedule(self._h, ctypes.byref(sc)))

    def graph_stats(self):
        """(captures, replays) of the fine stage's HIP graphs on this engine so far."""
        c, l = ctypes.c_int64(0), ctypes.c_int64(0)
        _check(lib().ofx_graph_stats(self._h, ctypes.byref(c), ctypes.byref(l)))
        return c.value, l.value

    def check_window(self, w, h, cfg):
        _check(lib().ofx_check_window(self._h, int(w), int(h), ctypes.byref(to_c(cfg))))

    def release_caches(self):
        _check(lib().ofx_release_caches(self._h))

    def refine_raw(self, i1, i2, ux, uy, cfg, shift=(0, 0)):
        """FineEngine::refine_raw (rust/crates/optflow-core/src/engine.rs:31-41): the fine stage from
        the start field (ux, uy) on i1's grid; i2 is any window of the second frame, pixel (x, y) of
        i1 sitting at (x + shift[0], y + shift[1]) in it. Returns (ux, uy, energy, resid), float64."""
        i1, i2, ux, uy = _f32(i1), _f32(i2), _f32(ux), _f32(uy)
        h, w = i1.shape
        if ux.shape != (h, w) or uy.shape != (h, w):
            raise ValueError(f"refine: i1 is {w}x{h} but the field is {ux.shape[1]}x{ux.shape[0]}")
        o = [np.empty((h, w)) for _ in range(4)]
        c = to_c(cfg)
        _check(lib().ofx_refine_raw(self._h, _p(i1), w, h, _p(i2), i2.shape[1], i2.shape[0],
                                    ctypes.c_int64(int(shift[0])), ctypes.c_int64(int(shift[1])), _p(ux), _p(uy),
                                    ctypes.byref(c), *[_p(a) for a in o]))
        return tuple(o)

    def coarse_flow(self, I1, I2, cfg, init=None):
        """The coarse stage (poc.py:216-239) and the dense start field of track.py:131-137: the
        (ux, uy) field, float64, that the fine stage of a pair starts from."""
        I1, I2 = _f32(I1), _f32(I2)
        if I1.shape != I2.shape:
            raise ValueError(f"image sizes differ: {I1.shape} vs {I2.shape}")
        h, w = I1.shape
        ix = iy = None
        if init is not None:
            ix, iy = _f32(init[0]), _f32(init[1])
        ox, oy = np.empty((h, w)), np.empty((h, w))
        _check(lib().ofx_coarse_flow(self._h, _p(I1), _p(I2), w, h, _p(ix), _p(iy), ctypes.byref(to_c(cfg)),
                                     _p(ox), _p(oy)))
        return ox, oy

    def track_pair(self, I1, I2, cfg, mask=None, init=None) -> Result:
        """track.py:115-146: displacement of I1's pixels into I2."""
        I1, I2 = _f32(I1), _f32(I2)
        if I1.shape != I2.shape:
            raise ValueError(f"image sizes differ: {I1.shape} vs {I2.shape}")
        if mask is not None and np.shape(mask) != I1.shape:
            raise ValueError(f"ROI mask shape {np.shape(mask)} differs from the image shape {I1.shape}")
        h, w = I1.shape
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        ix = iy = None
        if init is not None:
            ix, iy = _f32(init[0]), _f32(init[1])
        r, arrs, valid = _new_result(h, w)
        c = to_c(cfg)
        _check(lib().ofx_track_pair(self._h, _p(I1), _p(I2), w, h, _p(m), _p(ix), _p(iy), ctypes.byref(c),
                                    ctypes.byref(r)))
        return _to_result(r, arrs, valid)

    def track_pairs(self, pairs, cfg):
        """Independent pairs [(I1, I2), (I1, I2, mask), (I1, I2, mask, init), ...] of any sizes,
        spread over the engine's lanes (ofx_track_pairs); returns a Result per pair, each
        bit-identical to track_pair on it."""
        n = len(pairs)
        cp = (OfxPair * max(n, 1))()
        rs = (OfxResult * max(n, 1))()
        keep = []
        for k, p in enumerate(pairs):
            I1, I2 = _f32(p[0]), _f32(p[1])
            if I1.shape != I2.shape:
                raise ValueError(f"pair {k}: image sizes differ: {I1.shape} vs {I2.shape}")
            h, w = I1.shape
            mask = p[2] if len(p) > 2 else None
            init = p[3] if len(p) > 3 else None
            m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
            if m is not None and m.shape != (h, w):
                raise ValueError(f"pair {k}: ROI mask shape {m.shape} differs from the image shape {(h, w)}")
            ix = iy = None
            if init is not None:
                ix, iy = _f32(init[0]), _f32(init[1])
            r, arrs, valid = _new_result(h, w)
            cp[k] = OfxPair(_p(I1), _p(I2), w, h, _p(m), _p(ix), _p(iy))
            rs[k] = r
            keep.append((I1, I2, m, ix, iy, arrs, valid))
        _check(lib().ofx_track_pairs(self._h, n, cp, ctypes.byref(to_c(cfg)), rs))
        return [_to_result(rs[k], keep[k][5], keep[k][6]) for k in range(n)]

    def track_pair_device(self, d_i1, d_i2, w, h, cfg, d_out, d_mask=0, d_init=(0, 0)):
        """Device-pointer form (integers from e.g. torch's data_ptr()); d_out maps result names to
        device pointers. Returns the stage times."""
        r = OfxResult()
        for k in _RES_F32:
            setattr(r, k, ctypes.cast(ctypes.c_void_p(d_out.get(k, 0) or None), _FP))
        r.valid = ctypes.cast(ctypes.c_void_p(d_out.get("valid", 0) or None), ctypes.POINTER(ctypes.c_uint8))
        c = to_c(cfg)
        _check(lib().ofx_track_pair_device(self._h, ctypes.c_void_p(d_i1), ctypes.c_void_p(d_i2), int(w), int(h),
                                           ctypes.c_void_p(d_mask or None), ctypes.c_void_p(d_init[0] or None),
                                           ctypes.c_void_p(d_init[1] or None), ctypes.byref(c), ctypes.byref(r)))
        return {"coarse": r.t_coarse, "fine": r.t_fine, "strain": r.t_strain, "coarse_confirmed": r.coarse_confirmed,
                "coarse_fallback": bool(r.coarse_fallback)}

    def track_sequence(self, frames, cfg, mask=None):
        """track.py:260-315 (seq_mode chain) and track.py:383-505 (hybrid; results carry fb and the
        hybrid timing shares): list of (k, Result) for k = 1 .. n - 1."""
        fr = _f32(np.stack(frames))
        n, h, w = fr.shape
        if mask is not None and np.shape(mask) != (h, w):
            raise ValueError(f"ROI mask shape {np.shape(mask)} differs from the frame shape {(h, w)}")
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        if m is not None and not m.any():
            raise ValueError("ROI mask is empty")
        rs = (OfxResult * max(n - 1, 1))()
        keep = []
        hybrid = cfg.seq_mode == "hybrid"
        for k in range(n - 1):
            r, arrs, valid = _new_result(h, w, with_fb=hybrid)
            rs[k] = r
            keep.append((arrs, valid))
        c = to_c(cfg)
        _check(lib().ofx_track_sequence(self._h, _p(fr), n, w, h, _p(m), ctypes.byref(c), rs))
        return [(k + 1, _to_result(rs[k], *keep[k])) for k in range(n - 1)]

    def tavg_window(self, frames, centre, model=None, cfg=None, mask=None, keep_unaligned=False):
        """tavg.rs:168-207 / tavg.py:78-96 on the GPU: (the mean of the window as float32, the
        number of frames in it). model None: plain; "affine" / "dense": every other frame aligned
        to frames[centre] first (tavg.rs:118-163)."""
        fr = _f32(np.stack(frames))
        nf, h, w = fr.shape
        align = {None: 0, "affine": 1, "dense": 2}[model]
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        if m is not None and m.shape != (h, w):
            raise ValueError(f"ROI mask shape {m.shape} differs from the frame shape {(h, w)}")
        out = np.empty((h, w), np.float32)
        used = ctypes.c_int()
        c = to_c(cfg or Config())
        _check(lib().ofx_tavg_window(self._h, _p(fr), nf, int(centre), w, h, align, _p(m), int(bool(keep_unaligned)),
                                     ctypes.byref(c), _p(out), ctypes.byref(used)))
        return out, used.value

    def frame_to_gray(self, pixels, factor=1):
        """io.rs:44-52, :259-281 on the GPU: uint8 / uint16 pixels (H, W) or (H, W, 1|3|4) ->
        float32 grey levels in [0, 1], box downscaled by the integer factor."""
        a = np.ascontiguousarray(pixels)
        if a.dtype not in (np.uint8, np.uint16):
            if np.issubdtype(a.dtype, np.floating) and a.ndim == 2 and factor == 1:
                return a.astype(np.float32)
            raise ValueError(f"unsupported pixel type {a.dtype}")
        if a.ndim not in (2, 3):
            raise ValueError(f"unsupported image shape {a.shape}")
        h, w = a.shape[:2]
        ch = 1 if a.ndim == 2 else a.shape[2]
        if factor < 1:
            raise ValueError(f"scale must be >= 1, got {factor}")
        out = np.empty((h // factor, w // factor), np.float32)
        _check(lib().ofx_frame_to_gray(self._h, a.ctypes.data_as(ctypes.c_void_p), 8 * a.itemsize, ch, w, h,
                                       int(factor), _p(out)))
        return out

    def median(self, values):
        """select.rs:19-42 on the GPU: numpy.median of a float64 array (NaN if any value is NaN or
        the array is empty)."""
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        out = ctypes.c_double(0.0)
        _check(lib().ofx_median(self._h, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.size),
                                ctypes.byref(out)))
        return out.value

    def median_device_noalloc(self, d_values, n, d_out):
        """ofx_median_device_noalloc: the median of n float64 values at device address d_values
        written to the device address d_out, with no device allocation (integers from e.g.
        torch's data_ptr())."""
        _check(lib().ofx_median_device_noalloc(self._h, ctypes.c_void_p(d_values or None), ctypes.c_size_t(int(n)),
                                               ctypes.c_void_p(d_out or None)))

    def load_coarse_net(self, params=None):
        """Loads the learned coarse network (coarse_model "learned"): a flat float32 parameter
        vector, or by default the trained one shipped with the package (weights/coarse_net.npy)."""
        if params is None:
            if not os.path.exists(COARSE_NET_FILE):
                raise RuntimeError(f"{COARSE_NET_FILE} is missing: train the network with "
                                   "`python -m optflow_soil_amd.train_coarse_net`")
            params = np.load(COARSE_NET_FILE, allow_pickle=False)
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        _check(lib().ofx_coarse_net_load(self._h, p.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(p.size)))

    def coarse_net_propose(self, i1, i2):
        """The loaded network's proposal for a pair: ((h + 7) // 8, (w + 7) // 8, 2), pixels."""
        i1, i2 = _f32(i1), _f32(i2)
        h, w = i1.shape
        if i2.shape != (h, w):
            raise ValueError("images differ in size")
        out = np.empty(((h + 7) // 8, (w + 7) // 8, 2), np.float32)
        _check(lib().ofx_coarse_net_propose(self._h, _p(i1), _p(i2), w, h, _p(out)))
        return out

    def coarse_net_gradient(self, params, a1, a2, target):
        """Loss, parameter gradient (flat) and prediction of one forward and backward pass on a
        batch of cell maps a1, a2 (nb, h8, w8) with target (nb, h8, w8, 2), in cells."""
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        a1, a2 = _f32(a1), _f32(a2)
        t = _f32(target)
        nb, h8, w8 = a1.shape
        if a2.shape != a1.shape or t.shape != (nb, h8, w8, 2):
            raise ValueError("a1, a2 must be (nb, h8, w8) and target (nb, h8, w8, 2)")
        loss = ctypes.c_double(0)
        grad = np.empty_like(p)
        pred = np.empty((nb, h8, w8, 2), np.float32)
        _check(lib().ofx_coarse_net_gradient(self._h, _p(p), ctypes.c_size_t(p.size), _p(a1), _p(a2), nb, w8, h8,
                                             _p(t), ctypes.byref(loss), _p(grad), _p(pred)))
        return loss.value, grad, pred

    def coarse_net_train(self, params, patterns, steps, batch=32, crop=512, n_pairs=20000, lr=2e-3, lr_final=2e-5,
                         max_disp=48.0, noise=0.01, seed=0):
        """Trains the network on the GPU on pairs from the renderer (ofx_coarse_net_train); returns
        (trained parameters, loss of every step)."""
        p = np.ascontiguousarray(params, dtype=np.float32).ravel().copy()
        pats = _f32(patterns)
        if pats.ndim == 2:
            pats = pats[None]
        n_pat, ph, pw = pats.shape
        t = OfxNetTrain(int(steps), int(batch), int(crop), int(n_pairs), float(lr), float(lr_final), float(max_disp),
                        float(noise), int(seed))
        loss = np.zeros(int(steps), np.float64)
        _check(lib().ofx_coarse_net_train(self._h, ctypes.byref(t), _p(np.ascontiguousarray(pats)), n_pat, pw, ph,
                                          _p(p), _p(loss)))
        return p, loss

    def sequence(self, frame0, cfg, mask=None):
        """Opens a streaming sequence on frame 0 (the Sequence iterator of track.rs:1112-1225): a
        Sequence whose step(frame) returns the Result of the next frame. It holds one lane until
        closed."""
        return Sequence(self, frame0, cfg, mask)

    def iter_sequence(self, frames, cfg, mask=None):
        """track.py:260-315 / 383-505 as a generator over any iterable of frames: yields (k, Result)
        for k = 1, 2, ... as each frame arrives, one frame in memory at a time."""
        it = iter(frames)
        try:
            f0 = next(it)
        except StopIteration:
            raise ValueError("no frames") from None
        with self.sequence(f0, cfg, mask) as q:
            for k, f in enumerate(it, start=1):
                yield k, q.step(f)


class Sequence:
    """A sequence in progress on one lane of an Engine (ofx_seq_begin / ofx_seq_step / ofx_seq_end)."""

    def __init__(self, engine, frame0, cfg, mask=None):
        f0 = _f32(frame0)
        h, w = f0.shape
        if mask is not None and np.shape(mask) != (h, w):
            raise ValueError(f"ROI mask shape {np.shape(mask)} differs from the frame shape {(h, w)}")
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        self._engine = engine  # keeps the engine alive while the sequence holds its lane
        self._shape = (h, w)
        self._hybrid = cfg.seq_mode == "hybrid"
        self._h = ctypes.c_void_p()
        _check(lib().ofx_seq_begin(engine._h, _p(f0), w, h, _p(m), ctypes.byref(to_c(cfg)), ctypes.byref(self._h)))

    def step(self, frame) -> Result:
        if not self._h:
            raise RuntimeError("sequence is closed")
        f = _f32(frame)
        if f.shape != self._shape:
            raise ValueError(f"frame shape {f.shape} differs from frame 0's {self._shape}")
        r, arrs, valid = _new_result(*self._shape, with_fb=self._hybrid)
        _check(lib().ofx_seq_step(self._h, _p(f), ctypes.byref(r)))
        return _to_result(r, arrs, valid)

    def step_device(self, d_frame, d_out):
        """ofx_seq_step_device: the next frame from a device pointer (an integer, e.g. torch's
        data_ptr()), the results into the device arrays of d_out (result name -> pointer, "fb"
        for the hybrid counter). Returns the timing and sequence figures."""
        if not self._h:
            raise RuntimeError("sequence is closed")
        r = OfxResult()
        for k in _RES_F32:
            setattr(r, k, ctypes.cast(ctypes.c_void_p(d_out.get(k, 0) or None), _FP))
        r.valid = ctypes.cast(ctypes.c_void_p(d_out.get("valid", 0) or None), ctypes.POINTER(ctypes.c_uint8))
        r.fb = ctypes.cast(ctypes.c_void_p(d_out.get("fb", 0) or None), ctypes.POINTER(ctypes.c_int32))
        _check(lib().ofx_seq_step_device(self._h, ctypes.c_void_p(d_frame), ctypes.byref(r)))
        t = {"coarse": r.t_coarse, "fine": r.t_fine, "strain": r.t_strain, "reference": r.reference,
             "mean_conf": r.mean_conf}
        if self._hybrid:
            t.update(n_fb_max=r.n_fb_max, fallback_frac=r.fallback_frac, fb_gate_rejected=r.fb_gate_rejected,
                     reacquired=r.reacquired)
        return t

    def close(self):
        if self._h:
            lib().ofx_seq_end(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_engine() -> Engine:
    global _default
    if _default is None:
        _default = Engine()
    return _default


def track_pair(I1, I2, cfg=None, mask=None, init=None) -> Result:
    return default_engine().track_pair(I1, I2, cfg or Config(), mask=mask, init=init)


def track_sequence(frames, cfg=None, mask=None):
    return default_engine().track_sequence(frames, cfg or Config(), mask=mask)


def refine(I1, I2, ux, uy, cfg=None, shift=(0, 0)):
    return default_engine().refine_raw(I1, I2, ux, uy, cfg or Config(), shift)
