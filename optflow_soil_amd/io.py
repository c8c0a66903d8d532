"""Writing results in the reference's output-folder format, and the automatic ROI crop.

Both run in libofx.so (host code, no device work), the same functions the command-line twin
uses, so a folder has the same bytes whichever front end wrote it:

- write_result: one result folder (rust/crates/optflow-cli/src/run.rs:74-110 and io.rs:557-621):
  ux, uy (NaN = invalid), conf, Exx, Eyy, Exy, gmax, vol as '<f4' .npy, valid as '|u1', fb as
  '<i4' for hybrid sequences, raw_ux / raw_uy on request, timing.json; a crop's result is written
  as full-frame maps (NaN outside the crop, conf and fb 0). The .npy bytes are numpy.save's.
- CropWindow.plan: the automatic ROI crop of the reference CLI (crop.rs:43-130); cut() takes an
  input down to the crop, to_json() / write_json() give crop.json.
- frame_folder: the folder of sequence frame k (frame_00001, ..., run.rs:687).
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import OfxResult, Result, _FP, _check, lib, to_c


class OfxCrop(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_int), ("y0", ctypes.c_int), ("x1", ctypes.c_int), ("y1", ctypes.c_int),
                ("margin", ctypes.c_int64)]


def pyramid_levels(cfg, shape):
    """The coarse stage's pyramid depth for a frame of this (h, w) shape (poc.rs:628-666)."""
    out = ctypes.c_int(0)
    _check(lib().ofx_pyramid_levels(ctypes.byref(to_c(cfg)), int(shape[1]), int(shape[0]), ctypes.byref(out)))
    return out.value


def poc_level_window(shape, win):
    """The POC window (wy, wx) at a pyramid level of this (h, w) shape, or None (poc.rs:610-624)."""
    wy, wx = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().ofx_poc_level_window(int(shape[1]), int(shape[0]), int(win), ctypes.byref(wy), ctypes.byref(wx)))
    return (wy.value, wx.value) if wy.value else None


def json_number(v):
    """v spelled as the reference CLI's JSON writer spells it (ofx_format_json_number)."""
    buf = ctypes.create_string_buffer(64)
    _check(lib().ofx_format_json_number(ctypes.c_double(v), buf, ctypes.c_size_t(64)))
    return buf.value.decode()


@dataclass(frozen=True)
class CropWindow:
    """A half-open rectangle [x0, x1) x [y0, y1) of the frame that is tracked instead of the whole
    frame, and the margin it was grown by."""
    x0: int
    y0: int
    x1: int
    y1: int
    margin: int

    @staticmethod
    def plan(mask, cfg, margin=None, seq=False):
        """The ROI's bounding box grown by margin (None: the reference CLI's default margin),
        its origin on the frame's coarse grid, clipped to the frame (ofx_crop_plan)."""
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        h, w = m.shape
        out = OfxCrop()
        mg = None if margin is None else ctypes.byref(ctypes.c_int64(int(margin)))
        _check(lib().ofx_crop_plan(m.ctypes.data_as(ctypes.c_void_p), w, h, ctypes.byref(to_c(cfg)), mg,
                                   int(bool(seq)), ctypes.byref(out)))
        return CropWindow(out.x0, out.y0, out.x1, out.y1, out.margin)

    @property
    def shape(self):
        return self.y1 - self.y0, self.x1 - self.x0

    def cut(self, a):
        """The part of a full-frame array inside the crop, contiguous."""
        return np.ascontiguousarray(a[self.y0:self.y1, self.x0:self.x1])

    def _c(self):
        return OfxCrop(self.x0, self.y0, self.x1, self.y1, self.margin)

    def to_json(self):
        return ("{\n" + "".join(f'  "{k}": {getattr(self, k)},\n' for k in ("x0", "y0", "x1", "y1", "margin"))
                + '  "mode": "auto"\n}')

    @staticmethod
    def write_json(folder, crop=None):
        """crop.json into folder: the crop, or {"mode": "off"} without one."""
        c = None if crop is None else ctypes.byref(crop._c())
        _check(lib().ofx_write_crop_json(os.fsencode(folder), c))


def write_result(folder, res: Result, raw=False, crop: CropWindow = None, frame_shape=None, sequence=False):
    """The result folder of one pair or sequence frame. res is the whole frame's result, or with
    crop the crop's, written as maps of frame_shape (h, w). sequence: timing.json also carries the
    reference frame and mean confidence (and the hybrid shares when res.fb is set)."""
    h, w = res.ux.shape
    keep = []

    def ptr(a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.shape != (h, w):
            raise ValueError(f"result map of shape {a.shape}, expected {(h, w)}")
        keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype)))

    r = OfxResult()
    r.ux, r.uy, r.conf = ptr(res.ux, np.float32), ptr(res.uy, np.float32), ptr(res.conf, np.float32)
    r.exx, r.eyy, r.exy = (ptr(res.strain[k], np.float32) for k in ("Exx", "Eyy", "Exy"))
    r.gmax, r.vol = ptr(res.strain["gmax"], np.float32), ptr(res.strain["vol"], np.float32)
    r.raw_ux = ptr(res.raw_ux, np.float32) if raw else ctypes.cast(None, _FP)
    r.raw_uy = ptr(res.raw_uy, np.float32) if raw else ctypes.cast(None, _FP)
    r.fb = ptr(res.fb, np.int32) if res.fb is not None else ctypes.cast(None, ctypes.POINTER(ctypes.c_int32))
    t = res.timing or {}
    r.t_coarse, r.t_fine, r.t_strain = (float(t.get(k, 0.0)) for k in ("coarse", "fine", "strain"))
    r.reference, r.mean_conf = int(t.get("reference", 0) or 0), float(t.get("mean_conf", 0.0) or 0.0)
    r.n_fb_max = int(t.get("n_fb_max", 0))
    r.fallback_frac, r.fb_gate_rejected, r.reacquired = (float(t.get(k, 0.0))
                                                         for k in ("fallback_frac", "fb_gate_rejected", "reacquired"))
    if crop is None:
        c, fh, fw = None, h, w
    else:
        if frame_shape is None:
            raise ValueError("a crop's result needs the frame shape")
        c, (fh, fw) = ctypes.byref(crop._c()), frame_shape
    _check(lib().ofx_write_result(os.fsencode(folder), ctypes.byref(r), w, h, c, int(fw), int(fh), int(bool(raw)),
                                  int(bool(sequence))))


def frame_folder(out_dir, k):
    """The folder of sequence frame k."""
    return os.path.join(out_dir, f"frame_{k:05d}")
