"""Output contract of a tracked pair or sequence frame, and the automatic ROI crop.

- save_result writes the reference's output folder (python/optflow_ref/io.py:203-223,
  rust/crates/optflow-cli/src/run.rs:74-110): ux, uy (NaN = invalid), conf and the strain fields
  Exx, Eyy, Exy, gmax, vol as '<f4' .npy; valid as 'u1' (1 where ux is finite); fb as '<i4' for
  hybrid sequences; raw_ux, raw_uy with raw=True; timing.json. The .npy files are numpy's format
  1.0, the bytes numpy.save writes (rust/crates/optflow-core/src/npy.rs:95-136 writes the same).
- Crop is the automatic ROI crop (python/optflow_ref/crop.py:1-140): track the ROI's bounding box
  plus a margin, then embed the result in the full frame (NaN outside the crop, confidence 0).
- sequence_dir names a sequence frame's folder (frame_00001, ..., rust/crates/optflow-cli/src/run.rs).
"""
import json
import math
import os
from dataclasses import dataclass

import numpy as np

from . import Result

PAD = 32  # reflect pad of the filter bank (phaseflow.py:47)


def save_result(out_dir, res: Result, raw=False):
    """io.py:203-223: one folder of .npy maps and timing.json."""
    os.makedirs(out_dir, exist_ok=True)
    arrays = {"ux": res.ux, "uy": res.uy, "conf": res.conf}
    arrays.update({k: res.strain[k] for k in ("Exx", "Eyy", "Exy", "gmax", "vol")})
    if raw:
        arrays.update({"raw_ux": res.raw_ux, "raw_uy": res.raw_uy})
    for name, a in arrays.items():
        np.save(os.path.join(out_dir, f"{name}.npy"), np.ascontiguousarray(a, dtype="<f4"))
    np.save(os.path.join(out_dir, "valid.npy"), np.isfinite(res.ux).astype(np.uint8))
    if res.fb is not None:
        np.save(os.path.join(out_dir, "fb.npy"), np.ascontiguousarray(res.fb, dtype="<i4"))
    with open(os.path.join(out_dir, "timing.json"), "w") as f:
        json.dump(res.timing, f, indent=2)


def sequence_dir(out_dir, k):
    """The folder of sequence frame k (frame_{k:05})."""
    return os.path.join(out_dir, f"frame_{k:05d}")


# ---------------------------------------------------------------- automatic ROI crop
def _pow2_floor(n):
    return 1 << (int(n).bit_length() - 1) if n >= 1 else 0


def level_window(shape, win):
    """poc.py:172-187: the POC window (height, width) at a pyramid level of this shape, or None."""
    h, w = shape
    lim_h, lim_w = _pow2_floor(2 * h // 3), _pow2_floor(2 * w // 3)
    if win <= lim_h and win <= lim_w:
        return win, win
    if h <= w:
        wy = min(win, lim_h)
        wx = min(lim_w, max(win, win * win // max(wy, 1)))
    else:
        wx = min(win, lim_w)
        wy = min(lim_h, max(win, win * win // max(wx, 1)))
    return (wy, wx) if min(wy, wx) >= 16 else None


def n_levels(max_disp, win, shape):
    """poc.py:190-195: pyramid levels so that max_disp / 2^L <= win / 4 while a window fits."""
    L = int(max(0, math.ceil(math.log2(max(max_disp, 1e-9) / (win / 4)))))
    while L > 0 and level_window((shape[0] >> L, shape[1] >> L), win) is None:
        L -= 1
    return L


def default_margin(cfg, shape, seq=False):
    """crop.py:28-39: largest displacement + 4 sigma_w + the guard band + PAD + the coarse stage's
    footprint (poc_win / 2 + grid_step) 2^L; twice that for a sequence."""
    L = n_levels(cfg.max_disp, cfg.poc_win, shape)
    m = (math.ceil(cfg.max_disp) + math.ceil(4 * cfg.sigma_w) + math.ceil(cfg.edge_guard_px()) + PAD
         + (cfg.poc_win // 2 + cfg.grid_step) * 2 ** L)
    return 2 * m if seq else m


@dataclass(frozen=True)
class Crop:
    """A half-open pixel rectangle [x0, x1) x [y0, y1) of the frame (crop.py:42-78)."""
    x0: int
    y0: int
    x1: int
    y1: int
    margin: int

    @property
    def shape(self):
        return self.y1 - self.y0, self.x1 - self.x0

    def apply(self, a):
        """The part of a full-size array inside the crop (a copy)."""
        return a[self.y0:self.y1, self.x0:self.x1].copy()

    def json(self):
        return json.dumps({"x0": self.x0, "y0": self.y0, "x1": self.x1, "y1": self.y1, "margin": self.margin,
                           "mode": "auto"}, indent=2)

    def embed(self, res: Result, shape) -> Result:
        """The crop's result as one of the full frame: NaN outside the crop, confidence, valid and
        fb 0."""
        def put(a, fill):
            full = np.full(shape, fill, dtype=a.dtype)
            full[self.y0:self.y1, self.x0:self.x1] = a
            return full

        return Result(put(res.ux, np.nan), put(res.uy, np.nan), put(res.conf, 0.0),
                      {k: put(v, np.nan) for k, v in res.strain.items()}, put(res.raw_ux, np.nan),
                      put(res.raw_uy, np.nan), put(res.valid, False), res.timing,
                      None if res.fb is None else put(res.fb, 0))

    @staticmethod
    def plan(mask, cfg, margin=None, seq=False):
        """crop.py:81-103: the ROI's bounding box grown by the margin, x0 and y0 rounded down to a
        multiple of grid_step 2^L (the coarse grid of the crop lies on the frame's), clipped."""
        H, W = mask.shape
        cols = np.flatnonzero(mask.any(axis=0))
        if cols.size == 0:
            raise ValueError("ROI mask is empty")
        rows = np.flatnonzero(mask.any(axis=1))
        if margin is not None and margin < 0:
            raise ValueError(f"--crop-margin must be >= 0, got {margin}")
        if min(H, W) < cfg.poc_win:
            raise ValueError(f"image smaller than poc_win ({cfg.poc_win}): {(H, W)}")
        if margin is None:
            margin = default_margin(cfg, mask.shape, seq)
        A = cfg.grid_step * 2 ** n_levels(cfg.max_disp, cfg.poc_win, mask.shape)
        x0 = max(int(cols[0]) - margin, 0) // A * A
        y0 = max(int(rows[0]) - margin, 0) // A * A
        c = Crop(x0, y0, min(int(cols[-1]) + 1 + margin, W), min(int(rows[-1]) + 1 + margin, H), margin)
        h, w = c.shape
        if min(h, w) < cfg.poc_win:
            raise ValueError(f"the crop {w}x{h} is smaller than poc_win ({cfg.poc_win}): "
                             "increase --crop-margin or use --crop off")
        return c
