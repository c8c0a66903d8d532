"""Frame input and temporal averaging in front of a sequence.

- to_gray / read_image / list_images: the reference's image input (python/optflow_ref/io.py:15-97,
  rust/crates/optflow-cli/src/io.rs:44-52, :259-281). Decoding stays on the host (Pillow); the
  conversion to grey levels and the box downscale run on the GPU (ofx_frame_to_gray).
- averaged: the windowing of rust/crates/optflow-cli/src/tavg.rs:210-352 (io.py:100-162): every
  sampled frame is the mean of the 2 m + 1 frames around it, clipped to the input; a window that is
  one-sided because the input ends has its neighbours aligned by the dense coarse model first.
  The mean itself is ofx_tavg_window (or any function with its signature, for tests).
- iter_frames: io.py:100-121 on a directory or glob of images.
"""
import dataclasses
import glob
import os

import numpy as np

from . import Engine, default_engine
from .config import Config

VIDEO_EXT = {".mp4", ".mov", ".avi", ".mkv", ".m4v"}
IMAGE_EXT = (".png", ".jpg", ".jpeg", ".tif", ".tiff", ".bmp")


def to_gray(pixels, factor=1, engine: Engine = None):
    """io.py:15-30 then io.py:44-52: uint8 / uint16 pixels (H, W) or (H, W, 3|4) -> float32 grey
    levels in [0, 1], box downscaled by the integer factor (remainder cropped)."""
    return (engine or default_engine()).frame_to_gray(pixels, factor)


def read_image(path, factor=1, engine: Engine = None):
    """io.py:33-41: one image file as grey levels (Pillow decodes; 16-bit modes stay 16-bit)."""
    if path.lower().endswith((".tif", ".tiff")):
        try:
            import tifffile
        except ImportError:
            raise ValueError(f"reading {path} needs tifffile, which is not installed") from None
        return to_gray(tifffile.imread(path), factor, engine)
    from PIL import Image
    with Image.open(path) as im:
        if im.mode in ("I;16", "I;16B", "I;16L", "I"):
            return to_gray(np.asarray(im).astype(np.uint16), factor, engine)
        if im.mode not in ("L", "RGB", "RGBA"):
            im = im.convert("RGB")
        return to_gray(np.asarray(im), factor, engine)


def list_images(source):
    """io.py:82-92: the image files of a directory or glob pattern in plain string order."""
    if os.path.isdir(source):
        paths = sorted(p for p in glob.glob(os.path.join(source, "*")) if p.lower().endswith(IMAGE_EXT))
    else:
        paths = sorted(glob.glob(source))
    if not paths:
        raise ValueError(f"no images found for {source}")
    return paths


def gpu_average(engine: Engine = None):
    """The mean of a window on the GPU, with the signature `averaged` expects."""
    def average(frames, centre, model, cfg, mask, keep_unaligned):
        return (engine or default_engine()).tavg_window(frames, centre, model, cfg, mask, keep_unaligned)[0]
    return average


def averaged(raw, start=0, stop=None, step=1, m=0, align=False, cfg=None, mask_fn=None, edge_align=True,
             average=None):
    """tavg.rs:210-352 / io.py:124-162. raw yields the frames with index max(0, start - m), ...;
    yields the means for the sampled indices start, start + step, ... < stop (the windows may reach
    past stop and are clipped at the last frame). align: every window aligned (affine or dense as
    cfg.tavg_model says, fit restricted to mask_fn(shape)); else the one-sided windows at the ends
    are aligned with the dense model (edge_align), a neighbour whose alignment fails entering
    unaligned."""
    if step < 1:
        raise ValueError(f"step must be >= 1, got {step}")
    if start < 0:
        raise ValueError(f"start must be >= 0, got {start}")
    if m < 0:
        raise ValueError(f"tavg must be >= 0, got {m}")
    average = average or gpu_average()
    cfg = cfg or Config()
    edge_cfg = dataclasses.replace(cfg, tavg_model="dense") if edge_align else None
    buf = {}
    nxt = start
    last = None

    def emit(s):
        lo, hi = max(0, s - m), min(s + m, last)
        fr = [buf[i] for i in range(lo, hi + 1)]
        if len(fr) == 1:
            return fr[0]
        one_sided = s - lo != hi - s
        # an image smaller than the POC window has no grid to estimate the motion on
        edge = (one_sided and not align and edge_cfg is not None and min(fr[0].shape) >= edge_cfg.poc_win)
        if align:
            mask = mask_fn(fr[0].shape) if mask_fn else None
            return average(fr, s - lo, cfg.tavg_model, cfg, mask, False)
        if edge:
            return average(fr, s - lo, "dense", edge_cfg, None, True)
        return average(fr, s - lo, None, cfg, None, False)

    def due(ready):
        return (stop is None or nxt < stop) and ready is not None and nxt <= ready

    read = max(0, start - m)
    for f in raw:
        buf[read], last = f, read
        read += 1
        while due(last - m if last - m >= 0 else None):
            yield emit(nxt)
            nxt += step
            for j in [j for j in buf if j < nxt - m]:
                del buf[j]
    while due(last):
        yield emit(nxt)
        nxt += step


def iter_frames(source, start=0, stop=None, step=1, scale=1, tavg=0, align=False, cfg=None, mask_fn=None,
                engine: Engine = None):
    """io.py:100-121 for a directory or glob of images: frames start, start + step, ... < stop,
    downscaled by scale, and with tavg = m > 0 averaged over 2 m + 1 frames (averaged)."""
    if os.path.splitext(source)[1].lower() in VIDEO_EXT:
        raise ValueError(f"{source}: video input needs a decoder (OpenCV), which is not installed")
    if step < 1:
        raise ValueError(f"step must be >= 1, got {step}")
    if start < 0:
        raise ValueError(f"start must be >= 0, got {start}")
    paths = list_images(source)
    if not tavg:
        for p in paths[start:stop:step]:
            yield read_image(p, scale, engine)
        return
    lo = max(0, start - tavg)
    hi = None if stop is None else stop + tavg
    raw = (read_image(p, scale, engine) for p in paths[lo:hi])
    yield from averaged(raw, start, stop, step, tavg, align, cfg, mask_fn, True, gpu_average(engine))
