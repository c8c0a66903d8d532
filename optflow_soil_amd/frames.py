"""Frame input and temporal averaging in front of a sequence.

- to_gray / read_image / list_images: the reference's image input (python/optflow_ref/io.py:15-97,
  rust/crates/optflow-cli/src/io.rs:44-52, :259-281). Decoding stays on the host (Pillow); the
  conversion to grey levels and the box downscale run on the GPU (ofx_frame_to_gray).
- averaged: temporal averaging with the windows of rust/crates/optflow-cli/src/tavg.rs:210-352:
  every sampled frame is the mean of the 2 m + 1 frames around it, clipped to the input; a window
  that is one-sided because the input ends has its neighbours aligned by the dense coarse model
  first. A bounded queue holds the frames later windows still need. The mean itself is
  ofx_tavg_window (or any function with its signature, for tests).
- iter_frames: io.py:100-121 on a directory or glob of images. Files are decoded ahead on a worker
  thread (decoded_ahead) while the caller tracks the current frame.
"""
import collections
import concurrent.futures
import dataclasses
import glob
import itertools
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


def decode_image(path):
    """io.py:33-41: the pixels of one image file as an array (Pillow decodes; 16-bit modes stay
    16-bit). Host work only."""
    if path.lower().endswith((".tif", ".tiff")):
        try:
            import tifffile
        except ImportError:
            raise ValueError(f"reading {path} needs tifffile, which is not installed") from None
        return tifffile.imread(path)
    from PIL import Image
    with Image.open(path) as im:
        if im.mode in ("I;16", "I;16B", "I;16L", "I"):
            return np.asarray(im).astype(np.uint16)
        if im.mode not in ("L", "RGB", "RGBA"):
            im = im.convert("RGB")
        return np.asarray(im)


def read_image(path, factor=1, engine: Engine = None):
    """io.py:33-41: one image file as grey levels (decoded on the host, converted on the GPU)."""
    return to_gray(decode_image(path), factor, engine)


def decoded_ahead(paths, decode=decode_image, depth=2):
    """decode(p) for each path in order, with up to `depth` files decoded ahead on one worker
    thread, so that reading and decoding the next frames overlaps whatever the caller does with the
    current one (tracking it on the GPU). A decode error is raised when its frame is reached; the
    frames before it are yielded first."""
    it = iter(paths)
    with concurrent.futures.ThreadPoolExecutor(max_workers=1, thread_name_prefix="ofx-decode") as pool:
        pending = collections.deque(pool.submit(decode, p) for p in itertools.islice(it, depth))
        try:
            while pending:
                current = pending.popleft()
                for p in itertools.islice(it, 1):
                    pending.append(pool.submit(decode, p))
                yield current.result()
        finally:
            for f in pending:
                f.cancel()


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
    """Temporal averaging of a frame stream (the behaviour of tavg.rs:210-352): raw yields the
    frames with index max(0, start - m), max(0, start - m) + 1, ...; the result is the mean of
    the frames within m of each sampled index s = start, start + step, ... (s < stop when stop
    is given), clipped to the frames that exist. A mean is produced as soon as frame s + m has
    arrived, or at the end of the stream with the window cut at the last frame, and only the
    frames some later window still needs are held.

    align: every window aligned before averaging (cfg.tavg_model, fit restricted to
    mask_fn(shape)); otherwise a window made one-sided by the ends of the input is aligned with
    the dense coarse model (edge_align), a neighbour whose alignment fails entering unaligned.
    average(frames, centre, model, cfg, mask, keep_unaligned) forms one mean (ofx_tavg_window on
    the GPU by default)."""
    if step < 1:
        raise ValueError(f"step must be >= 1, got {step}")
    if start < 0:
        raise ValueError(f"start must be >= 0, got {start}")
    if m < 0:
        raise ValueError(f"tavg must be >= 0, got {m}")
    mean = average or gpu_average()
    cfg = cfg or Config()
    dense = dataclasses.replace(cfg, tavg_model="dense") if edge_align else None
    held = collections.deque()  # (index, frame), consecutive indices
    wanted = (lambda s: True) if stop is None else (lambda s: s < stop)

    def window_mean(s, newest):
        lo, hi = max(0, s - m), min(s + m, newest)
        frames = [f for i, f in held if lo <= i <= hi]
        centre = s - lo
        if len(frames) == 1:
            return frames[0]
        if align:
            return mean(frames, centre, cfg.tavg_model, cfg, mask_fn(frames[0].shape) if mask_fn else None, False)
        # a window cut by the ends of the input is centred on s only after alignment; an image
        # smaller than the POC window has no grid to estimate the motion on
        if dense is not None and hi - s != centre and min(frames[0].shape) >= dense.poc_win:
            return mean(frames, centre, "dense", dense, None, True)
        return mean(frames, centre, None, cfg, None, False)

    target = start
    newest = -1
    for newest, frame in enumerate(raw, start=max(0, start - m)):
        held.append((newest, frame))
        while wanted(target) and target + m <= newest:
            yield window_mean(target, newest)
            target += step
            while held and held[0][0] < target - m:
                held.popleft()
    while wanted(target) and target <= newest:
        yield window_mean(target, newest)
        target += step


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
        for px in decoded_ahead(paths[start:stop:step]):
            yield to_gray(px, scale, engine)
        return
    lo = max(0, start - tavg)
    hi = None if stop is None else stop + tavg
    raw = (to_gray(px, scale, engine) for px in decoded_ahead(paths[lo:hi]))
    yield from averaged(raw, start, stop, step, tavg, align, cfg, mask_fn, True, gpu_average(engine))
