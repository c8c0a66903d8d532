"""Frame input ahead of the tracker (optflow_soil_amd.frames.decoded_ahead): files are decoded on a
worker thread while the caller works on the current frame. Host logic only, no GPU: order, overlap,
errors at the frame that caused them, and an early stop that leaves no thread behind."""
import threading
import time

import numpy as np
import pytest

from optflow_soil_amd.frames import decode_image, decoded_ahead


def test_frames_come_in_order_and_from_another_thread():
    seen = []

    def decode(p):
        seen.append(threading.current_thread().name)
        return p * 10

    assert list(decoded_ahead(range(7), decode)) == [0, 10, 20, 30, 40, 50, 60]
    assert len(seen) == 7 and all(n.startswith("ofx-decode") for n in seen)
    assert list(decoded_ahead([], decode)) == []


def test_decoding_overlaps_the_callers_work():
    def decode(p):
        time.sleep(0.05)
        return p

    t0 = time.perf_counter()
    for _ in decoded_ahead(range(8), decode):
        time.sleep(0.05)  # the caller's work on the frame
    # serial: 8 x (0.05 + 0.05) = 0.8 s; overlapped: about 0.05 + 8 x 0.05 = 0.45 s
    assert time.perf_counter() - t0 < 0.7


def test_an_error_is_raised_at_its_frame_after_the_frames_before_it():
    def decode(p):
        if p == 3:
            raise ValueError("frame 3 is broken")
        return p

    got = []
    with pytest.raises(ValueError, match="frame 3 is broken"):
        for v in decoded_ahead(range(6), decode):
            got.append(v)
    assert got == [0, 1, 2]


def test_stopping_early_decodes_at_most_depth_more_and_returns():
    calls = []

    def decode(p):
        calls.append(p)
        return p

    gen = decoded_ahead(range(100), decode, depth=2)
    assert [next(gen) for _ in range(3)] == [0, 1, 2]
    gen.close()
    assert len(calls) <= 5
    assert not [t for t in threading.enumerate() if t.name.startswith("ofx-decode")]


def test_decode_image_keeps_16_bit_pixels(tmp_path):
    from PIL import Image
    a = (np.arange(48 * 32, dtype=np.uint32).reshape(32, 48) * 37 % 65536).astype(np.uint16)
    Image.fromarray(a).save(tmp_path / "f.png")
    b = decode_image(str(tmp_path / "f.png"))
    assert b.dtype == np.uint16 and np.array_equal(a, b)
    rgb = np.random.default_rng(1).integers(0, 256, (20, 30, 3), dtype=np.uint8)
    Image.fromarray(rgb).save(tmp_path / "c.png")
    assert np.array_equal(decode_image(str(tmp_path / "c.png")), rgb)
