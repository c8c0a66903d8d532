"""The byte models bench.py reports its roofline figures with: the accumulate kernel's algorithmic
bytes per launch and the whole pair's, against the figures the design states for the 2048^2 pair
(1.60 GB per accumulate launch; 67.9 GB per pair untiled, 79.1 GB in tiles of 1024), computed
independently here from their terms."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_accumulate_bytes_at_2048():
    b = bench()
    n, np_, F = 2048 * 2048, 2112 * 2112, 12
    # per pixel: u (2 f64), s and its spline (4 f32), the reference pack (4 f32 per filter) and the
    # seven f64 sums; per padded node: one c32 sample plane per filter
    want = n * (2 * 8 + 4 * 4 + 16 * F + 7 * 8) + F * 8 * np_
    assert b.accumulate_bytes(2048, 2048, F) == want == 1602617344


def test_pipeline_bytes_untiled_and_tiled():
    b = bench()
    assert b.pipeline_bytes(2048, 2048) == 6388 * 2048 ** 2 + 9120 * 2112 ** 2 + 100 * 2048 ** 2
    assert abs(b.pipeline_bytes(2048, 2048) / 1e9 - 67.89) < 0.01
    # four edge tiles of 1088 x 1088 (core 1024 plus a 64 px margin on the inner sides)
    tiled = 4 * (6388 * 1088 ** 2 + 9120 * 1152 ** 2) + 100 * 2048 ** 2
    assert b.pipeline_bytes(2048, 2048, 1024) == tiled
    # a frame no side of which exceeds the tile is not tiled
    assert b.pipeline_bytes(1000, 800, 1024) == b.pipeline_bytes(1000, 800)
