"""Temporal averaging and frame conversion on the CPU: the oracle (oracle/tavg.cpp) against closed
forms and the reference's own unit-test cases (rust/crates/optflow-cli/src/tavg.rs:356-553,
rust/crates/optflow-cli/src/io.rs:675-690), and the windowing of optflow_soil_amd.frames.averaged
driven by the oracle's mean (no GPU)."""
import numpy as np
import pytest

import oracle as O
from optflow_soil_amd import synth
from optflow_soil_amd.config import Config
from optflow_soil_amd.frames import averaged


def oracle_average(frames, centre, model, cfg, mask, keep_unaligned):
    return O.average_window(frames, centre, model, cfg, mask, keep_unaligned)[0]


def test_affine_fit_recovers_a_plane_and_drops_an_outlier():
    xs, ys = np.meshgrid(16.0 * np.arange(8) + 20.0, 16.0 * np.arange(6) + 30.0)
    xs, ys = xs.ravel(), ys.ravel()
    ux = 0.5 + 0.001 * xs - 0.002 * ys
    uy = -1.0 + 0.003 * ys
    ux[10] += 5.0
    cx, cy = O.fit_affine(xs, ys, ux, uy)
    assert np.allclose(cx, [0.5, 0.001, -0.002], atol=1e-9, rtol=0)
    assert np.allclose(cy, [-1.0, 0.0, 0.003], atol=1e-9, rtol=0)
    # a wrong sign or a swapped coordinate would show here: the fit reproduces every inlier
    keep = np.arange(len(xs)) != 10
    assert np.abs(cx[0] + cx[1] * xs + cx[2] * ys - ux)[keep].max() < 1e-9
    assert O.fit_affine(xs[:4], ys[:4], ux[:4], uy[:4]) is None  # fewer than 6 points
    # collinear points leave the system singular: no fit
    assert O.fit_affine(np.arange(8.0), np.zeros(8), np.ones(8), np.ones(8)) is None


def test_to_gray_is_the_box_mean_of_the_luminance():
    a = np.arange(16, dtype=np.uint8).reshape(4, 4)
    assert np.allclose(O.to_gray(a, 2) * 255, [[2.5, 4.5], [10.5, 12.5]], rtol=1e-15, atol=0)
    b = np.add.outer(np.arange(3), np.arange(5)).astype(np.uint8)  # 3 x 5, remainder cropped
    d = O.to_gray(b, 2)
    assert d.shape == (1, 2) and np.allclose(d * 255, [[1.0, 3.0]], rtol=1e-15, atol=0)
    rgb = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [10, 20, 30]]], np.uint8)
    want = [0.299, 0.587, 0.114, 0.299 * (10 / 255) + 0.587 * (20 / 255) + 0.114 * (30 / 255)]
    assert np.array_equal(O.to_gray(rgb)[0], want)  # left to right, as numpy evaluates it
    rgba = np.concatenate([rgb, np.full((1, 4, 1), 7, np.uint8)], axis=2)
    assert np.array_equal(O.to_gray(rgba), O.to_gray(rgb))  # alpha ignored
    w16 = np.array([[0, 65535, 32768]], np.uint16)
    assert np.array_equal(O.to_gray(w16)[0], [0.0, 1.0, 32768 / 65535])


def test_plain_mean_is_the_f64_sum_centre_first():
    rng = np.random.default_rng(0)
    fr = [rng.random((9, 11)).astype(np.float32) for _ in range(5)]
    mean, used = O.average_window(fr, 2)
    s = fr[2].astype(np.float64)
    for i in (0, 1, 3, 4):
        s = s + fr[i]
    assert used == 5 and np.array_equal(mean, (s / 5).astype(np.float32).astype(np.float64))
    # the mean of noisy frames has less noise (tavg.rs:400-418)
    assert np.sqrt(np.mean((mean - 0.5) ** 2)) < np.sqrt(np.mean((fr[2] - 0.5) ** 2)) / 1.8


def indexed(n, first=0):
    return [np.full((2, 2), i, np.float32) for i in range(first, first + n)]


def values(gen):
    return [float(f[0, 0]) for f in gen]


def test_windows_are_centred_and_clipped():
    """The window cases of tavg.rs:374-398 (frames whose value is their index)."""
    a = averaged(iter(indexed(10)), 1, None, 3, 2, average=oracle_average)
    assert values(a) == [1.5, 4.0, 7.0]
    a = averaged(iter(indexed(10)), 0, None, 4, 1, average=oracle_average)
    assert values(a) == [0.5, 4.0, 8.0]
    a = averaged(iter(indexed(9)), 2, 6, 2, 3, average=oracle_average)
    assert values(a) == [2.5, 4.0]
    # a reader that begins later numbers its frames from start - m
    a = averaged(iter(indexed(8, 3)), 6, None, 1, 3, average=oracle_average)
    assert values(a) == [6.0, 7.0, 7.5, 8.0, 8.5]
    a = averaged(iter(indexed(3)), 0, None, 1, 0, average=oracle_average)
    assert values(a) == [0.0, 1.0, 2.0]


def texture(n, seed=11):
    return synth.as_float(synth.render(synth.make_pattern(n, n, seed=seed))).astype(np.float64)


def moving(n, count, v, seed=11):
    base = texture(n, seed)
    return [O.warp(base, np.full((n, n), v * i), np.zeros((n, n))).astype(np.float32) for i in range(count)]


def rms_inside(a, b, e=20):
    return float(np.sqrt(np.mean((a[e:-e, e:-e].astype(np.float64) - b[e:-e, e:-e]) ** 2)))


def test_one_sided_windows_are_aligned_and_inner_ones_are_plain():
    """tavg.rs:456-491: frames moving 0.6 px per frame; the plain mean of a one-sided window lags,
    the dense-aligned one does not; symmetric windows are the plain mean bit for bit; with align
    every window is aligned, with or without the edge rule."""
    n, m = 128, 3
    fr = moving(n, 12, 0.6)
    cfg = Config(max_disp=8.0)
    plain = list(averaged(iter(fr), 0, None, 1, m, edge_align=False, cfg=cfg, average=oracle_average))
    edge = list(averaged(iter(fr), 0, None, 1, m, edge_align=True, cfg=cfg, average=oracle_average))
    assert len(plain) == 12
    for s in (0, 1, 2, 9, 10, 11):
        p, e = rms_inside(plain[s], fr[s]), rms_inside(edge[s], fr[s])
        assert p > 0.004, (s, p)
        assert e < 0.2 * p, (s, e, p)
    for s in range(m, 12 - m):
        assert np.array_equal(plain[s], edge[s])
    a1 = list(averaged(iter(fr), 0, None, 1, m, align=True, edge_align=False, cfg=cfg, average=oracle_average))
    a2 = list(averaged(iter(fr), 0, None, 1, m, align=True, edge_align=True, cfg=cfg, average=oracle_average))
    for s in range(12):
        assert np.array_equal(a1[s], a2[s])
    assert not np.array_equal(a1[5], plain[5])
    assert rms_inside(a1[5], fr[5]) < 0.2 * rms_inside(plain[5], fr[5]) + 1e-3


def test_integer_shift_alignment_reproduces_the_centre():
    """A frame moved by whole pixels is aligned exactly (the POC peak is integral, the affine fit
    is the translation, and the spline reproduces its samples) away from the border the move
    uncovers."""
    n = 128
    base = texture(n, seed=5)
    other = O.warp(base, np.full((n, n), 3.0), np.full((n, n), -2.0))
    a = O.align_to(other, base, Config(), model="affine")
    assert a is not None
    assert np.abs(a[8:-8, 8:-8] - other[8:-8, 8:-8].astype(np.float32)).max() < 1e-6
    # 96 x 96: 5 of the 9 POC windows leave the image (peak 0), too few points for a fit
    assert O.align_to(other[:96, :96], base[:96, :96], Config(), model="affine") is None


def test_small_frames_are_averaged_unaligned():
    """tavg.rs:493-503: frames smaller than the POC window have no grid; the ends stay plain."""
    fr = moving(16, 8, 0.3)
    plain = list(averaged(iter(fr), 0, None, 1, 2, edge_align=False, average=oracle_average))
    edge = list(averaged(iter(fr), 0, None, 1, 2, edge_align=True, average=oracle_average))
    for a, b in zip(plain, edge):
        assert np.array_equal(a, b)


def test_averaged_rejects_bad_ranges():
    with pytest.raises(ValueError, match="step must be >= 1"):
        list(averaged(iter(indexed(3)), 0, None, 0, 1, average=oracle_average))
    with pytest.raises(ValueError, match="tavg must be >= 0"):
        list(averaged(iter(indexed(3)), 0, None, 1, -1, average=oracle_average))
