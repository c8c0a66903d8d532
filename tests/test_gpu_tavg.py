"""Temporal averaging and frame conversion on the GPU (ofx_tavg_window, ofx_frame_to_gray) against
the oracle. Bars: the plain mean and the grey levels are bit-identical (the same f64 operations in
the same order, rounded to f32 once); the aligned means use the f32 spline of the GPU warp against
the oracle's f64 spline, within 2e-5 of the [0, 1] grey scale (interior, away from the 4 px the
spline's mirrored border reaches)."""
import numpy as np
import pytest

import oracle as O
from optflow_soil_amd import synth
from optflow_soil_amd.config import Config
from optflow_soil_amd.frames import averaged, gpu_average

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import optflow_soil_amd as ofx
    e = ofx.Engine(0)
    yield e
    e.close()


def moving(n, count, v, seed=11):
    base = synth.as_float(synth.render(synth.make_pattern(n, n, seed=seed))).astype(np.float64)
    return [O.warp(base, np.full((n, n), v * i), np.full((n, n), -0.5 * v * i)).astype(np.float32)
            for i in range(count)]


@pytest.mark.parametrize("bits,channels,factor", [(8, 1, 1), (8, 3, 2), (8, 4, 3), (16, 1, 2), (16, 3, 1)])
def test_frame_to_gray_is_bit_identical(eng, bits, channels, factor):
    rng = np.random.default_rng(bits * 10 + channels)
    shape = (67, 93) if channels == 1 else (67, 93, channels)
    px = rng.integers(0, 2 ** bits, size=shape, dtype=np.uint8 if bits == 8 else np.uint16)
    g = eng.frame_to_gray(px, factor)
    o = O.to_gray(px, factor).astype(np.float32)
    print(f"gray {bits} bit x{channels} factor {factor}: shape {g.shape}", flush=True)
    assert g.shape == o.shape == (67 // factor, 93 // factor)
    assert np.array_equal(g, o)


def test_plain_window_mean_is_bit_identical(eng):
    rng = np.random.default_rng(1)
    fr = [rng.random((50, 70)).astype(np.float32) for _ in range(7)]
    for centre in (0, 3, 6):
        g, used = eng.tavg_window(fr, centre)
        o, oused = O.average_window(fr, centre)
        assert used == oused == 7
        assert np.array_equal(g, o.astype(np.float32))


@pytest.mark.parametrize("model", ["affine", "dense"])
def test_aligned_window_mean(eng, model):
    fr = moving(128, 5, 0.7)
    cfg = Config(max_disp=8.0)
    mask = np.zeros((128, 128), bool)
    mask[16:112, 24:120] = True
    for centre, msk in ((2, None), (0, mask)):
        g, used = eng.tavg_window(fr, centre, model, cfg, msk if model == "affine" else None)
        o, oused = O.average_window(fr, centre, model, cfg, msk if model == "affine" else None)
        d = np.abs(g.astype(np.float64) - o)[4:-4, 4:-4]
        print(f"tavg {model} centre {centre}: used {used}/{oused} max {d.max():.2e} rms {np.sqrt((d ** 2).mean()):.2e}",
              flush=True)
        assert used == oused == 5
        assert d.max() <= 2e-5


@pytest.mark.parametrize("model", ["affine", "dense"])
def test_alignment_failure_leaves_the_frame_out_or_unaligned(eng, model):
    """48 x 48 frames are smaller than the POC window: no motion estimate, so each neighbour is
    left out (the mean is the centre) or, with keep_unaligned, enters as it is."""
    fr = moving(48, 3, 0.5)
    g, used = eng.tavg_window(fr, 1, model, Config())
    assert used == 1 and np.array_equal(g, fr[1])
    g, used = eng.tavg_window(fr, 1, model, Config(), keep_unaligned=True)
    o, oused = O.average_window(fr, 1, model, Config(), keep_unaligned=True)
    assert used == oused == 3 and np.array_equal(g, o.astype(np.float32))


def test_averaged_sequence_front_end(eng):
    """The windowing with the GPU mean against the same windowing with the oracle's: edge windows
    aligned (dense), inner windows plain."""
    fr = moving(128, 9, 0.6)
    cfg = Config(max_disp=8.0)

    def oracle_average(frames, centre, model, c, mask, keep):
        return O.average_window(frames, centre, model, c, mask, keep)[0]

    got = list(averaged(iter(fr), 0, None, 1, 2, cfg=cfg, average=gpu_average(eng)))
    want = list(averaged(iter(fr), 0, None, 1, 2, cfg=cfg, average=oracle_average))
    assert len(got) == len(want) == 9
    for k, (a, b) in enumerate(zip(got, want)):
        d = np.abs(a.astype(np.float64) - b)[4:-4, 4:-4].max()
        print(f"averaged frame {k}: max {d:.2e}", flush=True)
        assert d <= 2e-5
