"""The sequence and batched workloads at their full sizes, by properties that do not need the oracle
at that size, and a reduced sequence that exercises what the full one does not (a reference that
moves, a fallback that fires) against the oracle frame by frame.

- The 200-frame 2048 x 2048 hybrid sequence (tavg 3, prewarp_freeze_at 3, ref_update_conf 0.3):
  streamed once with the fine stage replayed from graphs and once eagerly, bit for bit the same on
  every frame (a digest of every output plane), the cumulative displacement at frame 199 within
  0.1 px RMS of the truth, and the valid fraction reported.
- A 384 x 384 hybrid sequence whose field grows and in which a foreign block appears at frame 6:
  the reference moves (ref_update_conf 0.9) and the adjacent leg takes pixels over; the GPU
  against the oracle on every frame (references, mean confidence, fb maps, displacement).
- The 64 distinct uneven pairs through the batched entry point: each bit for bit its single-pair
  result (a sample of them), the pairs distinct."""
import hashlib

import numpy as np
import pytest

from optflow_soil_amd.config import Config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import optflow_soil_amd as ofx
    e = ofx.Engine(0, lanes=2)
    yield e
    e.close()


def digest(r):
    h = hashlib.sha256()
    for a in (r.ux, r.uy, r.conf, r.raw_ux, r.raw_uy, r.valid, *r.strain.values()):
        h.update(np.ascontiguousarray(a).tobytes())
    if r.fb is not None:
        h.update(r.fb.tobytes())
    return h.hexdigest()


@pytest.mark.timeout(600)
def test_the_200_frame_hybrid_sequence(eng, monkeypatch):
    from optflow_soil_amd import synth
    from optflow_soil_amd.frames import averaged, gpu_average
    n, nf = 2048, 200
    cfg = Config(seq_mode="hybrid", prewarp_freeze_at=3, ref_update_conf=0.3)
    print(f"rendering {nf} frames of {n}^2 on the GPU", flush=True)
    raw = list(synth.growing_bowl_and_band_frames_gpu(n, nf))
    print("averaging (tavg 3)", flush=True)
    frames = list(averaged(iter(raw), 0, None, 1, 3, cfg=cfg, average=gpu_average(eng)))
    del raw

    def run():
        out, last = [], None
        for k, r in eng.iter_sequence(iter(frames), cfg):
            out.append((digest(r), r.timing["reference"], r.timing["fallback_frac"]))
            last = r
            if k % 50 == 0:
                print(f"  frame {k}", flush=True)
        return out, last
    print("streaming with graphs", flush=True)
    a, last = run()
    monkeypatch.setenv("OFX_GRAPHS", "0")
    print("streaming eagerly", flush=True)
    b, _ = run()
    monkeypatch.delenv("OFX_GRAPHS")
    assert [x[0] for x in a] == [x[0] for x in b]
    gx, gy = synth.bowl_and_band(n, n)(*np.mgrid[0:n, 0:n][::-1].astype(np.float64))
    fin = np.isfinite(last.ux)
    drift = float(np.sqrt(np.mean((last.ux[fin] - gx[fin]) ** 2 + (last.uy[fin] - gy[fin]) ** 2)))
    refs = sorted({x[1] for x in a})
    print(f"frame {nf - 1}: drift {drift:.4f} px, valid {fin.mean():.4f}, references {refs}, "
          f"fallback at the last frame {a[-1][2]:.4f}", flush=True)
    assert drift < 0.1 and fin.mean() > 0.9


def block_sequence():
    """384 x 384, 12 frames: the bowl-and-band field scaled up to 8 px depth and 6 px slip grows
    linearly; from frame 6 a foreign 100 x 110 px block replaces part of the scene; brightness drift
    -3 % .. +3 %, noise 0.01, 8-bit."""
    from optflow_soil_amd import synth
    h = w = 384
    n = 12
    pat = synth.make_pattern(h, w, seed=5)
    other = synth.make_pattern(h, w, seed=99)
    full = synth.bowl_and_band(h, w, depth=8.0, radius=120.0, slip=6.0, band_width=20.0)
    frames = []
    for k in range(n):
        s = k / (n - 1)
        f = (lambda s: (lambda X, Y: tuple(s * c for c in full(X, Y))))(s)
        v = synth.render(pat, f if k else None, noise=0.0, quantise=False)
        if k >= n // 2:
            v[40:140, 220:330] = other[64 + 40:64 + 140, 64 + 220:64 + 330]
        v = v * (0.97 + 0.06 * s) + np.random.default_rng(100 + k).normal(0.0, 0.01, v.shape)
        frames.append(np.clip(np.rint(v * 255.0), 0, 255).astype(np.uint8) / np.float32(255.0))
    return frames


@pytest.mark.timeout(400)
def test_hybrid_sequence_where_the_reference_moves_and_the_fallback_fires(eng):
    import oracle as O
    frames = block_sequence()
    cfg = Config(seq_mode="hybrid", prewarp_freeze_at=3, ref_update_conf=0.9)
    print("block sequence 384^2 x 12: GPU", flush=True)
    got = eng.track_sequence(frames, cfg)
    print("block sequence: oracle", flush=True)
    ref = O.track_sequence([f.astype(np.float64) for f in frames], cfg)
    refs, fbs = set(), []
    for (k, r), o in zip(got, ref):
        t, ot = r.timing, o["timing"]
        both = np.isfinite(r.ux) & np.isfinite(o["ux"])
        agree = float((np.isfinite(r.ux) == np.isfinite(o["ux"])).mean())
        d = float(np.sqrt(np.mean((r.ux[both] - o["ux"][both]) ** 2 + (r.uy[both] - o["uy"][both]) ** 2)))
        fb_agree = float((r.fb == o["fb"]).mean())
        print(f"  frame {k}: reference {t['reference']} / {ot['reference']}, mean conf {t['mean_conf']:.4f} / "
              f"{ot['mean_conf']:.4f}, fallback {t['fallback_frac']:.4f} / {ot['fallback_frac']:.4f}, valid agreement "
              f"{agree:.5f}, fb agreement {fb_agree:.5f}, displacement RMS {d:.2e}", flush=True)
        assert t["reference"] == ot["reference"]
        assert abs(t["mean_conf"] - ot["mean_conf"]) < 1e-3
        assert agree >= 0.999 and fb_agree >= 0.999 and d <= 1e-3
        refs.add(t["reference"])
        fbs.append(t["fallback_frac"])
    assert len(refs) >= 2 and max(fbs) > 0


@pytest.mark.timeout(600)
def test_64_distinct_uneven_pairs_batched(eng):
    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth
    print("rendering the 64 distinct pairs", flush=True)
    raw = synth.workload_c4()
    assert len({a.tobytes() for a, _, _ in raw}) == 64
    pairs = [(synth.as_float(a), synth.as_float(b), roi) for a, b, roi in raw]
    e4 = ofx.Engine(0, lanes=4)
    try:
        print("64 pairs on 4 lanes", flush=True)
        out = e4.track_pairs(pairs, Config())
    finally:
        e4.close()
    valid = [float(np.isfinite(r.ux)[p[2]].mean()) for r, p in zip(out, pairs)]
    print(f"valid fraction over the ROIs: min {min(valid):.4f}, mean {np.mean(valid):.4f}", flush=True)
    assert np.mean(valid) > 0.85 and min(valid) > 0.5
    for k in (0, 13, 27, 41, 63):
        one = eng.track_pair(pairs[k][0], pairs[k][1], Config(), mask=pairs[k][2])
        assert digest(one) == digest(out[k]), k
