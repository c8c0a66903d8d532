"""The fine stage replayed from a HIP graph (engine.hip refine_graph) against the same calls run
eagerly (OFX_GRAPHS=0): bit-identical on every replay, with other eager work and synchronisation
between the replays, for whole pairs, the FineEngine refine call and a hybrid sequence (whose
reference pack is cached between steps)."""
import numpy as np
import pytest

from optflow_soil_amd.config import Config

pytestmark = pytest.mark.gpu

PLANES = ("ux", "uy", "conf", "raw_ux", "raw_uy")


@pytest.fixture(scope="module")
def eng():
    import optflow_soil_amd as ofx
    e = ofx.Engine(0)
    yield e
    e.close()


def pair(h, w, seed, shift=(2.5, -1.5)):
    from optflow_soil_amd import synth
    pat = synth.make_pattern(h, w, seed=seed)
    f = synth.stretch_translate(0.01, shift, (w / 2, h / 2))
    return synth.as_float(synth.render(pat)), synth.as_float(synth.render(pat, f, noise=0.005, seed=seed + 1))


def same(r, q, label):
    for k in PLANES:
        assert np.array_equal(getattr(r, k), getattr(q, k), equal_nan=True), f"{label}: {k}"
    for k in r.strain:
        assert np.array_equal(r.strain[k], q.strain[k], equal_nan=True), f"{label}: {k}"


def test_pairs_replayed_from_a_graph_are_bit_identical(eng, monkeypatch):
    a, b = pair(192, 256, 41)
    c, d = pair(160, 224, 43)
    monkeypatch.setenv("OFX_GRAPHS", "0")
    want = eng.track_pair(a, b, Config())
    other = eng.track_pair(c, d, Config())
    monkeypatch.delenv("OFX_GRAPHS")
    cap0, run0 = eng.graph_stats()
    for k in range(5):
        print(f"graph replay {k}", flush=True)
        same(eng.track_pair(a, b, Config()), want, f"call {k}")
        same(eng.track_pair(c, d, Config()), other, f"other size, call {k}")  # eager work between replays
    cap, run = eng.graph_stats()
    print(f"captures {cap - cap0}, replays {run - run0}", flush=True)
    assert cap - cap0 >= 2 and run - run0 >= 6


def test_refine_raw_replayed_from_a_graph_is_bit_identical(eng, monkeypatch):
    a, b = pair(160, 192, 45)
    ux, uy = np.full(a.shape, 2.0, np.float32), np.full(a.shape, -1.0, np.float32)
    monkeypatch.setenv("OFX_GRAPHS", "0")
    want = eng.refine_raw(a, b, ux, uy, Config())
    monkeypatch.delenv("OFX_GRAPHS")
    for k in range(4):
        got = eng.refine_raw(a, b, ux, uy, Config())
        for g, w_ in zip(got, want):
            assert np.array_equal(g, w_, equal_nan=True), f"call {k}"


def test_hybrid_sequence_with_graphs_is_bit_identical(eng, monkeypatch):
    """Six frames of growing motion in hybrid mode: the direct leg's reference pack stays cached
    between steps (a graph without the pack's transforms), the adjacent leg is not cached."""
    from optflow_soil_amd import synth
    h, w = 192, 224
    pat = synth.make_pattern(h, w, seed=47)
    frames = [synth.as_float(synth.render(pat, synth.stretch_translate(0.004 * k, (0.7 * k, -0.4 * k),
                                                                        (w / 2, h / 2)) if k else None,
                                          noise=0.005, seed=k)) for k in range(6)]
    cfg = Config(seq_mode="hybrid")
    monkeypatch.setenv("OFX_GRAPHS", "0")
    want = [r for _, r in eng.iter_sequence(iter(frames), cfg)]
    monkeypatch.delenv("OFX_GRAPHS")
    cap0, run0 = eng.graph_stats()
    for rep in range(2):
        got = [r for _, r in eng.iter_sequence(iter(frames), cfg)]
        for k, (g, w_) in enumerate(zip(got, want)):
            same(g, w_, f"pass {rep} frame {k + 1}")
            assert np.array_equal(g.fb, w_.fb)
    cap, run = eng.graph_stats()
    print(f"sequence: captures {cap - cap0}, replays {run - run0}", flush=True)
    assert run - run0 >= 4


def test_schedules_do_not_change_results_and_the_autotuner_picks_one(eng):
    """Tile groups (1 .. 4) and the graph switch change how a tiled pair runs, not its result; the
    autotuner times the candidates on a sample and returns the fastest."""
    import optflow_soil_amd as ofx
    from optflow_soil_amd import autotune
    a, b = pair(288, 416, 51)
    cfg = Config(tile=128)
    ref = None
    for tg in (1, 2, 4):
        for gr in (True, False):
            eng.set_schedule(tile_groups=tg, graphs=gr)
            r = eng.track_pair(a, b, cfg)
            if ref is None:
                ref = r
            else:
                same(r, ref, f"tile groups {tg}, graphs {gr}")
    eng.set_schedule(tile_groups=3, graphs=True)
    assert eng.schedule() == {"tile_groups": 3, "graphs": True}
    pairs = [pair(192, 256, 53) + (None,), pair(160, 224, 55) + (None,), pair(192, 256, 57) + (None,)]
    best, table = autotune.tune(pairs, Config(), lanes=(1, 2), repeats=1)
    print(f"autotune table {table}", flush=True)
    assert len(table) == 4 and best in table and best["mpix_per_s"] == max(r["mpix_per_s"] for r in table)
    one = ofx.Engine(0, lanes=best["lanes"])
    try:
        one.set_schedule(graphs=best["graphs"])
        got = one.track_pairs(pairs, Config())
    finally:
        one.close()
    for (x, y, _), r in zip(pairs, got):
        same(r, eng.track_pair(x, y, Config()), "tuned batch")
