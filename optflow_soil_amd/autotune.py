"""Autotuner of the engine's schedule for a workload of many pairs.

The knobs are those that change how the engine runs a batch, not what it computes (results are
bit-identical under every choice): the number of lanes (engine streams, each with its own
workspace, that take pairs side by side: ofx_create), the number of tile groups refined side by
side when the fine stage is tiled, and whether the fine stage replays HIP graphs
(ofx_set_schedule). tune() times every combination on a stratified sample of the workload (the
largest pairs of each size class, as ofx_track_pairs takes them largest first) after a warm-up of
the same sample, and returns the fastest with the whole table. Timing is by the wall clock around
ofx_track_pairs, as the batched workload's figure is.
"""
import itertools
import time


def sample_pairs(pairs, per_size=2):
    """Up to per_size pairs of each distinct (h, w), the largest sizes first."""
    by_size = {}
    for p in pairs:
        by_size.setdefault(p[0].shape, []).append(p)
    out = []
    for shape in sorted(by_size, key=lambda s: -s[0] * s[1]):
        out += by_size[shape][:per_size]
    return out


def tune(pairs, cfg, lanes=(2, 3, 4, 6), tile_groups=None, graphs=(True, False), per_size=2, repeats=2,
         device=0, log=None):
    """The fastest schedule for ofx_track_pairs on `pairs` (a list of (I1, I2, roi)) under cfg.
    tile_groups defaults to (1, 2, 3, 4) when cfg tiles and to (3,) otherwise (no effect then).
    Returns ({"lanes", "tile_groups", "graphs", "mpix_per_s"}, table of every candidate)."""
    import optflow_soil_amd as ofx
    if tile_groups is None:
        tile_groups = (1, 2, 3, 4) if cfg.tile else (3,)
    sample = sample_pairs(pairs, per_size)
    mpix = sum(p[0].size for p in sample) / 1e6
    table = []
    for n_lanes in lanes:
        eng = ofx.Engine(device, lanes=n_lanes)
        try:
            for tg, gr in itertools.product(tile_groups, graphs):
                eng.set_schedule(tile_groups=tg, graphs=gr)
                eng.track_pairs(sample, cfg)  # warm-up: plans, buffers, graph keys
                best = float("inf")
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    eng.track_pairs(sample, cfg)
                    best = min(best, time.perf_counter() - t0)
                row = {"lanes": n_lanes, "tile_groups": tg, "graphs": bool(gr), "mpix_per_s": round(mpix / best, 3)}
                table.append(row)
                if log:
                    log(f"autotune: {row}")
        finally:
            eng.close()
    best = max(table, key=lambda r: r["mpix_per_s"])
    return best, table
