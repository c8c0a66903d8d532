"""Trains the learned coarse network on the GPU (ofx_coarse_net_train) and writes its parameters.

The training pairs come from the package's speckle renderer (synth.make_pattern for the
patterns; the GPU renderer of cnn.hip for the warped, noisy, 8-bit frames): crops of 512 x 512 px
of four 2048 x 2048 patterns under fields of six Gaussian bumps (scales 150-600 px) over a
4000 x 3000 px frame whose largest |u| is 28-56 px, and noise of sd 0.01, the large-motion
workload's kind of field. 20 000 distinct pairs, seen twenty times (12 500 steps of 32).
Parameters start from He-normal weights (seeded) and zero biases; Adam, learning rate 2e-3
falling to 2e-5 by a cosine. Everything is seeded, and the GPU training is deterministic, so a run
reproduces the shipped weights/coarse_net.npy bit for bit on the same GPU and software.

    python -m optflow_soil_amd.train_coarse_net [--out PATH] [--steps N] ...
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

LAYERS = [(169, 32), (32, 32), (32, 32), (32, 32), (32, 32), (32, 2)]


def initial_params(seed):
    """He-normal weights (sd sqrt(2 / (9 cin))), the output layer's scaled by 0.1, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, (ci, co) in enumerate(LAYERS):
        sd = np.sqrt(2.0 / (9 * ci)) * (0.1 if k == len(LAYERS) - 1 else 1.0)
        parts += [rng.standard_normal(co * 9 * ci) * sd, np.zeros(co)]
    return np.concatenate(parts).astype(np.float32)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "coarse_net.npy"))
    ap.add_argument("--steps", type=int, default=12500)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--crop", type=int, default=512)
    ap.add_argument("--n-pairs", type=int, default=20000)
    ap.add_argument("--patterns", type=int, default=4)
    ap.add_argument("--pattern-size", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=2e-5)
    ap.add_argument("--max-disp", type=float, default=56.0)
    ap.add_argument("--noise", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)

    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth

    t0 = time.perf_counter()
    pats = np.stack([synth.make_pattern(a.pattern_size, a.pattern_size, seed=1000 + k)
                     for k in range(a.patterns)]).astype(np.float32)
    print(f"# rendered {a.patterns} patterns of {pats.shape[1]}x{pats.shape[2]} in {time.perf_counter() - t0:.1f} s",
          flush=True)
    p0 = initial_params(a.seed)
    eng = ofx.Engine(0)
    t0 = time.perf_counter()
    p, loss = eng.coarse_net_train(p0, pats, a.steps, batch=a.batch, crop=a.crop, n_pairs=a.n_pairs, lr=a.lr,
                                   lr_final=a.lr_final, max_disp=a.max_disp, noise=a.noise, seed=a.seed)
    dt = time.perf_counter() - t0
    eng.close()
    k = max(1, a.steps // 10)
    curve = [float(loss[i:i + k].mean()) for i in range(0, a.steps, k)]
    print(f"# trained {a.steps} steps of {a.batch} pairs in {dt:.1f} s; loss (cells) by tenths: "
          + " ".join(f"{v:.4f}" for v in curve), flush=True)
    if not np.isfinite(p).all():
        raise SystemExit("training diverged: non-finite parameters")
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, p)
    record = {"settings": vars(a), "seconds": round(dt, 2), "loss_by_tenths_cells": curve,
              "final_loss_cells": float(loss[-k:].mean()), "parameters": int(p.size)}
    with open(os.path.splitext(a.out)[0] + ".json", "w") as fh:
        json.dump(record, fh, indent=2)
    print(json.dumps({k2: record[k2] for k2 in ("seconds", "final_loss_cells", "parameters")}))


if __name__ == "__main__":
    main()
