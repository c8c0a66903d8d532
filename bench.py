"""Benchmark: Mpix/s of flow + strain for one 2048 x 2048 speckle pair on one MI355X.

Workload (the plain method of the benchmark's single-pair configuration): a 2048 x 2048 uint8
speckle pair (seeded sum of Gaussian dots) with a settlement bowl uy = -6 exp(-r^2 / (2 300^2)) px
and a 40 px wide shear band at 35 deg with a 3 px tanh slip, noise sigma 0.005, a seeded convex
8-gon ROI of ~60 %, default settings. One step = one whole pair (coarse POC, fine stage,
confidence, guard, strain) with the inputs already resident in HBM, through the device-pointer
C ABI. Reports the per-stage GPU times, the roofline of the dominant kernel, the EPE against the
synthetic truth and a CPU baseline (the FP64 oracle on a 1024 x 1024 crop of the same pair, timed on
the host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpix/s flow+strain at 2048², plain GPU vs learned-coarse; EPE px vs truth"


def accumulate_bytes(w, h, F):
    """Algorithmic bytes of one accumulate launch: every input read once, every output written once.
    Per pixel: ux, uy (f64) 16 B, s (2 x f32) 8 B, spline of s (2 x f32) 8 B, reference pack 16 B per
    filter (re, im, gx, gy as f32), seven f64 sums 56 B; plus each node of the F padded response
    planes (c32, 8 B) once."""
    n, np_ = w * h, (w + 64) * (h + 64)
    return n * (16 + 8 + 8 + 16 * F + 56) + 8 * F * np_


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", choices=("c2", "c4"), default="c2",
                    help="c2: one 2048^2 pair (default); c4: 64 uneven pairs through the batched entry point")
    ap.add_argument("--lanes", type=int, default=4, help="engine lanes for --workload c4")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (twice) and exit: the reference kernel that scales "
                         "the HBM byte counters of a profiler run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this engine and its benchmark have no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4}))
        return

    if args.workload == "c4":
        bench_c4(args)
        return

    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth
    from optflow_soil_amd.config import Config

    size = args.size
    t = time.perf_counter()
    a, b, roi, (tx, ty) = synth.workload_c2(size)
    print(f"# rendered {size}x{size} pair in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    cfg = Config()
    dev = torch.device("cuda:0")
    d_a = torch.from_numpy(synth.as_float(a)).to(dev)
    d_b = torch.from_numpy(synth.as_float(b)).to(dev)
    d_roi = torch.from_numpy(roi.astype(np.uint8)).to(dev)
    outs = {k: torch.empty((size, size), dtype=torch.float32, device=dev)
            for k in ("ux", "uy", "conf", "raw_ux", "raw_uy", "exx", "eyy", "exy", "gmax", "vol")}
    outs["valid"] = torch.empty((size, size), dtype=torch.uint8, device=dev)
    ptrs = {k: v.data_ptr() for k, v in outs.items()}
    eng = ofx.Engine(0)
    eng.check_window(size, size, cfg)

    def step():
        return eng.track_pair_device(d_a.data_ptr(), d_b.data_ptr(), size, size, cfg, ptrs, d_mask=d_roi.data_ptr())

    for i in range(args.warmup):
        step()
        print(f"# warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    eng.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = eng.profile_read()
    eng.profile(False)
    dt = (t1 - t0) / args.steps
    value = size * size / 1e6 / dt

    # accuracy of the last step against the synthetic truth (valid pixels inside the ROI)
    ux, uy = outs["ux"].cpu().numpy(), outs["uy"].cpu().numpy()
    fin = np.isfinite(ux) & np.isfinite(uy)
    epe = float(np.sqrt(np.mean((ux[fin] - tx[fin]) ** 2 + (uy[fin] - ty[fin]) ** 2))) if fin.any() else float("nan")
    valid_frac = float(fin[roi].mean())

    # roofline of the accumulate kernel (one launch per inner iteration)
    acc_ms, acc_calls = prof["accumulate"]
    acc_avg_s = acc_ms / max(acc_calls, 1) * 1e-3
    acc_bytes = accumulate_bytes(size, size, cfg.n_filters())
    achieved = acc_bytes / acc_avg_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "accumulate_traffic_2048.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    stages = {k: round(v[0] / args.steps, 3) for k, v in prof.items() if v[1]}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(a, b, roi, cfg)

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"single pair {size}x{size} uint8 speckle, settlement bowl + 35 deg shear band, "
                               f"convex 8-gon ROI ~60 %, plain POC-initialised method, default settings, untiled",
                   "pixels_per_step": size * size},
        "roofline": {"kernel": "k_accumulate", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": acc_bytes, "avg_launch_ms": round(acc_avg_s * 1e3, 4)},
        "stage_ms_per_step": stages,
        "epe_px": round(epe, 5), "valid_frac_roi": round(valid_frac, 5),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def bench_c4(args):
    """64 independent pairs of uneven sizes (512^2 to 1024 x 2048, mean 1.31 Mpix) with their own
    affine fields and ROIs, through ofx_track_pairs on --lanes lanes (one host thread and stream
    each). One step = the whole batch, host arrays in and out (the copies overlap other lanes'
    work). Also times the same batch one pair at a time on one lane."""
    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth
    from optflow_soil_amd.config import Config

    t = time.perf_counter()
    raw = synth.workload_c4()
    pairs = [(synth.as_float(a), synth.as_float(b), roi) for a, b, roi in raw]
    mpix = sum(p[0].size for p in pairs) / 1e6
    print(f"# rendered {len(pairs)} pairs, {mpix:.1f} Mpix, in {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)
    cfg = Config()
    eng = ofx.Engine(0, lanes=args.lanes)
    for i in range(args.warmup):
        eng.track_pairs(pairs, cfg)
        print(f"# warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = eng.track_pairs(pairs, cfg)
        print(f"# step {i + 1}/{args.steps} done", file=sys.stderr, flush=True)
    dt = (time.perf_counter() - t0) / args.steps
    one = ofx.Engine(0, lanes=1)
    one.track_pair(*pairs[0][:2], cfg, mask=pairs[0][2])
    t0 = time.perf_counter()
    for a, b, roi in pairs:
        one.track_pair(a, b, cfg, mask=roi)
    dt_seq = time.perf_counter() - t0
    print("# sequential pass done", file=sys.stderr, flush=True)
    valid = float(np.mean([np.isfinite(r.ux)[p[2]].mean() for r, p in zip(out, pairs)]))
    line = {
        "metric": "aggregate Mpix/s, 64 uneven pairs batched", "value": round(mpix / dt, 3), "unit": "Mpix/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 uneven pairs (512^2, 768x1024, 1024^2, 1536^2, 1024x2048), affine fields, "
                               "own ROIs 40-90 %, default settings, host arrays in and out",
                   "lanes": args.lanes, "mpix_per_step": round(mpix, 2)},
        "sequential_value": round(mpix / dt_seq, 3), "batched_over_sequential": round(dt_seq / dt, 3),
        "valid_frac_roi": round(valid, 5),
    }
    print(json.dumps(line))


def cpu_baseline(a, b, roi, cfg):
    """The FP64 oracle (C++/OpenMP restatement of the reference CPU path) on the central 1024 x 1024
    crop of the same pair, all host cores up to 16, one timed run."""
    import oracle
    from optflow_soil_amd import synth
    cores = min(os.cpu_count() or 1, 16)
    oracle.set_threads(cores)
    c = 1024
    y0 = (a.shape[0] - c) // 2
    sl = np.s_[y0:y0 + c, y0:y0 + c]
    A = synth.as_float(a[sl]).astype(np.float64)
    Bm = synth.as_float(b[sl]).astype(np.float64)
    t = time.perf_counter()
    oracle.track_pair(A, Bm, cfg, mask=roi[sl])
    dt = time.perf_counter() - t
    return {"value": round(c * c / 1e6 / dt, 4), "unit": "Mpix/s", "cores": cores, "kind": "port",
            "sample": f"central {c}x{c} crop of the benchmark pair (same ROI), one pair, {dt:.2f} s"}


if __name__ == "__main__":
    main()
