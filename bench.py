"""Benchmark: Mpix/s of flow + strain for one 2048 x 2048 speckle pair on one MI355X.

Workload (the plain method of the benchmark's single-pair configuration): a 2048 x 2048 uint8
speckle pair (seeded sum of Gaussian dots) with a settlement bowl uy = -6 exp(-r^2 / (2 300^2)) px
and a 40 px wide shear band at 35 deg with a 3 px tanh slip, noise sigma 0.005, a seeded convex
8-gon ROI of ~60 %, default settings. One step = one whole pair (coarse POC, fine stage,
confidence, guard, strain) with the inputs already resident in HBM, through the device-pointer
C ABI. `value` is the plain method (the reference's POC-pyramid coarse stage); `learned_coarse`
times the same pair with the learned coarse proposal in place of the pyramid's coarse levels.
Reports the per-stage GPU times, the roofline of the dominant kernel, the EPE against the synthetic
truth, the error against the FP64 oracle and a CPU baseline (the oracle on a 1024 x 1024 crop of
the same pair, or the whole frame with --cpu-crop 0, timed on the host's cores).
"""
import argparse
import dataclasses
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


FP64_PEAK_TFLOPS = 78.6
# f64 multiply-adds per pixel of one window-moment pass pair (sigma_w 4, 33 taps): the y pass forms
# 15 planes from the 7 sums (15 x 33), the x pass the 26 moments from those (26 x 33)
MOMENT_FMA_PER_PIXEL = 15 * 33 + 26 * 33


def pipeline_bytes(w, h, tile=0):
    """Algorithmic HBM bytes of one whole pair with the default bank (F = 12 filters, 8 outer and
    2 inner iterations), each intermediate read once and written once: 6388 B per frame pixel
    plus 9120 B per padded-plane node of the fine stage, and about 100 B per pixel for the
    coarse stage, guard and strain. Tiled, the planes are those of the tile extensions."""
    if tile and max(w, h) > tile:
        m = 64
        n, np_ = 0, 0
        for ky in range(-(-h // tile)):
            for kx in range(-(-w // tile)):
                ew = min(w, (kx + 1) * tile + m) - max(0, kx * tile - m)
                eh = min(h, (ky + 1) * tile + m) - max(0, ky * tile - m)
                n += ew * eh
                np_ += (ew + 64) * (eh + 64)
        return 6388 * n + 9120 * np_ + 100 * w * h
    return 6388 * w * h + 9120 * (w + 64) * (h + 64) + 100 * w * h


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-crop", type=int, default=1024,
                    help="side of the central crop the CPU baseline times (0: the whole frame, tiled at 1024 "
                         "as the reference CLI tiles it; about four times as long at 2048^2)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-coarse form of the pair")
    ap.add_argument("--frames", type=int, default=12, help="frames of the --workload c3 sequence")
    ap.add_argument("--workload", choices=("c2", "c3", "c4", "c5"), default="c2",
                    help="c2: one 2048^2 pair (default); c4: 64 uneven pairs through the batched entry point; "
                         "c3: a 2048^2 hybrid sequence with temporal averaging (--frames frames); "
                         "c5: one 4000 x 3000 pair with displacements up to 40 px (plain POC-initialised method)")
    ap.add_argument("--lanes", default="auto",
                    help="engine lanes for --workload c4: a number, or auto (the autotuner picks lanes and the "
                         "graph switch on a sample of the batch)")
    ap.add_argument("--tile", type=int, default=0,
                    help="fine-stage tile size for c2/c5 (0: whole frame; 1024 is the reference CLI's tiling)")
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
    if args.workload == "c3":
        bench_c3(args)
        return

    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth
    from optflow_soil_amd.config import Config

    print(f"# rendering the {args.workload} pair", file=sys.stderr, flush=True)
    t = time.perf_counter()
    if args.workload == "c5":
        h, w = 3000, 4000
        a, b, (tx, ty) = synth.workload_c5()
        roi = None
        cfg = Config(max_disp=48.0, tile=args.tile)
        metric = "Mpix/s flow+strain at 4000x3000, plain GPU vs learned-coarse; EPE px vs truth"
        desc = ("single pair 4000x3000 uint8 speckle, six Gaussian-bump fields up to 40 px, noise 0.01, "
                "plain POC-initialised method, max_disp 48 (two POC levels), full frame, "
                + (f"tiles of {args.tile}" if args.tile else "untiled"))
    else:
        h = w = args.size
        a, b, roi, (tx, ty) = synth.workload_c2(args.size)
        cfg = Config(tile=args.tile)
        metric = METRIC
        desc = (f"single pair {w}x{h} uint8 speckle, settlement bowl + 35 deg shear band, "
                f"convex 8-gon ROI ~60 %, plain POC-initialised method, default settings, "
                + (f"tiles of {args.tile}" if args.tile else "untiled"))
    print(f"# rendered {w}x{h} pair in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    d_a = torch.from_numpy(synth.as_float(a)).to(dev)
    d_b = torch.from_numpy(synth.as_float(b)).to(dev)
    d_roi = None if roi is None else torch.from_numpy(roi.astype(np.uint8)).to(dev)
    outs = {k: torch.empty((h, w), dtype=torch.float32, device=dev)
            for k in ("ux", "uy", "conf", "raw_ux", "raw_uy", "exx", "eyy", "exy", "gmax", "vol")}
    outs["valid"] = torch.empty((h, w), dtype=torch.uint8, device=dev)
    ptrs = {k: v.data_ptr() for k, v in outs.items()}
    eng = ofx.Engine(0)
    eng.check_window(w, h, cfg)

    def timed(c, label):
        """warm-up, then args.steps timed pairs; returns (seconds per pair, stage profile of a
        second, profiled pass, last step's timing record, EPE, valid fraction)"""
        def step():
            return eng.track_pair_device(d_a.data_ptr(), d_b.data_ptr(), w, h, c, ptrs,
                                         d_mask=0 if d_roi is None else d_roi.data_ptr())
        for i in range(args.warmup):
            step()
            print(f"# {label}: warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            rec = step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # the stage times from a second pass with the profiler on (events between the stages; the
        # fine stage then runs eagerly instead of from its graph), not part of the timed value
        eng.profile(True)
        for i in range(args.steps):
            step()
        torch.cuda.synchronize()
        prof = eng.profile_read()
        eng.profile(False)
        # accuracy of the last step against the synthetic truth (valid pixels inside the ROI)
        ux, uy = outs["ux"].cpu().numpy(), outs["uy"].cpu().numpy()
        fin = np.isfinite(ux) & np.isfinite(uy)
        epe = float(np.sqrt(np.mean((ux[fin] - tx[fin]) ** 2 + (uy[fin] - ty[fin]) ** 2))) if fin.any() else float("nan")
        valid_frac = float(fin[roi].mean()) if roi is not None else float(fin.mean())
        return (t1 - t0) / args.steps, prof, rec, epe, valid_frac

    dt, prof, _, epe, valid_frac = timed(cfg, "plain")
    value = w * h / 1e6 / dt
    learned = None
    if not args.no_learned:
        eng.load_coarse_net()
        cfg_l = dataclasses.replace(cfg, coarse_model="learned")
        dt_l, prof_l, rec_l, epe_l, valid_l = timed(cfg_l, "learned coarse")
        learned = {
            "value": round(w * h / 1e6 / dt_l, 3), "ms_per_step": round(dt_l * 1e3, 3),
            "over_plain": round(dt / dt_l, 4), "epe_px": round(epe_l, 5), "valid_frac_roi": round(valid_l, 5),
            "coarse_ms_per_step": round(prof_l["coarse"][0] / args.steps, 3),
            "plain_coarse_ms_per_step": round(prof["coarse"][0] / args.steps, 3),
            "coarse_confirmed": round(rec_l["coarse_confirmed"], 4), "fallback": rec_l["coarse_fallback"],
            "network": "6 x conv3x3 bf16 (169 -> 32 -> 32 -> 32 -> 32 -> 32 -> 2), fp32 accumulate, "
                       "optflow_soil_amd/weights/coarse_net.npy"}

    # roofline of the accumulate kernel: one launch per inner iteration over the frame, or per tile
    # and inner iteration when tiled (then the frame's bytes per iteration over the tiles' total time)
    acc_ms, acc_calls = prof["accumulate"]
    acc_avg_s = acc_ms / max(acc_calls, 1) * 1e-3
    acc_bytes = accumulate_bytes(w, h, cfg.n_filters())
    n_tiles = -(-w // cfg.tile) * -(-h // cfg.tile) if cfg.tile else 1
    frame_passes = acc_calls / n_tiles
    achieved = acc_bytes * frame_passes / (acc_ms * 1e-3) / 1e9 if acc_ms > 0 else float("nan")
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "accumulate_traffic_2048.json")
    if args.workload == "c2" and args.size == 2048 and os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    stages = {k: round(v[0] / args.steps, 3) for k, v in prof.items() if v[1]}
    # the second kernel pair by time: the window moments (y pass, then x pass with the 6 x 6 solve),
    # bound by f64 arithmetic; flops of the moments only (the solve's are not counted)
    mom_ms, mom_calls = prof["moments"]
    mom_flop = 2.0 * MOMENT_FMA_PER_PIXEL * w * h * (mom_calls / n_tiles)
    moments_roof = None
    if mom_ms > 0 and cfg.shape_order == 1 and cfg.sigma_w == 4.0:
        got = mom_flop / (mom_ms * 1e-3) / 1e12
        moments_roof = {"kernels": "k_moments_y_rb<16> + k_moments_x_rb<16>", "bound": "fp64", "achieved": round(got, 2),
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(got / FP64_PEAK_TFLOPS, 4),
                        "flop_per_launch_pair": 2 * MOMENT_FMA_PER_PIXEL * w * h,
                        "avg_launch_pair_ms": round(mom_ms / max(mom_calls, 1), 4)}

    writer = output_writer_rate(outs, w, h)
    cpu = err = None
    if not args.no_cpu:
        cpu, crop_info = cpu_baseline(a, b, roi, cfg, crop=args.cpu_crop)
        err = err_vs_oracle(eng, a, b, crop_info)

    line = {
        "metric": metric, "value": round(value, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "pixels_per_step": w * h},
        "roofline": {"kernel": "k_accumulate", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": ("stored measurement: profiles/accumulate_traffic_2048.json, two rocprofv3 "
                                        "--pmc passes (FETCH_SIZE, WRITE_SIZE) scaled by a 1 GiB copy")
                     if traffic is not None else None,
                     "algorithmic_bytes_per_launch": acc_bytes, "avg_launch_ms": round(acc_avg_s * 1e3, 4)},
        "roofline_moments": moments_roof,
        "stage_ms_per_step": stages,
        "epe_px": round(epe, 5), "valid_frac_roi": round(valid_frac, 5), "err_vs_oracle": err,
        "learned_coarse": learned,
        "pipeline": {"bound": "hbm", "bytes_per_step": pipeline_bytes(w, h, cfg.tile),
                     "achieved": round(pipeline_bytes(w, h, cfg.tile) / dt / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s",
                     "frac": round(pipeline_bytes(w, h, cfg.tile) / dt / 1e9 / HBM_PEAK_GBS, 4)},
        "output_writer": writer,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def output_writer_rate(outs, w, h, runs=3):
    """The host rate of writing one pair's result folder (the reference CLI's files, the --raw maps
    included) through ofx_write_result, the writer the command-line twin uses: the last pair's maps
    copied to the host, then the folder written into a temporary directory `runs` times, the
    fastest kept. Not part of the timed value (the metric excludes I/O)."""
    import shutil
    import tempfile
    from optflow_soil_amd import Result, io
    host = {k: v.cpu().numpy() for k, v in outs.items()}
    strain = {"Exx": host["exx"], "Eyy": host["eyy"], "Exy": host["exy"], "gmax": host["gmax"], "vol": host["vol"]}
    res = Result(host["ux"], host["uy"], host["conf"], strain, host["raw_ux"], host["raw_uy"], host["valid"] != 0,
                 {"coarse": 0.0, "fine": 0.0, "strain": 0.0})
    d = tempfile.mkdtemp(prefix="ofx-writer-")
    try:
        best = float("inf")
        for k in range(runs):
            t0 = time.perf_counter()
            io.write_result(os.path.join(d, f"run{k}"), res, raw=True)
            best = min(best, time.perf_counter() - t0)
        folder = os.path.join(d, "run0")
        nbytes = sum(os.path.getsize(os.path.join(folder, f)) for f in os.listdir(folder))
        nfiles = len(os.listdir(folder))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"mpix_per_s": round(w * h / 1e6 / best, 1), "gb_per_s": round(nbytes / best / 1e9, 2),
            "ms_per_folder": round(best * 1e3, 2), "files": nfiles, "bytes": nbytes,
            "sample": f"the last pair's result folder with the raw maps, min of {runs} writes to a temporary "
                      "directory on the host"}


def bench_c3(args):
    """The long-sequence workload at 2048^2 (--frames frames, default 12; 200 is the full workload,
    rendered on the GPU above 24 frames): the field grows linearly to the bowl-and-band field at
    the last frame, noise 0.01, brightness drift -3 % .. +3 %.
    The frames are averaged over +-3 frames on the GPU (host frames in and out, timed on its own),
    then tracked one frame at a time from HBM with seq_mode hybrid, prewarp_freeze_at 3 and
    ref_update_conf 0.3 (ofx_seq_step_device). value: tracked frames per second."""
    import torch
    import optflow_soil_amd as ofx
    from optflow_soil_amd import synth
    from optflow_soil_amd.config import Config
    from optflow_soil_amd.frames import averaged, gpu_average

    n, nf = 2048, args.frames
    t = time.perf_counter()
    pat = synth.make_pattern(n, n, seed=0)
    full = synth.bowl_and_band(n, n, radius=300.0, band_width=40.0)

    def field_at(k):
        s = k / (nf - 1)
        return lambda X, Y: tuple(s * c for c in full(X, Y))
    frames = []
    if nf <= 24:
        for k in range(nf):
            v = synth.render(pat, field_at(k) if k else None, noise=0.0, quantise=False)
            v = v * (0.97 + 0.06 * k / (nf - 1)) + np.random.default_rng(100 + k).normal(0.0, 0.01, v.shape)
            frames.append(synth.as_float(np.clip(np.rint(v * 255.0), 0, 255).astype(np.uint8)))
            print(f"# rendered frame {k + 1}/{nf}", file=sys.stderr, flush=True)
    else:
        for k, f in enumerate(synth.growing_bowl_and_band_frames_gpu(n, nf)):
            frames.append(f)
            if (k + 1) % 20 == 0:
                print(f"# rendered frame {k + 1}/{nf}", file=sys.stderr, flush=True)
    print(f"# rendered {nf} frames in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    cfg = Config(seq_mode="hybrid", prewarp_freeze_at=3, ref_update_conf=0.3)
    eng = ofx.Engine(0)
    t0 = time.perf_counter()
    avg = list(averaged(iter(frames), 0, None, 1, 3, cfg=cfg, average=gpu_average(eng)))
    t_avg = time.perf_counter() - t0
    print(f"# averaged {nf} frames in {t_avg:.2f} s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    d_frames = [torch.from_numpy(np.ascontiguousarray(f, dtype=np.float32)).to(dev) for f in avg]
    outs = {k: torch.empty((n, n), dtype=torch.float32, device=dev)
            for k in ("ux", "uy", "conf", "raw_ux", "raw_uy", "exx", "eyy", "exy", "gmax", "vol")}
    outs["valid"] = torch.empty((n, n), dtype=torch.uint8, device=dev)
    outs["fb"] = torch.empty((n, n), dtype=torch.int32, device=dev)
    ptrs = {k: v.data_ptr() for k, v in outs.items()}
    with eng.sequence(avg[0], cfg) as q:  # warm-up: buffers, plans
        for k in (1, 2):
            q.step_device(d_frames[k].data_ptr(), ptrs)
    print("# warm-up sequence done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    shares = []
    with eng.sequence(avg[0], cfg) as q:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(1, nf):
            shares.append(q.step_device(d_frames[k].data_ptr(), ptrs))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    ux, uy = outs["ux"].cpu().numpy(), outs["uy"].cpu().numpy()
    gx, gy = field_at(nf - 1)(*np.mgrid[0:n, 0:n][::-1].astype(np.float64))
    fin = np.isfinite(ux) & np.isfinite(uy)
    drift = float(np.sqrt(np.mean((ux[fin] - gx[fin]) ** 2 + (uy[fin] - gy[fin]) ** 2)))
    line = {
        "metric": "frames/s, 2048^2 hybrid sequence with temporal averaging; drift vs truth at the last frame",
        "value": round((nf - 1) / dt, 3), "unit": "frames/s", "n_gpus": 1, "steps": nf - 1, "warmup": 2,
        "ms_per_step": round(dt / (nf - 1) * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{nf} frames 2048x2048 (the 200-frame workload{'' if nf >= 200 else ', shortened'}), "
                               f"field growing to the "
                               f"bowl-and-band field, noise 0.01, brightness drift +-3 %, tavg 3, seq_mode hybrid, "
                               f"prewarp_freeze_at 3, ref_update_conf 0.3, untiled",
                   "frames": nf},
        "mpix_per_s": round((nf - 1) * n * n / 1e6 / dt, 2),
        "averaging_s_per_frame": round(t_avg / nf, 3),
        "drift_px_last_frame": round(drift, 5), "valid_frac_last_frame": round(float(fin.mean()), 5),
        "references": sorted({s["reference"] for s in shares}),
        "fallback_frac_last": round(shares[-1]["fallback_frac"], 5),
    }
    print(json.dumps(line))


def bench_c4(args):
    """64 independent pairs of uneven sizes (512^2 to 1024 x 2048, mean 1.31 Mpix) with their own
    affine fields and ROIs, through ofx_track_pairs on --lanes lanes (one host thread and stream
    each; by default the autotuner's choice of lanes and graph switch). One step = the whole batch, host arrays in and out (the copies overlap other lanes'
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
    tuned = None
    if args.lanes == "auto":
        from optflow_soil_amd import autotune
        t = time.perf_counter()
        best, table = autotune.tune(pairs, cfg, log=lambda m: print("# " + m, file=sys.stderr, flush=True))
        tuned = {"chosen": best, "candidates": table, "seconds": round(time.perf_counter() - t, 1),
                 "sample": "the two largest pairs of each size class"}
        lanes, graphs = best["lanes"], best["graphs"]
    else:
        lanes, graphs = int(args.lanes), True
    eng = ofx.Engine(0, lanes=lanes)
    eng.set_schedule(graphs=graphs)
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
                   "lanes": lanes, "graphs": graphs, "mpix_per_step": round(mpix, 2)},
        "sequential_value": round(mpix / dt_seq, 3), "batched_over_sequential": round(dt_seq / dt, 3),
        "valid_frac_roi": round(valid, 5), "autotune": tuned,
    }
    print(json.dumps(line))


def host_threads():
    """Threads the CPU leg may use: OMP_NUM_THREADS when set (the host's CPU share for this
    process), otherwise every core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(a, b, roi, cfg, crop=1024, runs=3):
    """The FP64 oracle (this repository's C++/OpenMP restatement of the reference CPU path: the
    same algorithm and constants, f64 storage and arithmetic) on the central crop x crop window of
    the same pair, tiled as the reference CLI tiles (Config.tile = 1024: a 1024 x 1024 crop is one
    of the four tiles of the 2048 x 2048 frame), one untimed warm-up, then the minimum of `runs`
    timed runs, as the reference's bench_pair (examples/bench_pair.rs:12-17). Returns the
    baseline record and the oracle's result of the crop (for the error of the GPU against it)."""
    import oracle
    from optflow_soil_amd import synth
    cores = host_threads()
    oracle.set_threads(cores)
    full = crop <= 0 or crop >= min(a.shape)
    if full:
        sl = np.s_[:, :]
    else:
        y0, x0 = (a.shape[0] - crop) // 2, (a.shape[1] - crop) // 2
        sl = np.s_[y0:y0 + crop, x0:x0 + crop]
    A = synth.as_float(a[sl]).astype(np.float64)
    Bm = synth.as_float(b[sl]).astype(np.float64)
    m = None if roi is None else np.ascontiguousarray(roi[sl])
    ccfg = dataclasses.replace(cfg, tile=1024)
    times = []
    o = None
    for k in range(runs + 1):
        t = time.perf_counter()
        o = oracle.track_pair(A, Bm, ccfg, mask=m)
        dt = time.perf_counter() - t
        if k:
            times.append(dt)
        print(f"# cpu baseline run {k}/{runs}{' (warm-up)' if k == 0 else ''}: {dt:.2f} s", file=sys.stderr,
              flush=True)
    best = min(times)
    if full:
        what = f"the whole {a.shape[1]}x{a.shape[0]} benchmark pair in the reference CLI's 1024 px tiles"
    else:
        what = (f"central {crop}x{crop} crop of the benchmark pair (same ROI; one tile of the reference "
                f"CLI's 1024 px tiling)")
    rec = {"value": round(A.size / 1e6 / best, 4), "unit": "Mpix/s", "cores": cores, "kind": "port",
           "host_cpus": os.cpu_count(),
           "sample": (f"{what}, FP64 oracle, {cores} OpenMP threads, min of {runs} runs after a "
                      f"warm-up: " + ", ".join(f"{t:.2f}" for t in times) + " s"),
           "published_reference_cpu": "0.49-0.67 Mpix/s, Rust CPU path on 10 Apple M4 threads (rust/README.md:166-171)"}
    return rec, (sl, m, ccfg, o)


def err_vs_oracle(eng, a, b, crop_info):
    """The GPU's result on the CPU leg's crop against the oracle's: ux / uy RMS and largest
    difference over pixels valid on both sides, valid-mask agreement over the crop's ROI pixels,
    and the strain planes' RMS."""
    from optflow_soil_amd import synth
    sl, m, ccfg, o = crop_info
    r = eng.track_pair(synth.as_float(a[sl]), synth.as_float(b[sl]), ccfg, mask=m)
    vg, vo = np.isfinite(r.ux), np.isfinite(o["ux"])
    both = vg & vo
    region = np.ones(vg.shape, bool) if m is None else m.astype(bool)
    dx = r.ux.astype(np.float64)[both] - o["ux"][both]
    dy = r.uy.astype(np.float64)[both] - o["uy"][both]
    fin = np.isfinite(r.strain["Exx"]) & np.isfinite(o["Exx"])
    out = {"sample": "the CPU leg's crop, same settings (tiled at 1024)",
           "ux_rms": float(np.sqrt(np.mean(dx ** 2))), "uy_rms": float(np.sqrt(np.mean(dy ** 2))),
           "disp_max": float(max(np.abs(dx).max(), np.abs(dy).max())),
           "valid_agree_roi": float((vg == vo)[region].mean())}
    for p in ("Exx", "Eyy", "Exy", "gmax", "vol"):
        out[p + "_rms"] = float(np.sqrt(np.mean((r.strain[p].astype(np.float64)[fin] - o[p][fin]) ** 2)))
    return {k: (float(f"{v:.3g}") if isinstance(v, float) else v) for k, v in out.items()}


if __name__ == "__main__":
    main()
