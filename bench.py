"""FDK benchmark on one MI355X: 512^3 volume from 720 views of 768^2.

A step is one full FDK pass over resident projections: FDK cosine weight +
ramp filter (FFT length 2048) into the staged layout, then the voxel-driven
cone backprojection of all 720 views.  Inputs are synthetic: 30 seeded random
ellipsoids rendered at 0.5 mm, projected by the ray-driven forward projector,
Poisson noise at I0 = 1e5 and log.  The headline value is the FP16-stored
form (filtered projections in FP16 with a per-view scale, FP32 interpolation
and accumulation); the FP32-stored and BF16-stored forms are timed beside it,
and the 16-bit forms once more with the 16-bit interpolation arithmetic
(workspace option CTR_OPTION_INTERPOLATION = CTR_INTERP_FAST) as a secondary
figure.

Prints ONE JSON line.  Roofline for the dominant kernel (backprojection):
algorithmic bytes per voxel-view update = 4 taps x element size + 8/N bytes
of volume traffic, divided by the kernel's average duration (HIP events on
the launch stream), against 8 TB/s.  The taps are served from LDS, so that
fraction exceeds 1; `traffic_frac` is the HBM bytes the counters measured per
launch over the same duration, against the same peak.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

HBM_PEAK_GBPS = 8000.0
WORKLOAD = "fdk_512"
FFT_LENGTH = 2048
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "backproject_fdk512_traffic.md")


def read_traffic(key):
    """HBM bytes per backprojection launch measured with rocprofv3 FETCH_SIZE
    and WRITE_SIZE (each scaled by the same counter's reading for a plain copy
    of known size), as recorded in the profile summary: a line `key = N`."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            parts = line.strip().strip("`").split("=")
            if len(parts) == 2 and parts[0].strip() == key:
                return int(parts[1].strip())
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs(ctr, g, dev, seed=20240602):
    import torch
    ell = ctr.random_ellipsoids(seed, 30, (g.X, g.Y, g.Z), g.voxel, (5, 120))
    ell[:, 9] *= 0.02  # density 1.0 = 0.02 / mm
    vol = torch.zeros((g.Z, g.Y, g.X), dtype=torch.float32, device=dev)
    ctr.render_ellipsoids(ell, vol, g.volume_width)
    t0 = time.time()
    proj = ctr.forward_project(vol, g, step=1.0)
    ctr.poisson_log(proj, 1e5, seed)
    torch.cuda.synchronize()
    log(f"inputs: phantom + forward projection + noise in {time.time() - t0:.1f} s")
    del vol
    return proj


def time_fdk(ctr, g, proj, dtype, steps, warmup, ws, scale):
    import torch
    staged = None
    vs = None
    vol = torch.empty((g.Z, g.Y, g.X), dtype=torch.float32, device=proj.device)
    stream = torch.cuda.current_stream()
    step_ms, bp_ms = [], []
    for it in range(warmup + steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e2 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        staged, vs = ctr.filter_projections(proj, g, ctr.RAM_LAK, FFT_LENGTH, dtype,
                                            fdk=True, ws=ws, staged=staged, view_scale=vs)
        e1.record(stream)
        ctr.backproject(staged, g, dtype, view_scale=vs, scale=scale, volume=vol, ws=ws)
        e2.record(stream)
        e2.synchronize()
        if it >= warmup:
            step_ms.append(e0.elapsed_time(e2))
            bp_ms.append(e1.elapsed_time(e2))
    return step_ms, bp_ms, vol


def cpu_baseline(g, proj_cpu, scale, threads, budget_s=15.0):
    """The oracle's FP32 voxel loop + reference-algorithm row filter on a
    bounded sample of the same workload, extrapolated to the full job."""
    import oracle
    n_rows = g.nv * g.n_views
    sample_rows = min(n_rows, 4096)
    rows = proj_cpu.reshape(-1, g.nu)[:sample_rows]
    t0 = time.perf_counter()
    oracle.filter_rows(rows, oracle.RAM_LAK, FFT_LENGTH, threads=threads)
    t_filter = (time.perf_counter() - t0) * n_rows / sample_rows
    q = proj_cpu  # filter output values do not change the voxel loop's cost
    slices = 1
    while True:
        t0 = time.perf_counter()
        oracle.backproject_cone_f32(q, g.source_distance, g.detector_distance,
                                    g.detector_width, g.volume_width, 1.0, 0.0, scale,
                                    (g.X, g.Y, g.Z), (g.Z // 2, g.Z // 2 + slices),
                                    threads=threads)
        dt = time.perf_counter() - t0
        if dt > budget_s / 3 or slices >= g.Z:
            break
        slices = min(g.Z, max(slices * 2, int(slices * budget_s / 3 / max(dt, 1e-3))))
    t_bp = dt * g.Z / slices
    total = t_filter + t_bp
    return {
        "value": g.updates / total / 1e9,
        "unit": "GUPS",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle filter of {sample_rows} of {n_rows} detector rows "
                   f"(FFT {FFT_LENGTH}, reference radix-2 algorithm, FP64) + FP32 "
                   f"backprojection of {slices} of {g.Z} z-slices x {g.n_views} views; "
                   f"times scaled to the full job (filter {t_filter:.1f} s, "
                   f"backprojection {t_bp:.1f} s)"),
    }


def serving_main(args):
    """Poisson stream of FDK requests over three sizes (256^3/360/384^2,
    512^3/720/768^2, 1024^3/1200/1024^2), FP16 filtered projections, FP32
    volumes, 200 GB HBM budget: p50/p99 latency and aggregate GUPS."""
    import torch
    ctr = __graft_entry__.load()
    ctr.require_device()
    from ctrecon import serving
    small, medium, large = (int(x) for x in args.max_slots.split(","))
    srv = serving.Server(budget_bytes=args.budget_gb * 1e9, dtype=ctr.F16, pack=args.pack,
                         max_slots={"small": small, "medium": medium, "large": large},
                         pool=2, log=log)
    classes = ["small", "medium", "large"]
    t0 = time.time()
    srv.prepare_sources(classes)
    log(f"inputs ready in {time.time() - t0:.1f} s")
    srv.warm(classes)
    torch.cuda.synchronize()
    reqs = serving.poisson_arrivals(args.requests, args.rate, seed=7)
    log(f"serving {args.requests} requests at {args.rate}/s")
    makespan = srv.run(reqs)
    s = serving.summarize(reqs, makespan, srv.geoms)
    out = {
        "metric": "FDK serving latency (Poisson arrivals, sizes 256^3/512^3/1024^3, FP16 store)",
        "value": round(s["p99_ms"], 2),
        "unit": "ms (p99 latency)",
        "n_gpus": 1,
        "higher_is_better": False,
        "latency_ms": {"p50": round(s["p50_ms"], 2), "p99": round(s["p99_ms"], 2),
                       "mean": round(s["mean_ms"], 2)},
        "aggregate_gups": round(s["aggregate_gups"], 2),
        "throughput_rps": round(s["throughput_rps"], 3),
        "rate_rps": args.rate,
        "requests": args.requests,
        "makespan_s": round(makespan, 2),
        "hbm_budget_gb": args.budget_gb,
        "max_slots": args.max_slots,
        "slots": {c: sum(1 for x in srv.slots if x.cls == c) for c in classes},
        "slot_gb": round(srv.used / 1e9, 2),
        "per_class": s["per_class"],
        "dtype": "f32",
        "data": "synthetic: seeded ellipsoid phantoms, ray-driven projections, Poisson I0=1e5 "
                "(2 phantoms per size class, reused by request seed)",
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--mode", choices=["fdk", "serving", "copy-probe"], default="fdk")
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=3.0)
    ap.add_argument("--pack", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    # one 1024^3 slot: a 1024^3 reconstruction fills the GPU for ~0.35 s, and
    # several in flight delay every smaller request behind them
    ap.add_argument("--max-slots", default="2,6,1",
                    help="serving slots per class: small,medium,large")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(the reconstruction would shard into z-slabs with no exchange; "
              "a slab is the same kernel with fewer y-runs)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "serving":
        serving_main(args)
        return
    if args.mode == "copy-probe":
        # a plain elementwise kernel that reads 1 GiB and writes 1 GiB, the
        # reference for scaling the HBM byte counters (FETCH_SIZE /
        # WRITE_SIZE) of the profiled kernels.  (A clone goes through the
        # runtime's copy path instead, whose counter readings for the same
        # 1 GiB varied between runs by a quarter.)
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.neg(x, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 << 28, "checksum": float(y[:16].sum())}))
        return

    import torch
    ctr = __graft_entry__.load()
    ctr.require_device()
    dev = torch.device("cuda:0")
    g = ctr.config_geometry(WORKLOAD)
    scale = g.fbp_scale(FFT_LENGTH)
    ws = ctr.Workspace()
    ws.reserve(g.n_views, FFT_LENGTH)
    log(f"workload {WORKLOAD}: {g.X}x{g.Y}x{g.Z} from {g.n_views} views of "
        f"{g.nu}x{g.nv}, FFT {FFT_LENGTH}")
    proj = make_inputs(ctr, g, dev)
    ws.set_interpolation(ctr.INTERP_FP32)

    log("timing FP32-stored form")
    f32_ms, f32_bp, vol32 = time_fdk(ctr, g, proj, ctr.F32, args.steps, args.warmup, ws, scale)
    log(f"  step {statistics.median(f32_ms):.2f} ms, backprojection {statistics.median(f32_bp):.2f} ms")
    ref32 = vol32.clone()
    log("timing FP16-stored form")
    f16_ms, f16_bp, vol16 = time_fdk(ctr, g, proj, ctr.F16, args.steps, args.warmup, ws, scale)
    log(f"  step {statistics.median(f16_ms):.2f} ms, backprojection {statistics.median(f16_bp):.2f} ms")
    diff = (vol16.double() - ref32.double()).norm() / ref32.double().norm()
    rel_rmse_16 = float(diff)
    log("timing BF16-stored form")
    b16_ms, b16_bp, volb = time_fdk(ctr, g, proj, ctr.BF16, max(3, args.steps // 2), 1, ws, scale)
    rel_rmse_bf = float((volb.double() - ref32.double()).norm() / ref32.double().norm())
    del volb
    log("timing FP16 / BF16 with the 16-bit interpolation arithmetic")
    ws.set_interpolation(ctr.INTERP_FAST)
    fast = {}
    for name, dt in (("fp16", ctr.F16), ("bf16", ctr.BF16)):
        t_ms, t_bp, v = time_fdk(ctr, g, proj, dt, max(3, args.steps // 2), 1, ws, scale)
        fast[name] = {
            "gups": round(g.updates / (statistics.mean(t_ms) * 1e-3) / 1e9, 3),
            "ms_per_step": round(statistics.mean(t_ms), 3),
            "backproject_ms": round(statistics.mean(t_bp), 3),
            "rel_rmse_vs_fp32": float((v.double() - ref32.double()).norm()
                                      / ref32.double().norm())}
        del v
    ws.set_interpolation(ctr.INTERP_FP32)

    torch.cuda.synchronize()
    ms = statistics.mean(f16_ms)
    gups = g.updates / (ms * 1e-3) / 1e9
    gups32 = g.updates / (statistics.mean(f32_ms) * 1e-3) / 1e9
    bp_ms = statistics.mean(f16_bp)
    bp_ms32 = statistics.mean(f32_bp)
    b_alg16 = (4 * 2 + 8.0 / g.n_views) * g.updates
    b_alg32 = (4 * 4 + 8.0 / g.n_views) * g.updates
    ach16 = b_alg16 / (bp_ms * 1e-3) / 1e9
    ach32 = b_alg32 / (bp_ms32 * 1e-3) / 1e9
    traffic = read_traffic("fp16_bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"CPU baseline on {threads} threads")
        proj_cpu = proj.cpu().numpy()
        cpu = cpu_baseline(g, proj_cpu, scale, threads, args.cpu_budget)
        log(f"  {cpu['value']:.3f} GUPS")
        if threads > 1:
            log("CPU baseline on 1 thread")
            one = cpu_baseline(g, proj_cpu, scale, 1, min(6.0, args.cpu_budget / 2))
            cpu["single_core"] = {"value": one["value"], "unit": "GUPS", "cores": 1,
                                  "sample": one["sample"]}
            log(f"  {one['value']:.4f} GUPS")
        del proj_cpu

    out = {
        "metric": "FDK GUPS 512^3 x 720 views (FP16-stored filtered projections, FP32 "
                  "interpolation and accumulation)",
        "value": round(gups, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 30 seeded ellipsoids, ray-driven projections, Poisson I0=1e5",
        "config": {"workload": "FDK cone 512x512x512 @0.5mm, 720 views/360deg, "
                               "768x768 panel @0.8mm, SOD 500, SDD 1000, FFT 2048, Ram-Lak",
                   "storage": "fp16 + per-view scale"},
        "latency_ms": {"p50": round(statistics.median(f16_ms), 3),
                       "p99": round(float(np.percentile(f16_ms, 99)), 3)},
        "fp32_store": {"gups": round(gups32, 3), "ms_per_step": round(statistics.mean(f32_ms), 3),
                       "backproject_ms": round(bp_ms32, 3),
                       "roofline_frac": round(ach32 / HBM_PEAK_GBPS, 4),
                       "traffic": read_traffic("fp32_bytes_per_launch")},
        "fp16_vs_fp32_gups": round(gups / gups32, 3),
        "rel_rmse_fp16_vs_fp32": rel_rmse_16,
        "bf16_store": {"gups": round(g.updates / (statistics.mean(b16_ms) * 1e-3) / 1e9, 3),
                       "rel_rmse_vs_fp32": rel_rmse_bf},
        "interp_16bit_arithmetic": fast,
        "backproject_gups_fp16": round(g.updates / (bp_ms * 1e-3) / 1e9, 3),
        "roofline": {"bound": "hbm", "achieved": round(ach16, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(ach16 / HBM_PEAK_GBPS, 4),
                     "traffic": traffic,
                     "traffic_frac": (round(traffic / (bp_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
                                      if traffic else None),
                     "kernel": "backproject_lds_kernel<fp16, 32> (FP32 interpolation; LDS 48 "
                               "columns, pitch 51, 72 rows)",
                     "avg_ms": round(bp_ms, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
