"""FDK benchmark on one MI355X: 512^3 volume from 720 views of 768^2, plus the
serving stream.

A step is one full FDK pass over resident projections: FDK cosine weight +
ramp filter (FFT length 2048) into the staged layout, then the voxel-driven
cone backprojection of all 720 views.  Inputs are synthetic: 30 seeded random
ellipsoids rendered at 0.5 mm, projected by the ray-driven forward projector,
Poisson noise at I0 = 1e5 and log.  The headline value is the FP16-stored
form (filtered projections stored as FP16 value-and-step texels with a
per-view scale, FP32 interpolation and accumulation); the FP32-stored and
BF16-stored forms are timed beside it.

Roofline of the dominant kernel (the backprojection).  It is bound by VALU
issue, not HBM: its four bilinear taps per update come from LDS.  So
`roofline` is VALU issue: wave64 VALU instructions per launch (SQ_INSTS_VALU
per voxel-view update, measured by rocprofv3 and recorded in
profiles/backproject_fdk512_sq_counters.md, times the updates of one launch)
over the launch's duration (HIP events on its stream), against the issue
rate of MI355X's published 157.3 TFLOP/s FP32 vector peak: 157.3e12 / 2
FLOP per FMA / 64 lanes = 1229 G wave-instructions/s.  That published rate
counts packed FP32 (two FMAs per lane per instruction), which this build
does not use; beside it the line gives the rate a probe of the kernel's own
instruction mix reaches (probes/valu_issue, run on the same GPU after the
timing; its kernels are short and run at a higher clock than the
backprojector's, so the kernel can come close to it or pass it).  The HBM side is reported
too: bytes measured per launch (FETCH_SIZE + WRITE_SIZE, scaled by a copy of
known size, profiles/backproject_fdk512_traffic.md) over the same duration,
against 8 TB/s, and the tap bytes the LDS serves.

Also in the line: the forward projector on the same scan, the CPU baseline
(the oracle's restatement of the reference CPU path, the whole job on the
host cores) and the serving stream (Poisson arrivals over 256^3 / 512^3 /
1024^3 requests, each with its own seeded phantom, FP16 storage, 200 GB HBM
budget): p50 / p99 latency and aggregate GUPS.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

HBM_PEAK_GBPS = 8000.0
# MI355X published FP32 vector peak 157.3 TFLOP/s = 2 FLOP x 64 lanes x
# this many wave64 VALU instructions per second (packed FP32 counted as two)
VALU_PEAK_GINSTR = 157.3e3 / 2 / 64
WORKLOAD = "fdk_512"
FFT_LENGTH = 2048
# the larger FDK configurations `--workload` can time in place of the
# headline one: (FFT length, ellipsoids of the phantom, semi-axes in voxels,
# description); the headline line is always the 512^3 one
WORKLOADS = {
    "fdk_512": (2048, 30, (5, 120), "FDK cone 512x512x512 @0.5mm, 720 views/360deg, "
                                    "768x768 panel @0.8mm, SOD 500, SDD 1000, FFT 2048, Ram-Lak"),
    "fdk_1024_s": (2048, 60, (10, 240), "FDK cone 1024x1024x1024 @0.25mm, 1200 views/360deg, "
                                        "1024x1024 panel @0.6mm, SOD 500, SDD 1000, FFT 2048, "
                                        "Ram-Lak"),
    "fdk_1440": (4096, 60, (15, 340), "FDK cone 1440x1470x1440 @0.25mm, 1440 views/360deg, "
                                      "1536x1536 panel @0.4mm, SOD 500, SDD 1000, FFT 4096, "
                                      "Ram-Lak"),
}
PROFILES = os.path.join(ROOT, "profiles")


def profile_value(name, key):
    """A `key = value` line of a profiles/ summary (rocprofv3 counter
    figures the bench does not measure itself), or None."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            parts = line.strip().strip("`").split("=")
            if len(parts) == 2 and parts[0].strip() == key:
                return float(parts[1].strip())
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def probe_mix_ceiling():
    """The issue rate (wave64 VALU instructions per ns per SIMD, 5 waves per
    SIMD) that the probe probes/valu_issue reaches with the FP16 loop's
    instruction mix, run now as a child process on this GPU; None when the
    probe is not built or fails."""
    import subprocess
    exe = os.path.join(ROOT, "3dfilteredbackprojection-public_amd", "probes", "valu_issue")
    if not os.path.exists(exe):
        return None
    try:
        out = subprocess.run([exe], capture_output=True, text=True, timeout=60).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    for line in out.splitlines():
        if line.startswith("loop mix fp16"):
            return float(line.split()[4])
    return None


def make_inputs(ctr, g, dev, seed=20240602):
    """Phantom -> forward projection (timed: the forward-projector line) ->
    Poisson noise and log."""
    import torch
    _, n_ell, semi, _ = WORKLOADS[WORKLOAD]
    ell = ctr.random_ellipsoids(seed, n_ell, (g.X, g.Y, g.Z), g.voxel, semi)
    ell[:, 9] *= 0.02  # density 1.0 = 0.02 / mm
    vol = torch.zeros((g.Z, g.Y, g.X), dtype=torch.float32, device=dev)
    ctr.render_ellipsoids(ell, vol, g.volume_width)
    proj = ctr.forward_project(vol, g, step=1.0)  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ctr.forward_project(vol, g, step=1.0, out=proj)
    e1.record()
    e1.synchronize()
    fwd_ms = e0.elapsed_time(e1)
    ctr.poisson_log(proj, 1e5, seed)
    torch.cuda.synchronize()
    log(f"inputs ready; forward projection {fwd_ms:.1f} ms")
    del vol
    return proj, fwd_ms


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


def cpu_baseline(g, proj_cpu, scale, threads, full=True, budget_s=6.0):
    """The reference's CPU path as the oracle restates it: the row filter
    with the reference's radix-2 FFT algorithm (FP64, bit-exact with its
    compiled FFT) and the FP32 voxel-driven cone backprojection (FP64
    geometry).  full=True runs the whole job; otherwise a bounded sample of
    rows and z-slices, scaled to the job."""
    import oracle
    n_rows = g.nv * g.n_views
    rows = proj_cpu.reshape(-1, g.nu)
    sample_rows = n_rows if full else min(n_rows, 2048)
    t0 = time.perf_counter()
    q = oracle.filter_rows(rows[:sample_rows], oracle.RAM_LAK, FFT_LENGTH, threads=threads)
    t_filter = (time.perf_counter() - t0) * n_rows / sample_rows
    if full:
        q = q.astype(np.float32).reshape(proj_cpu.shape)
        t0 = time.perf_counter()
        oracle.backproject_cone_f32(q, g.source_distance, g.detector_distance,
                                    g.detector_width, g.volume_width, 1.0, 0.0, scale,
                                    (g.X, g.Y, g.Z), threads=threads)
        t_bp = time.perf_counter() - t0
        sample = (f"full job: all {n_rows} detector rows filtered (FFT {FFT_LENGTH}) and all "
                  f"{g.Z} z-slices x {g.n_views} views backprojected (filter {t_filter:.1f} s, "
                  f"backprojection {t_bp:.1f} s)")
    else:
        slices = 1
        while True:
            t0 = time.perf_counter()
            oracle.backproject_cone_f32(proj_cpu, g.source_distance, g.detector_distance,
                                        g.detector_width, g.volume_width, 1.0, 0.0, scale,
                                        (g.X, g.Y, g.Z), (g.Z // 2, g.Z // 2 + slices),
                                        threads=threads)
            dt = time.perf_counter() - t0
            if dt > budget_s / 3 or slices >= g.Z:
                break
            slices = min(g.Z, max(slices * 2, int(slices * budget_s / 3 / max(dt, 1e-3))))
        t_bp = dt * g.Z / slices
        sample = (f"sample scaled to the job: {sample_rows} of {n_rows} rows filtered and "
                  f"{slices} of {g.Z} z-slices backprojected (filter {t_filter:.1f} s, "
                  f"backprojection {t_bp:.1f} s extrapolated)")
    total = t_filter + t_bp
    return {"value": g.updates / total / 1e9, "unit": "GUPS", "cores": threads,
            "kind": "port", "seconds": round(total, 2), "sample": sample}


def run_serving(ctr, requests, rate, budget_gb, pack, max_slots, chunks):
    """Poisson stream of FDK requests over three sizes (256^3/360/384^2,
    512^3/720/768^2, 1024^3/1200/1024^2), each with its own seeded phantom,
    FP16 filtered projections, FP32 volumes: p50/p99 latency and aggregate
    GUPS."""
    import torch
    from ctrecon import serving
    classes = ["small", "medium", "large"]
    slots = dict(zip(classes, (int(x) for x in max_slots.split(","))))
    view_chunks = dict(zip(classes, (int(x) for x in chunks.split(","))))
    srv = serving.Server(budget_bytes=budget_gb * 1e9, dtype=ctr.F16, pack=pack,
                         max_slots=slots, chunks=view_chunks, log=log)
    srv.warm(classes)
    torch.cuda.synchronize()
    reqs = serving.poisson_arrivals(requests, rate, seed=7)
    log(f"serving {requests} requests at {rate}/s")
    makespan = srv.run(reqs)
    s = serving.summarize(reqs, makespan, srv.geoms)
    return {
        "metric": "FDK serving latency, Poisson arrivals",
        "rate_rps": rate,
        "requests": requests,
        "latency_ms": {"p50": round(s["p50_ms"], 2), "p99": round(s["p99_ms"], 2),
                       "mean": round(s["mean_ms"], 2)},
        "aggregate_gups": round(s["aggregate_gups"], 2),
        "throughput_rps": round(s["throughput_rps"], 3),
        "makespan_s": round(makespan, 2),
        "per_class": {c: {k: round(v, 2) if isinstance(v, float) else v for k, v in d.items()}
                      for c, d in s["per_class"].items()},
        "hbm_budget_gb": budget_gb,
        "slots": {c: sum(1 for x in srv.slots if x.cls == c) for c in classes},
        "slot_gb": round(srv.used / 1e9, 2),
        "pack": pack,
        "view_chunks": view_chunks,
        "config": "sizes uniform over 256^3/360 views/384^2, 512^3/720/768^2, "
                  "1024^3/1200/1024^2; FP16 store, FP32 volume, FFT 2 nu",
        "data": "each request its own seeded ellipsoid phantom; its exact projections are "
                "generated on the device at admission (counted in its latency)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--mode", choices=["fdk", "serving", "copy-probe"], default="fdk")
    ap.add_argument("--workload", choices=list(WORKLOADS), default="fdk_512",
                    help="FDK configuration to time (the headline is fdk_512; the others "
                         "print the same line without the CPU baseline, the serving stream "
                         "and the counter-based figures, which are the 512^3 scan's)")
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--pack", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    # one 1024^3 slot: a 1024^3 reconstruction fills the GPU for ~0.35 s, and
    # several in flight delay every smaller request behind them
    ap.add_argument("--max-slots", default="2,6,1",
                    help="serving slots per class: small,medium,large")
    # a 1024^3 reconstruction in 8 accumulating launches of 150 views: other
    # slots' work waits for one ~40 ms chunk instead of the whole ~0.3 s
    ap.add_argument("--chunks", default="1,1,8",
                    help="serving view chunks per reconstruction: small,medium,large")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(the reconstruction would shard into z-slabs with no exchange; "
              "a slab is the same kernel with fewer y-runs)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "copy-probe":
        # a plain elementwise kernel that reads 1 GiB and writes 1 GiB, the
        # reference for scaling the HBM byte counters (FETCH_SIZE /
        # WRITE_SIZE) of the profiled kernels
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
    if args.mode == "serving":
        print(json.dumps(run_serving(ctr, args.requests, args.rate, args.budget_gb, args.pack,
                                     args.max_slots, args.chunks)), flush=True)
        return

    global WORKLOAD, FFT_LENGTH
    WORKLOAD = args.workload
    FFT_LENGTH = WORKLOADS[WORKLOAD][0]
    headline = WORKLOAD == "fdk_512"
    if not headline:
        args.no_cpu_baseline = args.no_serving = True
    dev = torch.device("cuda:0")
    g = ctr.config_geometry(WORKLOAD)
    scale = g.fbp_scale(FFT_LENGTH)
    ws = ctr.Workspace()
    ws.reserve(g.n_views, FFT_LENGTH)
    log(f"workload {WORKLOAD}: {g.X}x{g.Y}x{g.Z} from {g.n_views} views of "
        f"{g.nu}x{g.nv}, FFT {FFT_LENGTH}")
    proj, fwd_ms = make_inputs(ctr, g, dev)

    log("timing FP32-stored form")
    f32_ms, f32_bp, vol32 = time_fdk(ctr, g, proj, ctr.F32, args.steps, args.warmup, ws, scale)
    log(f"  step {statistics.median(f32_ms):.2f} ms, backprojection {statistics.median(f32_bp):.2f} ms")
    ref32 = vol32.clone()
    log("timing FP16-stored form")
    f16_ms, f16_bp, vol16 = time_fdk(ctr, g, proj, ctr.F16, args.steps, args.warmup, ws, scale)
    log(f"  step {statistics.median(f16_ms):.2f} ms, backprojection {statistics.median(f16_bp):.2f} ms")
    rel_rmse_16 = float((vol16.double() - ref32.double()).norm() / ref32.double().norm())
    del vol16
    log("timing BF16-stored form")
    b16_ms, b16_bp, volb = time_fdk(ctr, g, proj, ctr.BF16, max(3, args.steps // 2), 1, ws, scale)
    rel_rmse_bf = float((volb.double() - ref32.double()).norm() / ref32.double().norm())
    del volb
    del ref32, vol32
    torch.cuda.synchronize()

    # upload of the projections from pinned host memory: the rate with the
    # host handing over its projections (the drop-in classes' textures load
    # from the host) is the step plus this copy; `value` excludes it
    host = torch.empty(proj.shape, dtype=proj.dtype, pin_memory=True)
    host.copy_(proj)
    h2d = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        proj.copy_(host, non_blocking=True)
        e1.record()
        e1.synchronize()
        h2d.append(e0.elapsed_time(e1))
    del host
    h2d_ms = statistics.median(h2d)
    proj_bytes = proj.numel() * 4
    h2d_gbps = proj_bytes / (h2d_ms * 1e-3) / 1e9

    ms = statistics.mean(f16_ms)
    gups = g.updates / (ms * 1e-3) / 1e9
    gups32 = g.updates / (statistics.mean(f32_ms) * 1e-3) / 1e9
    bp_ms = statistics.mean(f16_bp)
    bp_ms32 = statistics.mean(f32_bp)

    def valu(form, t_ms):
        if not headline:
            return None, None
        per_update = profile_value("backproject_fdk512_sq_counters.md",
                                   f"{form}_valu_per_update")
        if per_update is None:
            return None, None
        ginstr = per_update * g.updates / 64 / (t_ms * 1e-3) / 1e9
        return per_update, ginstr

    v16, achieved16 = valu("fp16", bp_ms)
    v32, achieved32 = valu("fp32", bp_ms32)
    # the issue rate a probe of the fp16 loop's own instruction mix reaches
    # (wave64 VALU instructions per ns per SIMD, independent chains): on this
    # GPU now, else the figure recorded with the counters
    mix_ceiling, mix_source = probe_mix_ceiling(), "probe run on this GPU"
    if mix_ceiling is None:
        mix_ceiling = profile_value("backproject_fdk512_sq_counters.md",
                                    "fp16_mix_issue_ceiling_per_ns")
        mix_source = "profiles/backproject_fdk512_sq_counters.md"
    traffic16 = traffic32 = fwd_beyond_l2 = None
    if headline:
        traffic16 = profile_value("backproject_fdk512_traffic.md", "fp16_bytes_per_launch")
        traffic32 = profile_value("backproject_fdk512_traffic.md", "fp32_bytes_per_launch")
        fwd_beyond_l2 = profile_value("forward_fdk512_counters.md", "slab_beyond_l2_bytes")
    tap_bytes16 = (4 * 2 + 8.0 / g.n_views) * g.updates
    compulsory = ctr.staged_layout(g.nu, g.nv, g.n_views)[2] + 4.0 * g.X * g.Y * g.Z

    # forward projector on the same scan: rays and trilinear samples per
    # second; bytes from beyond L2 against reading the volume once per view
    rays = g.nu * g.nv * g.n_views
    forward = {"kernel": "forward_slab_kernel (step 1 voxel)", "ms": round(fwd_ms, 2),
               "grays_per_s": round(rays / (fwd_ms * 1e-3) / 1e9, 3),
               "gsamples_per_s": round(rays * g.X / (fwd_ms * 1e-3) / 1e9, 1),
               "beyond_l2_bytes": fwd_beyond_l2,
               "volume_once_per_view_bytes": 4.0 * g.X * g.Y * g.Z * g.n_views,
               "beyond_l2_gbps": (round(fwd_beyond_l2 / (fwd_ms * 1e-3) / 1e9, 1)
                                  if fwd_beyond_l2 else None)}

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        proj_cpu = proj.cpu().numpy()
        log(f"CPU baseline, whole job on {threads} threads")
        cpu = cpu_baseline(g, proj_cpu, scale, threads, full=True)
        log(f"  {cpu['value']:.3f} GUPS ({cpu['seconds']} s)")
        if threads > 1:
            log("CPU baseline on 1 thread (sampled: the whole job would take ~7 min)")
            one = cpu_baseline(g, proj_cpu, scale, 1, full=False)
            cpu["single_core"] = {"value": one["value"], "unit": "GUPS", "cores": 1,
                                  "sample": one["sample"]}
            log(f"  {one['value']:.4f} GUPS")
        del proj_cpu
    del proj
    torch.cuda.empty_cache()

    serving_line = None
    if not args.no_serving:
        serving_line = run_serving(ctr, args.requests, args.rate, args.budget_gb, args.pack,
                                   args.max_slots, args.chunks)
        log(f"serving p50 {serving_line['latency_ms']['p50']} ms, "
            f"p99 {serving_line['latency_ms']['p99']} ms")

    out = {
        "metric": (f"FDK GUPS {g.X}x{g.Y}x{g.Z} x {g.n_views} views (FP16-stored filtered "
                   "projections, FP32 interpolation and accumulation)"
                   if not headline else
                   "FDK GUPS 512^3 x 720 views (FP16-stored filtered projections, FP32 "
                   "interpolation and accumulation)"),
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
        "data": f"synthetic: {WORKLOADS[WORKLOAD][1]} seeded ellipsoids, ray-driven "
                "projections, Poisson I0=1e5",
        "config": {"workload": WORKLOADS[WORKLOAD][3],
                   "storage": "fp16 value-and-step texels + per-view scale"},
        "latency_ms": {"p50": round(statistics.median(f16_ms), 3),
                       "p99": round(float(np.percentile(f16_ms, 99)), 3)},
        "fp32_store": {"gups": round(gups32, 3), "ms_per_step": round(statistics.mean(f32_ms), 3),
                       "backproject_ms": round(bp_ms32, 3),
                       "valu_ginstr_per_s": round(achieved32, 1) if achieved32 else None,
                       "valu_frac": round(achieved32 / VALU_PEAK_GINSTR, 4) if achieved32 else None,
                       "hbm_bytes_per_launch": traffic32,
                       "hbm_frac": (round(traffic32 / (bp_ms32 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
                                    if traffic32 else None)},
        "fp16_vs_fp32_gups": round(gups / gups32, 3),
        "rel_rmse_fp16_vs_fp32": rel_rmse_16,
        "bf16_store": {"gups": round(g.updates / (statistics.mean(b16_ms) * 1e-3) / 1e9, 3),
                       "backproject_ms": round(statistics.mean(b16_bp), 3),
                       "rel_rmse_vs_fp32": rel_rmse_bf},
        "backproject_gups_fp16": round(g.updates / (bp_ms * 1e-3) / 1e9, 3),
        "pcie_inclusive": {"h2d_ms": round(h2d_ms, 3),
                           "h2d_gb_per_s": round(h2d_gbps, 1),
                           "gups": round(g.updates / ((ms + h2d_ms) * 1e-3) / 1e9, 3),
                           "note": f"pinned host -> HBM copy of the "
                                   f"{proj_bytes / 1e9:.2f} GB projections plus the FP16 step"},
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved16, 1) if achieved16 else None,
            "peak": round(VALU_PEAK_GINSTR, 1),
            "unit": "G wave64 VALU instructions/s",
            "frac": round(achieved16 / VALU_PEAK_GINSTR, 4) if achieved16 else None,
            "issue_per_ns_per_simd": (round(achieved16 / 1024, 3) if achieved16 else None),
            "mix_ceiling_per_ns_per_simd": mix_ceiling,
            "mix_ceiling_source": mix_source if mix_ceiling else None,
            "frac_of_mix_ceiling": (round(achieved16 / 1024 / mix_ceiling, 4)
                                    if achieved16 and mix_ceiling else None),
            "valu_per_update": v16,
            "kernel": "backproject_lds_kernel<fp16, 32> (value-and-step texels, FP32 "
                      "interpolation; LDS 48 columns, pitch 51, 72 rows)",
            "avg_ms": round(bp_ms, 3),
            "traffic": traffic16,
            "hbm_gbps": round(traffic16 / (bp_ms * 1e-3) / 1e9, 1) if traffic16 else None,
            "hbm_frac": (round(traffic16 / (bp_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
                         if traffic16 else None),
            "compulsory_bytes": compulsory,
            "lds_served_tap_bytes": tap_bytes16,
        },
        "forward_projector": forward,
        "cpu_baseline": cpu,
        "serving": serving_line,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
