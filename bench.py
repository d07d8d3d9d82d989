#!/usr/bin/env python3
"""Benchmark of the reconstruction hot path on one MI355X.

Workload (one step): 128x128 confocal scan on a 2.0 m wall, 1024 bins of
16 ps -> 128^3 voxel backprojection, then the separable LoG (sigma = 1,
k = 5) and min-max normalisation, i.e. Laser::reconstructShapeAABB without
the file dump.  The scene is 4 seeded random planar quads sampled at 20k
surface points, rendered with 1/r^4 falloff and Poisson noise at a peak of
1000 counts; inputs are resident in HBM before timing starts.

Prints ONE JSON line.  value = voxel-updates/s of the whole step
(X*Y*Z*L updates per step).  roofline = the backprojection's algorithmic
FP32 rate (16 flops per voxel-update) against the MI355X FP32 peak, plus
the aligned-grid kernel's own bound (LDS bytes, 4 per update) and the HBM
traffic per launch from the committed rocprofv3 counter profile.
cpu_baseline = the FP64 OpenMP oracle on a bounded slab of the same
workload, run on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FLOPS_PER_UPDATE = 16          # confocal, FMA = 2 (reference formula)
PEAK_FP32_TFLOPS = 157.3       # MI355X FP32 (datasheet)
PEAK_LDS_TBPS = 78.6           # 256 CUs x 128 B/clk x 2.4 GHz
PEAK_HBM_GBPS = 8000.0         # MI355X HBM3E (datasheet)
PEAK_BF16_TFLOPS = 2500.0      # MI355X dense BF16 MFMA (datasheet, no sparsity)
VALU_LANE_SLOTS = 39.3e12      # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (non-packed issue)
# Issue slots per update: (VALU instructions - transcendentals) + 2.37 x
# transcendentals, from the counter profile of the kernels
# (profiles/counters_confocal_exhaustive_lct_bench.md; 2.37 = measured
# v_fma / v_rsq throughput ratio on this GPU).
GENERAL_SLOTS = 14.9           # confocal_voxel_kernel
EXH_SLOTS = 9.38               # exhaustive_group_kernel (profiles/exhaustive_group_c3_bench_counters.md)
DEPTH_EXH_SLOTS = 10.91        # depth_group_kernel (profiles/exhaustive_group_c3_bench_counters.md)
TRAFFIC_PROFILE = ROOT / "profiles" / "confocal_128cube_traffic.json"


def traffic_bytes(aligned):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a plain copy)."""
    try:
        prof = json.loads(TRAFFIC_PROFILE.read_text())
    except (OSError, ValueError):
        return None
    key = "aligned_conv_kernel" if aligned else "confocal_voxel_kernel"
    k = prof.get("kernels", {}).get(key)
    return None if k is None else k["hbm_bytes_per_launch"]
METRIC = "voxel-updates/s + %roofline; learned vs exhaustive BP time at rel-L2<=2%"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU time of the oracle sample")
    ap.add_argument("--res", type=int, default=128)
    ap.add_argument("--scan", type=int, default=128)
    ap.add_argument("--bins", type=int, default=1024)
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the FP16/BF16 storage leg (256^2 x 1024 -> 128^3)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-vs-plain leg (64^2 x 512 -> 64^3)")
    ap.add_argument("--copy-reference-mib", type=int, default=0,
                    help="run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--no-lct", action="store_true",
                    help="skip the LCT leg (128^2 x 1024 padded to 256x256x2048)")
    ap.add_argument("--no-depth", action="store_true",
                    help="skip the depth-plane mode + slice images leg")
    ap.add_argument("--no-splat", action="store_true",
                    help="skip the scatter-form (splat) backprojection leg")
    ap.add_argument("--no-mis", action="store_true",
                    help="skip the MIS measurement on the main capture")
    ap.add_argument("--no-filters", action="store_true",
                    help="skip the filter leg (Laplacian, LoG, normalise, LoGFFT, filter_H)")
    ap.add_argument("--no-exhaustive", action="store_true",
                    help="skip the exhaustive leg (16^2 lasers x 64^2 sensors x 1024 -> 64^3)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from fasterbackprojection_amd.api import Context, make_info
    from fasterbackprojection_amd.scene import full_confocal_geometry, random_quads

    torch.manual_seed(0)
    dev = torch.device("cuda", 0)
    g = full_confocal_geometry(args.scan, args.bins, args.res)
    info = make_info(g)
    L = len(g.laser_targets)
    X = Y = Z = args.res
    updates = X * Y * Z * L

    ctx = Context(0)
    rng = np.random.default_rng(0)
    scene = random_quads(rng, 4, 20000)
    laser = torch.from_numpy(g.laser_targets).to(dev)
    pts = torch.from_numpy(scene.points.astype(np.float32)).to(dev)
    alb = torch.from_numpy(scene.albedo.astype(np.float32)).to(dev)
    clean = ctx.render_confocal(laser, pts, alb, g.num_time_bins, g.time_step)
    lam = clean * (1000.0 / clean.max())
    I = torch.poisson(lam).contiguous()
    torch.cuda.synchronize()
    log(f"workload: {L} scan points x {g.num_time_bins} bins -> {X}x{Y}x{Z}, "
        f"{updates:.3e} voxel-updates/step")

    vol = torch.empty((Z, Y, X), dtype=torch.float32, device=dev)
    filt = torch.empty_like(vol)

    def step():
        ctx.backproject_confocal(info, I, laser, out=vol)
        ctx.log_normalize(vol, 5, 1.0, out=filt)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("warmup done")

    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = e0.elapsed_time(e1) / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step (wall {wall:.2f} s)")

    # Dominant kernel: backprojection alone, events on the launch stream.
    def time_bp(reps):
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            ctx.backproject_confocal(info, I, laser, out=vol)
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.mean(ms))

    bp_avg = time_bp(max(3, args.steps))
    aligned = ctx.last_confocal_path_aligned()
    f_ms = []
    for _ in range(max(3, args.steps)):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        ctx.log_normalize(vol, 5, 1.0, out=filt)
        b.record(stream)
        b.synchronize()
        f_ms.append(a.elapsed_time(b))
    achieved = FLOPS_PER_UPDATE * updates / (bp_avg * 1e-3) / 1e12
    log(f"backprojection {bp_avg:.3f} ms ({'aligned-grid path' if aligned else 'general kernel'}) "
        f"-> {achieved:.1f} TFLOP/s algorithmic; LoG+normalise {np.mean(f_ms):.3f} ms")
    # The general kernel on the same input, for comparison.
    ctx.set_aligned_path(False)
    general_ms = time_bp(3)
    ctx.set_aligned_path(True)
    log(f"general kernel on the same input: {general_ms:.3f} ms")
    mis = None if args.no_mis else mis_measure(ctx, info, I, laser, vol, stream)
    splat = None if args.no_splat else splat_measure(ctx, info, I, laser, vol, stream, general_ms,
                                                     bp_avg)
    # End to end from host memory: pinned transient -> device, one step,
    # normalised volume -> pinned host buffer.
    I_host = torch.empty(I.shape, dtype=I.dtype, pin_memory=True)
    I_host.copy_(I)
    vol_host = torch.empty(filt.shape, dtype=filt.dtype, pin_memory=True)
    I_dev = torch.empty_like(I)
    e2e = []
    for _ in range(3):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        I_dev.copy_(I_host, non_blocking=True)
        ctx.backproject_confocal(info, I_dev, laser, out=vol)
        ctx.log_normalize(vol, 5, 1.0, out=filt)
        vol_host.copy_(filt, non_blocking=True)
        b.record(stream)
        b.synchronize()
        e2e.append(a.elapsed_time(b))
    e2e_ms = float(np.median(e2e))
    del I_host, vol_host, I_dev
    log(f"end to end with host transfers: {e2e_ms:.3f} ms")
    if args.copy_reference_mib:
        # Plain device copy of a known size: calibrates FETCH_SIZE / WRITE_SIZE
        # when this run is profiled with rocprofv3 --pmc.
        n = args.copy_reference_mib << 18
        src = torch.ones(n, dtype=torch.float32, device=dev)
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    # The aligned path's first call on a fresh context builds its offset
    # tables (geometry only, kept across calls): cold and warm call.
    ctx_cold = Context(0)
    cold = []
    for _ in range(2):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        ctx_cold.backproject_confocal(info, I, laser, out=vol)
        b.record(stream)
        b.synchronize()
        cold.append(a.elapsed_time(b))
    ctx_cold.close()
    log(f"fresh context: first call {cold[0]:.3f} ms (table build), second {cold[1]:.3f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        log(f"cpu legs use the oracle built for this host: {oracle.use_native()}")
        cpu = cpu_baseline(info, I, g, args.cpu_seconds)
    precision = None if args.no_precision else precision_leg(ctx, args)
    learned = None if args.no_learned else learned_leg(ctx)
    exhaustive = None if args.no_exhaustive else exhaustive_leg(ctx, args)
    lct = None if args.no_lct else lct_leg(ctx, args)
    depth = None if args.no_depth else depth_leg(ctx, args)
    filters = None if args.no_filters else filters_leg(ctx, args)

    out = {
        "metric": METRIC,
        "value": updates / (ms_per_step * 1e-3),
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 4 seeded random quads, 20k samples, 1/r^4, Poisson peak 1000",
        "config": {
            "workload": f"confocal {args.scan}x{args.scan}x{args.bins} (16 ps) -> "
                        f"{X}^3 backprojection + LoG(sigma=1,k=5) + min-max normalise",
            "scan_points": L, "time_bins": args.bins, "voxels": [X, Y, Z],
            "voxel_updates_per_step": updates,
        },
        "roofline": {
            "bound": "valu",
            "kernel": ("aligned_conv_kernel (fbp_backproject_confocal_voxel, aligned-grid path)"
                       if aligned else "confocal_voxel_kernel (fbp_backproject_confocal_voxel)"),
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic_bytes(aligned),
            "flops_per_update": FLOPS_PER_UPDATE,
            "kernel_ms": bp_avg,
            # The aligned path does one LDS read and one FMA per update, so
            # its own binding resource is LDS bandwidth (4 B per update).
            "lds": {"achieved": 4.0 * updates / (bp_avg * 1e-3) / 1e12, "peak": PEAK_LDS_TBPS,
                    "unit": "TB/s", "frac": 4.0 * updates / (bp_avg * 1e-3) / 1e12 / PEAK_LDS_TBPS},
        },
        "breakdown_ms": {"backprojection": bp_avg, "log_and_normalize": float(np.mean(f_ms)),
                         "backprojection_general_kernel": general_ms,
                         "backprojection_fresh_context_first_call": cold[0],
                         "backprojection_fresh_context_second_call": cold[1],
                         "end_to_end_with_h2d_d2h": e2e_ms},
        # Issue-slot view (non-packed FP32: 39.3e12 lane-slots/s; a
        # transcendental takes 2.37 slots, measured).  Aligned path: 1.27 FMA
        # slots and 1.27 LDS reads per useful update (LDS, at 2 clocks per
        # wave-read per CU, binds).  General kernel: 14.9 slots per update
        # measured by counters (13.5 VALU instructions, 1.0 of them v_rsq).
        "issue_slot_diagnostic": {
            "aligned_conv_kernel": {
                "lds_bound_updates_per_s": PEAK_LDS_TBPS * 1e12 / 4.0 / 1.27,
                "frac": updates / (bp_avg * 1e-3) / (PEAK_LDS_TBPS * 1e12 / 4.0 / 1.27)},
            "confocal_voxel_kernel": {
                "valu_slots_per_update": GENERAL_SLOTS,
                "bound_updates_per_s": VALU_LANE_SLOTS / GENERAL_SLOTS,
                "frac": updates / (general_ms * 1e-3) / (VALU_LANE_SLOTS / GENERAL_SLOTS)}},
        "cpu_baseline": cpu,
        "precision": precision,
        "learned": learned,
        "exhaustive": exhaustive,
        "lct": lct,
        "depth_planes": depth,
        "filters": filters,
        "mis": mis,
        "splat": splat,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def splat_measure(ctx, info, I, laser, vol, stream, general_ms, bp_ms, reps=3):
    """Scatter form (fbp_backproject_confocal_splat, the mathematical form of
    DrawEllipsoids::solveBackprojection) beside the gather paths: on the
    main capture, on the same capture keeping only its largest 4 % of
    samples (a sparse transient), and on the 32^2 x 256 capture of 200
    points -> 32^3.  Event-timed medians; rel-L2 against the general gather
    kernel's volume."""
    import numpy as np
    import torch

    from fasterbackprojection_amd.api import make_info
    from fasterbackprojection_amd.scene import small_confocal

    def timed(fn):
        fn()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.median(ms))

    def case(inf, tr, las, general=None):
        r = inf.voxel_resolution
        upd = r.x * r.y * r.z * inf.num_laser_targets
        out = torch.empty((r.z, r.y, r.x), dtype=torch.float32, device=tr.device)
        t_s = timed(lambda: ctx.backproject_confocal_splat(inf, tr, las, out=out))
        ref = torch.empty_like(out)
        ctx.set_aligned_path(False)
        try:
            t_g = general if general is not None else timed(
                lambda: ctx.backproject_confocal(inf, tr, las, out=ref))
            ctx.backproject_confocal(inf, tr, las, out=ref)
        finally:
            ctx.set_aligned_path(True)
        rel = float((out - ref).norm() / ref.norm())
        return {"splat_ms": t_s, "gather_general_ms": t_g,
                "lit_fraction": float((tr > 1e-5).float().mean()),
                "voxel_updates_per_s": upd / (t_s * 1e-3), "rel_l2_vs_gather": rel}

    res = {"kernel": "splat_kernel (fbp_backproject_confocal_splat)"}
    res["capture_128x128x1024_to_128cube"] = case(info, I, laser, general_ms)
    res["capture_128x128x1024_to_128cube"]["gather_aligned_ms"] = bp_ms
    thr = torch.quantile(I.flatten()[::97].float(), 0.96)
    sparse = torch.where(I > thr, I, torch.zeros_like(I)).contiguous()
    res["capture_128x128x1024_to_128cube_largest_4pct"] = case(info, sparse, laser)
    cap = small_confocal(seed=0)
    info1 = make_info(cap.geometry)
    I1 = torch.from_numpy(cap.transient).to(I.device)
    l1 = torch.from_numpy(cap.geometry.laser_targets).to(I.device)
    res["capture_32x32x256_to_32cube"] = case(info1, I1, l1)
    for k, v in res.items():
        if isinstance(v, dict):
            log(f"splat {k}: {v['splat_ms']:.3f} ms (gather general {v['gather_general_ms']:.3f} "
                f"ms), lit {v['lit_fraction']:.3f}, rel-L2 {v['rel_l2_vs_gather']:.1e}")
    return res


def mis_measure(ctx, info, I, laser, vol, stream, reps=3):
    """MIS-sampled backprojection of the main capture (128^2 x 1024 -> 128^3):
    REFERENCE mode (the reference's per-voxel sampling and weights, L/16
    samples per voxel) and the tile-shared mode (L/16 samples per 16^3 tile
    through the general kernel).  Per sample the confocal chain (16 flops)
    is evaluated; the VALU roofline is priced on that count."""
    import numpy as np
    import torch

    from fasterbackprojection_amd.api import MIS_REFERENCE, MIS_TILED

    dev = torch.device("cuda", 0)
    L = info.num_laser_targets
    V = info.voxel_resolution.x * info.voxel_resolution.y * info.voxel_resolution.z
    gen = torch.Generator(device=dev)
    gen.manual_seed(21)
    noise = torch.rand(1 << 20, device=dev, generator=gen)
    n = L // 16
    res = {"workload": "128^2 x 1024 confocal -> 128^3, L/16 = %d samples per voxel (reference "
                       "mode) or per 16^3 tile (tile-shared mode)" % n}
    for name, mode in (("reference_mode", MIS_REFERENCE), ("tile_shared_mode", MIS_TILED)):
        ctx.backproject_confocal_mis(info, I, laser, noise, n, mode=mode, out=vol)
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            ctx.backproject_confocal_mis(info, I, laser, noise, n, mode=mode, out=vol)
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        t = float(np.median(ms))
        samples = V * n
        tf = FLOPS_PER_UPDATE * samples / (t * 1e-3) / 1e12
        res[name] = {"ms": t, "sample_updates": samples,
                     "sample_updates_per_s": samples / (t * 1e-3),
                     "roofline": {"bound": "valu", "achieved": tf, "peak": PEAK_FP32_TFLOPS,
                                  "unit": "TFLOP/s", "frac": tf / PEAK_FP32_TFLOPS,
                                  "flops_per_update": FLOPS_PER_UPDATE}}
    log(f"mis: {res}")
    return res


def filters_leg(ctx, args, reps=5):
    """The post-filters on a 128^3 volume (Laplacian k=5, separable LoG
    sigma=1 k=5, min-max normalise, LoGFFT sigma=1) and filter_H on the
    128^2 x 1024 transient (wl = 10 bins), each against the HBM roofline
    with its algorithmic bytes: read V*4 + write V*4 per filter; normalise
    reads twice (min/max, apply) and writes once; LoGFFT two FFTs of
    (read + write) plus the spectrum multiply over the half spectrum;
    filter_H two FFTs x (read + write) x 8 bytes x L x nt_pad."""
    import numpy as np
    import torch

    dev = torch.device("cuda", 0)
    n = 128
    gen = torch.Generator(device=dev)
    gen.manual_seed(9)
    vol = torch.rand((n, n, n), device=dev, generator=gen)
    out = torch.empty_like(vol)
    V = n ** 3

    def timed(fn, setup=None):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            if setup:
                setup()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.median(ms))

    def entry(ms, nbytes):
        gbps = nbytes / (ms * 1e-3) / 1e9
        return {"ms": ms, "roofline": {"bound": "hbm", "achieved": gbps, "peak": PEAK_HBM_GBPS,
                                       "unit": "GB/s", "frac": gbps / PEAK_HBM_GBPS,
                                       "algorithmic_bytes": nbytes}}

    res = {"workload": "128^3 volume (Laplacian, LoG, normalise, LoGFFT); "
                       "128^2 x 1024 transient at 16 ps (filter_H, wl = 10 bins)"}
    res["laplacian_k5"] = entry(timed(lambda: ctx.laplacian(vol, 5, out=out)), 8 * V)
    res["log_k5_sigma1"] = entry(timed(lambda: ctx.log(vol, 5, 1.0, out=out)), 8 * V)
    work = vol.clone()
    res["normalize"] = entry(timed(lambda: ctx.normalize(work),
                                   setup=lambda: work.copy_(vol)), 12 * V)
    # LoG + normalisation in two launches (the LoG kernel reduces its own
    # min / max): read V*4, write V*4, normalise read + write V*8.
    res["log_normalize_k5_sigma1"] = entry(timed(lambda: ctx.log_normalize(vol, 5, 1.0, out=out)),
                                           16 * V)
    half = n * n * (n // 2 + 1)
    res["log_fft_sigma1"] = entry(timed(lambda: ctx.log_fft(work),
                                        setup=lambda: work.copy_(vol)),
                                  2 * (4 * V + 8 * half) + 16 * half)
    L, T, dt = 128 * 128, 1024, 0.0048
    I = torch.rand((L, T), device=dev, generator=gen)
    Iw = I.clone()
    wl = 10 * dt
    sigma = wl / np.sqrt(2.0)
    t6 = int(round(6 * sigma / dt))
    t6 += t6 % 2
    nt_pad = T + 2 * (t6 - 1)
    res["filter_h"] = entry(timed(lambda: ctx.filter_h(Iw, dt, wl), setup=lambda: Iw.copy_(I)),
                            2 * 8 * L * nt_pad * 2)
    res["filter_h"]["nt_pad"] = nt_pad
    log(f"filters: {res}")
    del vol, out, work, I, Iw
    torch.cuda.empty_cache()
    return res


def precision_leg(ctx, args):
    """256x256 confocal scan, 1024 bins of 16 ps, 8 seeded quads, Poisson at
    peak 500 -> 128^3.  FP32, FP16 and BF16 storage (per-scan-point max
    scale, Kahan accumulation, FP64 chunk reduction); errors against the FP64
    oracle on the first depth slices."""
    import numpy as np
    import torch

    import oracle
    from fasterbackprojection_amd.api import STORAGE_BF16, STORAGE_F16, STORAGE_F32, make_info
    from fasterbackprojection_amd.scene import full_confocal_geometry, random_quads

    dev = torch.device("cuda", 0)
    g = full_confocal_geometry(256, 1024, 128)
    info = make_info(g)
    L = len(g.laser_targets)
    rng = np.random.default_rng(1)
    scene = random_quads(rng, 8, 20000)
    laser = torch.from_numpy(g.laser_targets).to(dev)
    pts = torch.from_numpy(scene.points.astype(np.float32)).to(dev)
    alb = torch.from_numpy(scene.albedo.astype(np.float32)).to(dev)
    clean = ctx.render_confocal(laser, pts, alb, 1024, g.time_step)
    I = torch.poisson(clean * (500.0 / clean.max())).contiguous()
    torch.cuda.synchronize()
    nz = 2
    sub = type(info).from_buffer_copy(info)
    sub.voxel_resolution.z = nz
    t = time.perf_counter()
    ref = oracle.confocal_voxel(sub, I.cpu().numpy(), g.laser_targets)
    cpu_s = time.perf_counter() - t
    log(f"precision leg: FP64 oracle on {nz} slices in {cpu_s:.1f} s")
    out = {"workload": "confocal 256x256x1024 (16 ps, 8 quads, Poisson peak 500) -> 128^3",
           "oracle_slices": nz,
           "cpu": {"sample": f"FP64 OpenMP oracle on the first {nz} of 128 depth slices",
                   "voxel_updates_per_s": 128 * 128 * nz * L / cpu_s,
                   "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                   "kind": "port"}}
    stream = torch.cuda.current_stream()
    for name, st in (("f32", STORAGE_F32), ("f16", STORAGE_F16), ("bf16", STORAGE_BF16)):
        if st == STORAGE_F32:
            data, sc = I, None
        else:
            data, sc = ctx.quantize(I, st)
        vol = ctx.backproject_confocal(info, data, laser, storage=st, scales=sc)
        ms = []
        for _ in range(3):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            ctx.backproject_confocal(info, data, laser, storage=st, scales=sc, out=vol)
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        got = vol[:nz].cpu().numpy().astype(np.float64)
        err = got - ref
        out[name] = {"ms": float(np.mean(ms)),
                     "path": "aligned" if ctx.last_confocal_path_aligned() else "general",
                     "voxel_updates_per_s": 128 ** 3 * L / (np.mean(ms) * 1e-3),
                     "rel_l2": float(np.linalg.norm(err) / np.linalg.norm(ref)),
                     "max_abs": float(np.abs(err).max()),
                     "max_abs_rel_to_peak": float(np.abs(err).max() / np.abs(ref).max())}
        log(f"precision {name}: {out[name]}")
    return out


def lct_leg(ctx, args, reps=5):
    """LCT of the full confocal capture (128x128 scan, 1024 bins of 16 ps on a
    2 m wall) zero-padded to 256x256x2048, Wiener SNR 8 (python/lct.py
    semantics).  HBM-bound: algorithmic bytes = 2 FFTs x (read + write) x
    8 |P| + filter read 8 |P| for the padded complex volume |P| (5.4 GB).
    CPU baseline: the NumPy/SciPy restatement of python/lct.py
    (oracle/lct.py, scipy.fft with the host's worker threads) on the same
    capture and on the small configuration (32x32 scan, 256 bins of 32 ps,
    pad 64x64x512), with the inverse filter built once beforehand on both
    sides (the GPU caches it): the full size is timed once (it takes seconds
    and ~10 GB), the small one as the median of 3 after a warm-up run."""
    import numpy as np
    import torch

    from fasterbackprojection_amd.scene import (C_LIGHT, full_confocal_geometry, random_quads,
                                                small_confocal)

    dev = torch.device("cuda", 0)
    g = full_confocal_geometry(128, 1024, 128)
    sc = random_quads(np.random.default_rng(2), 4, 20000)
    laser = torch.from_numpy(g.laser_targets).to(dev)
    clean = ctx.render_confocal(laser, torch.from_numpy(sc.points.astype(np.float32)).to(dev),
                                torch.from_numpy(sc.albedo.astype(np.float32)).to(dev), 1024,
                                g.time_step)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2)
    I = torch.poisson(clean * (1000.0 / clean.max()), generator=gen)
    # scan order l = iz * n + ix -> data[iz][ix][t] = (N, N, M)
    data = I.reshape(128, 128, 1024).contiguous()
    width, rmax = 1.0, 1024 * float(g.time_step)
    out = ctx.lct(data, width, rmax, 8.0)
    stream = torch.cuda.current_stream()
    ms = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        ctx.lct(data, width, rmax, 8.0, out=out)
        b.record(stream)
        b.synchronize()
        ms.append(a.elapsed_time(b))
    P = 256 * 256 * 2048
    alg = 5 * 8 * P
    t = float(np.median(ms))
    res = {"workload": "LCT 128x128x1024 (16 ps, 2 m wall) padded to 256x256x2048, snr 8",
           "ms": t, "roofline": {"bound": "hbm", "achieved": alg / (t * 1e-3) / 1e9,
                                 "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                                 "frac": alg / (t * 1e-3) / 1e9 / PEAK_HBM_GBPS,
                                 "algorithmic_bytes": alg},
           "finite": bool(torch.isfinite(out).all())}
    if not args.no_cpu_baseline:
        from oracle import lct as olct
        workers = cpu_workers()
        # full size: the same capture, inverse filter prebuilt
        data_h = data.cpu().numpy()
        t0 = time.perf_counter()
        inv_full = olct.inverse_filter(128, 1024, width, rmax, 8.0, workers=workers)
        filt_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref_full = olct.lct(data_h, width, rmax, 8.0, invpsf=inv_full, workers=workers)
        full_s = time.perf_counter() - t0
        res["cpu"] = {"sample": "the same 128x128x1024 capture, one run (inverse filter "
                                "prebuilt, as the GPU caches it)",
                      "s": full_s, "inverse_filter_build_s": filt_s, "cores": workers,
                      "kind": "port (NumPy/SciPy restatement of python/lct.py)",
                      "rel_l2_gpu_vs_cpu": float(np.linalg.norm(out.cpu().numpy() - ref_full) /
                                                 np.linalg.norm(ref_full))}
        log(f"lct cpu (full size, {workers} workers): {full_s:.2f} s")
        del data_h, inv_full, ref_full
        # small configuration: GPU vs the CPU restatement, same input
        cap = small_confocal(seed=0)
        d1 = cap.transient.reshape(32, 32, 256).astype(np.float32)
        w1, r1 = 0.5, 256 * C_LIGHT * 32e-12
        inv1 = olct.inverse_filter(32, 256, w1, r1, 8.0, workers=workers)
        ref = olct.lct(d1, w1, r1, 8.0, invpsf=inv1, workers=workers)   # warm-up
        cpu_runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            ref = olct.lct(d1, w1, r1, 8.0, invpsf=inv1, workers=workers)
            cpu_runs.append(time.perf_counter() - t0)
        cpu_s = float(np.median(cpu_runs))
        g1 = ctx.lct(torch.from_numpy(d1).to(dev), w1, r1, 8.0)
        torch.cuda.synchronize()
        ms1 = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            ctx.lct(torch.from_numpy(d1).to(dev), w1, r1, 8.0, out=g1)
            b.record(stream)
            b.synchronize()
            ms1.append(a.elapsed_time(b))
        res["small"] = {"workload": "LCT 32x32x256 (32 ps, 1 m wall) padded to 64x64x512",
                        "gpu_ms": float(np.median(ms1)), "cpu_s": cpu_s, "cpu_cores": workers,
                        "cpu_kind": "port (NumPy/SciPy restatement of python/lct.py), median of 3",
                        "rel_l2_gpu_vs_cpu": float(np.linalg.norm(g1.cpu().numpy() - ref) /
                                                   np.linalg.norm(ref))}
    log(f"lct: {res}")
    del I, data, out
    torch.cuda.empty_cache()
    return res


def depth_leg(ctx, args, reps=3):
    """Depth-plane mode on the full confocal capture (128x128 scan, 1024 bins
    of 16 ps): 200 planes at focus +- 0.1 (FB/Laser.cu:365-377), unweighted
    sums (aligned grid: the ring convolution of the voxel path with weight 1;
    the gather kernel timed beside it; priced like the voxel path, 16 flops
    per (pixel, scan point, depth) update against FP32 VALU peak), min-max
    normalise, then the per-depth RGBA images enlarged to 256x256
    (slice_bytes_kernel + slice_rgba_kernel, HBM-bound: activation read
    4 B + byte planes written and read 2 B per (pixel, depth), 4 B per output
    pixel written).  CPU: the oracle's depth gather on the first 2 planes."""
    import numpy as np
    import torch

    from fasterbackprojection_amd.api import make_info
    from fasterbackprojection_amd.scene import full_confocal_geometry, random_quads

    dev = torch.device("cuda", 0)
    g = full_confocal_geometry(128, 1024, 128)
    info = make_info(g)
    sc = random_quads(np.random.default_rng(3), 4, 20000)
    laser = torch.from_numpy(g.laser_targets).to(dev)
    clean = ctx.render_confocal(laser, torch.from_numpy(sc.points.astype(np.float32)).to(dev),
                                torch.from_numpy(sc.albedo.astype(np.float32)).to(dev), 1024,
                                g.time_step)
    gen = torch.Generator(device=dev)
    gen.manual_seed(3)
    I = torch.poisson(clean * (1000.0 / clean.max()), generator=gen).contiguous()
    L, D = 128 * 128, 200
    focus = abs((np.float32(g.hidden_min[1]) + np.float32(g.hidden_max[1])) * np.float32(0.5))
    lo = float(np.float32(focus - np.float32(0.1)))
    hi = float(np.float32(focus + np.float32(0.1)))
    depths = torch.tensor(lo + (hi - lo) / D * np.arange(D), dtype=torch.float64, device=dev)
    act = ctx.backproject_depth(info, I, laser, depths)
    ctx.normalize(act)
    img = ctx.depth_slices_rgba(act, 128, 128, 256, 256)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    def timed(fn):
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.median(ms))

    t_bp = timed(lambda: ctx.backproject_depth(info, I, laser, depths, out=act))
    aligned = ctx.last_confocal_path_aligned()
    ctx.set_aligned_path(False)
    try:
        t_gather = timed(lambda: ctx.backproject_depth(info, I, laser, depths, out=act))
    finally:
        ctx.set_aligned_path(True)
    ctx.backproject_depth(info, I, laser, depths, out=act)
    ctx.normalize(act)
    t_norm = timed(lambda: ctx.normalize(act))
    t_img = timed(lambda: ctx.depth_slices_rgba(act, 128, 128, 256, 256, out=img))
    updates = L * L * D
    achieved = updates * FLOPS_PER_UPDATE / (t_bp * 1e-3) / 1e12
    img_bytes = L * D * (4 + 2) + D * 256 * 256 * 4
    img_gbps = img_bytes / (t_img * 1e-3) / 1e9
    res = {"workload": "depth planes: confocal 128x128x1024 (16 ps), 200 depths at focus +- 0.1 "
                       "-> activation [16384][200] + 200 RGBA images 256x256",
           "ms": {"depth_backprojection": t_bp, "normalize": t_norm, "slice_images": t_img,
                  "depth_backprojection_gather_kernel": t_gather},
           "path": "aligned ring convolution" if aligned else "gather kernel",
           "updates": updates,
           "roofline_depth_backprojection": {"bound": "valu", "achieved": achieved,
                                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                     "frac": achieved / PEAK_FP32_TFLOPS},
           "roofline_slice_images": {"bound": "hbm", "achieved": img_gbps, "peak": PEAK_HBM_GBPS,
                                     "unit": "GB/s", "frac": img_gbps / PEAK_HBM_GBPS,
                                     "algorithmic_bytes": img_bytes}}
    if not args.no_cpu_baseline:
        import oracle
        Ih = I.cpu().numpy()
        d2 = depths[:2].cpu().numpy()
        t0 = time.perf_counter()
        ref = oracle.depth(info, Ih, g.laser_targets, d2)
        cpu_s = time.perf_counter() - t0
        raw = ctx.backproject_depth(info, I, laser, depths[:2].contiguous())
        got = raw.cpu().numpy().astype(np.float64)
        res["cpu"] = {"sample": "oracle depth gather, first 2 of the 200 planes",
                      "updates_per_s": L * L * 2 / cpu_s,
                      "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                      "kind": "port",
                      "rel_l2_gpu_vs_cpu": float(np.linalg.norm(got - ref) / np.linalg.norm(ref))}
    log(f"depth: {res}")
    del I, act, img
    torch.cuda.empty_cache()
    return res


def exhaustive_leg(ctx, args, reps=3):
    """Exhaustive capture: 16x16 laser targets x 64x64 sensor targets on a
    2.0 m wall, 1024 bins of 16 ps (1,048,576 histograms, 4 GiB FP32), 8
    seeded random spheres (8000 samples) with 1/r^2 1/r^2 falloff and Poisson
    noise at a peak of 1000 -> 64^3 ellipsoidal backprojection + Laplacian
    (k = 5).  25 flops per pair-update (two distances)."""
    import numpy as np
    import torch

    from fasterbackprojection_amd.api import make_info
    from fasterbackprojection_amd.scene import exhaustive_geometry, random_spheres

    dev = torch.device("cuda", 0)
    g = exhaustive_geometry()
    info = make_info(g)
    lasers = torch.from_numpy(g.laser_targets).to(dev)
    sensors = torch.from_numpy(g.sensor_targets).to(dev)
    sc = random_spheres(np.random.default_rng(3), 8, 8000)
    pts = torch.from_numpy(sc.points.astype(np.float32)).to(dev)
    alb = torch.from_numpy(sc.albedo.astype(np.float32)).to(dev)
    I = ctx.render_exhaustive(lasers, sensors, pts, alb, g.num_time_bins, g.time_step)
    I.mul_(1000.0 / I.max())
    gen = torch.Generator(device=dev)
    gen.manual_seed(3)
    I = torch.poisson(I, generator=gen)
    torch.cuda.synchronize()
    L, S = len(g.laser_targets), len(g.sensor_targets)
    res = g.voxel_resolution[0]
    updates = L * S * res ** 3
    vol = torch.empty((res,) * 3, dtype=torch.float32, device=dev)
    filt = torch.empty_like(vol)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.mean(ms))

    bp = timed(lambda: ctx.backproject_exhaustive(info, I, lasers, sensors, out=vol))
    kernel = ("exhaustive_group_kernel (laser groups of 8)" if ctx.last_exhaustive_laser_groups()
              else "exhaustive_voxel_kernel (per pair)")
    # The per-pair kernel on the same capture, for comparison.
    ctx.set_exhaustive_laser_groups(False)
    vol_pp = torch.empty_like(vol)
    bp_pair = timed(lambda: ctx.backproject_exhaustive(info, I, lasers, sensors, out=vol_pp))
    ctx.set_exhaustive_laser_groups(True)
    pair_rel = float(torch.linalg.vector_norm((vol_pp - vol).double())
                     / torch.linalg.vector_norm(vol.double()))
    del vol_pp
    lap = timed(lambda: ctx.laplacian(vol, 5, out=filt))
    # Depth-plane mode of the same capture (backprojectExhaustive): 200
    # planes at focus +- 0.1 over the laser-grid pixels, every pair.
    focus = abs((np.float32(g.hidden_min[1]) + np.float32(g.hidden_max[1])) * np.float32(0.5))
    lo_d = float(np.float32(focus - np.float32(0.1)))
    hi_d = float(np.float32(focus + np.float32(0.1)))
    depths = torch.tensor(lo_d + (hi_d - lo_d) / 200 * np.arange(200), dtype=torch.float64,
                          device=dev)
    act = torch.empty((L, 200), dtype=torch.float32, device=dev)
    t_depth = timed(lambda: ctx.backproject_depth(info, I, lasers, depths, sensor_targets=sensors,
                                                  out=act))
    depth_kernel = ("depth_group_kernel (laser groups of 8)" if ctx.last_exhaustive_laser_groups()
                    else "depth_exhaustive_window_kernel")
    ctx.set_exhaustive_laser_groups(False)
    act_pp = torch.empty_like(act)
    t_depth_pair = timed(lambda: ctx.backproject_depth(info, I, lasers, depths,
                                                       sensor_targets=sensors, out=act_pp))
    ctx.set_exhaustive_laser_groups(True)
    depth_pair_rel = float(torch.linalg.vector_norm((act_pp - act).double())
                           / torch.linalg.vector_norm(act.double()))
    del act_pp
    depth_updates = L * 200 * L * S
    # backprojectExhaustive per (pixel, depth, pair): |p - l| (3 sub, 3 mul,
    # 2 add, sqrt) + |p - s| (cached per sensor) - t0, / dt, + sample = 13.
    depth_tflops = 13 * depth_updates / (t_depth * 1e-3) / 1e12
    # Issue slots per counted pair-update of the depth-plane group kernel from
    # its counters (10.7 VALU, 0.16 transcendental; 200 of its 256 depth slots
    # per tile hold planes).
    depth_slot_bound = VALU_LANE_SLOTS / DEPTH_EXH_SLOTS
    depth_check = None
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        cols = [3, 151]
        ref_d = oracle.depth(info, I.cpu().numpy(), g.laser_targets,
                             depths.cpu().numpy()[cols], sensor=g.sensor_targets, exhaustive=True)
        got_d = act[:, cols].cpu().numpy().astype(np.float64)
        err_d = np.abs(got_d - ref_d) / np.maximum(np.abs(ref_d), 1e-3 * np.abs(ref_d).max())
        depth_check = {"sample": "oracle on planes 3 and 151 (all pixels, all pairs)",
                       "max_rel_err": float(err_d.max())}
        # The oracle on the first depth slice (64x64 voxels x all pairs).
        import oracle
        sub = type(info).from_buffer_copy(info)
        sub.voxel_resolution.z = 1
        Ih = I.cpu().numpy()
        t0 = time.perf_counter()
        ref = oracle.exhaustive_voxel(sub, Ih, g.laser_targets, g.sensor_targets)
        cpu_s = time.perf_counter() - t0
        del Ih
        got = vol[:1].cpu().numpy().astype(np.float64)
        cpu = {"sample": "FP64 OpenMP oracle on the first of 64 depth slices (64x64 voxels x "
                         "all 1,048,576 pairs)",
               "pair_updates_per_s": L * S * res * res / cpu_s,
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
               "kind": "port",
               "rel_l2_gpu_vs_cpu": float(np.linalg.norm(got - ref) / np.linalg.norm(ref))}
    tflops = 25 * updates / (bp * 1e-3) / 1e12
    # Issue slots per pair-update of the laser-group kernel measured by
    # counters over the whole kernel (profiles/exhaustive_group_c3_bench_counters.md):
    # 9.2 VALU instructions, 0.13 of them transcendental (the sensor square roots) at 2.37 slots.
    slot_bound = VALU_LANE_SLOTS / EXH_SLOTS
    out = {"workload": "exhaustive 16^2 lasers x 64^2 sensors x 1024 (16 ps, 4 GiB) -> 64^3 "
                       "+ Laplacian(k=5)",
           "pair_updates": updates, "kernel": kernel, "backprojection_ms": bp,
           "per_pair_kernel": {"ms": bp_pair, "rel_l2_vs_group_kernel": pair_rel},
           "laplacian_ms": lap,
           "pair_updates_per_s": updates / ((bp + lap) * 1e-3),
           "roofline": {"bound": "valu", "achieved": tflops, "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": tflops / PEAK_FP32_TFLOPS,
                        "flops_per_update": 25},
           "issue_slot_diagnostic": {"valu_slots_per_update": EXH_SLOTS, "bound_updates_per_s": slot_bound,
                                     "frac": updates / (bp * 1e-3) / slot_bound},
           "depth_planes": {"workload": "200 depth planes at focus +- 0.1 over the 16x16 laser "
                                        "pixels, all 1,048,576 pairs",
                            "kernel": depth_kernel, "ms": t_depth,
                            "staged_window_kernel": {"ms": t_depth_pair,
                                                     "rel_l2_vs_group_kernel": depth_pair_rel},
                            "pair_updates": depth_updates,
                            "roofline": {"bound": "valu", "achieved": depth_tflops,
                                         "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                         "frac": depth_tflops / PEAK_FP32_TFLOPS,
                                         "flops_per_update": 13},
                            "issue_slot_diagnostic": {
                                "valu_slots_per_update": DEPTH_EXH_SLOTS,
                                "bound_updates_per_s": depth_slot_bound,
                                "frac": depth_updates / (t_depth * 1e-3) / depth_slot_bound},
                            "oracle_check": depth_check},
           "finite": bool(torch.isfinite(vol).all()),
           "cpu": cpu}
    log(f"exhaustive: {out}")
    del I
    torch.cuda.empty_cache()
    return out


def learned_leg(ctx, n_scenes=16):
    """64x64 confocal scan, 512 bins of 16 ps, 64^3 volume: plain path (all
    scan points + LoG + normalisation) against the learned path (MIS at 1/16
    of the scan-point x bin pairs + refiner) on held-out seeds."""
    import numpy as np
    import torch

    from fasterbackprojection_amd.learned import WEIGHTS, LearnedReconstructor, rel_l2

    if not (WEIGHTS / "refiner_c5.f32").exists():
        log("learned leg skipped: no trained weights")
        return None
    lrn = LearnedReconstructor(ctx)
    p = lrn.p
    stream = torch.cuda.current_stream()

    def timed(fn, reps=5):
        fn()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.median(ms))

    t_plain, t_learn, t_check, e_learn, e_base, fb = [], [], [], [], [], 0
    t_general = []
    for s in range(2_000_000, 2_000_000 + n_scenes):
        I = p.render(s)
        y = p.plain(I).clone()
        t_plain.append(timed(lambda: p.plain(I)))
        if s < 2_000_000 + 4:
            # The plain path on the general kernel: what a capture whose scan
            # grid is not aligned with the voxel lattice gets.
            ctx.set_aligned_path(False)
            try:
                t_general.append(timed(lambda: p.plain(I)))
            finally:
                ctx.set_aligned_path(True)
        t_learn.append(timed(lambda: lrn(I, s, check=False)))
        t0 = time.perf_counter()
        out, fell_back, _ = lrn(I, s, check=True)
        t_check.append((time.perf_counter() - t0) * 1e3)
        fb += int(fell_back)
        out, _, _ = lrn(I, s, check=False)
        e_learn.append(rel_l2(out, y))
        e_base.append(rel_l2(p.mis_input(I, s), y))
    res = {"workload": "confocal 64x64x512 (16 ps, 1-6 quads/spheres) -> 64^3; "
                       "learned = MIS (L/16 samples) + LoG + 3-level U-Net refiner",
           "held_out_scenes": n_scenes,
           "plain_ms": float(np.mean(t_plain)), "learned_ms": float(np.mean(t_learn)),
           "speedup": float(np.mean(t_plain) / np.mean(t_learn)),
           "plain_general_kernel_ms": float(np.mean(t_general)),
           "speedup_vs_plain_general_kernel": float(np.mean(t_general) / np.mean(t_learn)),
           "learned_with_residual_check_ms": float(np.mean(t_check)),
           "rel_l2_learned_vs_plain": float(np.mean(e_learn)),
           "rel_l2_unrefined_mis_vs_plain": float(np.mean(e_base)),
           "fallbacks": fb, "residual_threshold": lrn.threshold}
    # The refiner's forward pass alone against the dense BF16 MFMA peak:
    # flops = sum over layers of 2 * V_out * k^3 * Cin * Cout at 64^3
    # (16/32/64 channels; the 1->16 and 16->1 layers run on VALU).
    S = 64
    layers = [(S, 3, 1, 16), (S // 2, 3, 16, 32), (S // 4, 3, 32, 64), (S // 2, 3, 64, 32),
              (S, 1, 32, 16), (S, 3, 16, 1)]
    flops = sum(2 * r ** 3 * k ** 3 * ci * co for r, k, ci, co in layers)
    gen = torch.Generator(device=torch.device("cuda", 0))
    gen.manual_seed(3)
    x = torch.rand((S, S, S), device=torch.device("cuda", 0), generator=gen)
    yb = torch.empty_like(x)
    t_ref = timed(lambda: lrn.net.forward(x, out=yb))
    tf = flops / (t_ref * 1e-3) / 1e12
    res["refiner_forward"] = {"ms": t_ref, "flops": flops,
                              "roofline": {"bound": "mfma", "achieved": tf,
                                           "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                                           "frac": tf / PEAK_BF16_TFLOPS}}
    log(f"learned: {res}")
    return res


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_workers():
    """Host threads a CPU leg uses: OMP_NUM_THREADS when set (the GPU box
    sets it to its share of cores), else the cores this process may run on."""
    import os
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(info, I, g, target_seconds):
    """FP64 OpenMP oracle (its -O3 -march=native build for this host) on the
    first z-slices of the same volume: one slice as warm-up and calibration,
    then the median of 3 timed runs sized to take about target_seconds
    together; one core: the median of 3 runs on a quarter slice."""
    import numpy as np

    import oracle
    host = I.cpu().numpy()
    laser = g.laser_targets
    X, Y = info.voxel_resolution.x, info.voxel_resolution.y
    L = info.num_laser_targets
    sub = type(info).from_buffer_copy(info)

    def run(nz):
        sub.voxel_resolution.z = nz
        t = time.perf_counter()
        oracle.confocal_voxel(sub, host, laser)
        return time.perf_counter() - t

    t1 = run(1)
    nz = int(max(1, min(info.voxel_resolution.z, target_seconds / 3 / max(t1, 1e-3))))
    times = [run(nz) for _ in range(3)]
    t = float(np.median(times))
    cores = oracle.max_threads()
    log(f"cpu oracle: {nz} slices in {t:.2f} s (median of 3) on {cores} threads")
    # One core on a smaller slab (a quarter of one slice's rows).
    oracle.set_threads(1)
    try:
        sub.voxel_resolution.z = 1
        sub.voxel_resolution.y = max(1, Y // 4)
        t_ones = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.confocal_voxel(sub, host, laser)
            t_ones.append(time.perf_counter() - t0)
        t_one = float(np.median(t_ones))
    finally:
        oracle.set_threads(cores)
    one_core = X * max(1, Y // 4) * L / t_one
    log(f"cpu oracle, 1 thread: {one_core:.3e} voxel-updates/s")
    return {
        "value": X * Y * nz * L / t,
        "value_one_core": one_core,
        "cpu_model": cpu_model(),
        "unit": "voxel-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 OpenMP oracle (oracle/fbp_oracle.c, -O3 -march=native) on the first "
                  f"{nz} of {info.voxel_resolution.z} depth slices ({X}x{Y}x{nz} voxels x {L} "
                  f"scan points), median of 3 runs after a warm-up slice",
        "runs_s": times,
    }


if __name__ == "__main__":
    main()
