#!/usr/bin/env python3
"""Frame-graph step throughput of the SLAM inner loop on one MI355X.

Workload: 64 keyframes, 384 edges, 64x80 maps (1/8 of 512x640), 128-channel
float32 features, fx=fy=80, cx=40, cy=32, seed 0 (vipe_amd.synth
"kf64_e384_64x80").  One step = one correlation lookup (4-level pyramid,
radius 3, all 384 edges -> [384, 196, 64, 80]) + 8 Gauss-Newton iterations of
dense BA (``slam_ext.ba_extended``, FP64 reduced-camera Cholesky).  The
pyramid is built once before timing (its own throughput is reported under
``corr_build``).  Inputs are resident in HBM when timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode fp32|mixed]

Prints one JSON line.  ``value`` = BA GN iterations/s of the combined step;
``corr_lookups_per_s`` = lookups/s of the same run.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import torch

METRIC = "BA GN iters/s + corr lookups/s at 64 KF; mixed vs FP32 GPU at equal px error"
HBM_PEAK = 8.0e12
BF16_PEAK = 2.5e15
FP32_PEAK = 157.3e12
GN_ITERS = 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["fp32", "mixed"], default="fp32")
    ap.add_argument("--workload", default="kf64_e384_64x80")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-edges", type=int, default=32, help="edges of the bounded CPU lookup sample")
    ap.add_argument("--build-variants", action="store_true",
                    help="also time the volume build with f16 / bf16 / fp8 contractions and report their error")
    return ap.parse_args()


def lookup_bytes(E, H, W, L, elem, r=3):
    """Algorithmic bytes of one lookup: per edge and level, (2r+2)^2 taps per pixel
    read once, (2r+1)^2 outputs written, 2 coordinates read."""
    taps = (2 * r + 2) ** 2
    outs = (2 * r + 1) ** 2
    return E * H * W * L * (taps * elem + outs * elem + 8)


def ba_bytes(E, N, H, W):
    """Compulsory bytes of one GN iteration: targets + weights once, disps read +
    write, disps_sens and eta once (float32)."""
    return E * 4 * H * W * 4 + N * H * W * 4 * 4


def profiled_traffic(kernel: str):
    """HBM bytes per dispatch of ``kernel`` from the committed rocprofv3 summaries
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration); None when no
    summary covers it."""
    import glob

    here = os.path.dirname(os.path.abspath(__file__))
    for fn in sorted(glob.glob(os.path.join(here, "profiles", "*.md"))):
        with open(fn) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_dispatch ") and ":" in line:
                    name, val = line[len("traffic_bytes_per_dispatch "):].rsplit(":", 1)
                    if name.strip() == kernel:
                        return float(val)
    return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(how the path would shard over several GPUs is described, not implemented).", file=sys.stderr)
        sys.exit(2)
    from vipe_amd import synth
    from vipe_amd.corr import CorrBlock
    from vipe_amd.ext import droid_net_ext, slam_ext

    dev = torch.device("cuda", 0)
    g = synth.make_graph(a.workload, seed=0)
    N, H, W = g.disps.shape
    E = g.ii.shape[0]
    d = {k: (v.to(dev) if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}
    depth_active = torch.ones(N, device=dev)
    vol_dtype = torch.float32 if a.mode == "fp32" else torch.float16
    fm = d["fmaps"]
    f1 = fm[d["ii"]].contiguous()[None]
    f2 = fm[d["jj"]].contiguous()[None]
    if a.mode == "mixed":
        f1, f2 = f1.half(), f2.half()

    # ---- pyramid build (timed separately, after a warm build so that the
    # allocator's cached blocks serve the pyramid and workspace) ----
    s = torch.cuda.current_stream()
    corr = CorrBlock(f1, f2, num_levels=4, radius=3, volume_dtype=vol_dtype)
    del corr
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    corr = CorrBlock(f1, f2, num_levels=4, radius=3, volume_dtype=vol_dtype)
    e1.record(s)
    torch.cuda.synchronize()
    build_ms = e0.elapsed_time(e1)
    build_variants = build_variant_report(f1, f2, vol_dtype) if a.build_variants else None
    del f1, f2
    coords = d["coords"][None]  # [1, E, H, W, 2]

    poses0 = d["poses"].clone()
    disps0 = d["disps"].clone()
    lk_events = []

    def step(record=False):
        if record:
            ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ea.record(s)
        out = corr(coords)
        if record:
            eb.record(s)
            lk_events.append((ea, eb))
        slam_ext.ba_extended(d["poses"], d["disps"], d["intrinsics"], d["disps_sens"], d["targets"], d["weights"],
                             d["eta"], d["ii"], d["jj"], depth_active, g.t0, g.t1, GN_ITERS, 1e-4, 0.1, False, 0.05,
                             False, 1e-6, 1e-6, 1.0, False)
        return out

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    d["poses"].copy_(poses0)
    d["disps"].copy_(disps0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(s)
    for _ in range(a.steps):
        step(record=True)
    ev1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / a.steps
    lk_ms = sum(x.elapsed_time(y) for x, y in lk_events) / len(lk_events)
    elem = 4 if vol_dtype == torch.float32 else 2
    lk_alg = lookup_bytes(E, H, W, 4, elem)
    lk_gbs = lk_alg / (lk_ms * 1e-3)
    lk_kernel = (f"corr_lookup_tiled_kernel<{'float' if elem == 4 else '__half'}, 3>" if corr.layout == "tiled"
                 else f"corr_lookup_pyramid_kernel<{'float' if elem == 4 else '__half'}, 3, false>")

    # BA-only timing (same inputs, separate region)
    d["poses"].copy_(poses0)
    d["disps"].copy_(disps0)
    torch.cuda.synchronize()
    eb0, eb1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    eb0.record(s)
    nba = max(3, a.steps // 2)
    for _ in range(nba):
        slam_ext.ba_extended(d["poses"], d["disps"], d["intrinsics"], d["disps_sens"], d["targets"], d["weights"],
                             d["eta"], d["ii"], d["jj"], depth_active, g.t0, g.t1, GN_ITERS, 1e-4, 0.1, False, 0.05,
                             False, 1e-6, 1e-6, 1.0, False)
    eb1.record(s)
    torch.cuda.synchronize()
    ba_ms_per_iter = eb0.elapsed_time(eb1) / (nba * GN_ITERS)

    value = GN_ITERS * a.steps / (gpu_ms * 1e-3)
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GN iters/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.mode == "fp32" else "f16-volume/f32/f64",
        "data": "synthetic (seeded frame graph, random features)",
        "config": {"workload": f"{a.workload}: 64 KF, 384 edges, 64x80, 128-ch, 4-level r=3 lookup + "
                               f"{GN_ITERS} GN iters (FP64 Cholesky)", "mode": a.mode,
                   "volume_dtype": str(vol_dtype).replace("torch.", "")},
        "corr_lookups_per_s": round(a.steps / (gpu_ms * 1e-3), 2),
        "lookup_ms": round(lk_ms, 4),
        "ba_ms_per_gn_iter": round(ba_ms_per_iter, 4),
        "ba_gn_iters_per_s_alone": round(1e3 / ba_ms_per_iter, 1),
        "corr_build_ms": round(build_ms, 3),
        "corr_build_tflops": round(2.0 * E * (H * W) ** 2 * 128 / (build_ms * 1e-3) / 1e12, 2),
        "corr_build_variants": build_variants,
        "wall_s": round(wall, 3),
        "roofline": {
            "kernel": lk_kernel,
            "bound": "hbm",
            "achieved": round(lk_gbs / 1e9, 1),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(lk_gbs / HBM_PEAK, 4),
            "traffic": profiled_traffic(lk_kernel),
        },
    }
    if not a.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(g, a.cpu_edges)
    print(json.dumps(result), flush=True)


def build_variant_report(f1, f2, vol_dtype, sample_edges=2):
    """Warm-timed pyramid builds of the same features with each MFMA contraction
    type, and the max / relative-L2 error of level 0 against an FP64 volume of
    the unrounded features on ``sample_edges`` edges."""
    from oracle import corr as OC
    from vipe_amd.ext import droid_net_ext

    B, N, Cc, H, W = f1.shape
    x1 = f1.reshape(B * N, Cc, H, W).float().contiguous()
    x2 = f2.reshape(B * N, Cc, H, W).float().contiguous()
    ref = OC.corr_volume(x1[:sample_edges].double().cpu()[None], x2[:sample_edges].double().cpu()[None])[0]
    ref = ref.reshape(sample_edges, H * W, H * W)
    out = {}
    s = torch.cuda.current_stream()
    for name, mma in (("f32", torch.float32), ("f16", torch.float16), ("bf16", torch.bfloat16),
                      ("fp8_e4m3", torch.float8_e4m3fn)):
        lv = droid_net_ext.corr_build(x1, x2, 4, out_dtype=vol_dtype, contraction_dtype=mma)
        del lv
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        lv = droid_net_ext.corr_build(x1, x2, 4, out_dtype=vol_dtype, contraction_dtype=mma)
        e1.record(s)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        got = lv[0][:sample_edges].double().cpu().reshape(sample_edges, H * W, H * W)
        d = got - ref
        out[name] = {"ms": round(ms, 3), "tflops": round(2.0 * B * N * (H * W) ** 2 * Cc / (ms * 1e-3) / 1e12, 1),
                     "max_err": float(d.abs().max()), "rel_l2_err": float(d.norm() / ref.norm())}
        del lv, got, d
    return out


def cpu_baseline(g, n_edges):
    """Oracle restatement (float64 torch on the host CPU) on a bounded sample of
    the same workload: 1 GN iteration of the full 384-edge graph and one
    4-level lookup (pyramid built outside the timer, as on the GPU) over
    ``n_edges`` edges, scaled to 384."""
    from oracle import ba as OB
    from oracle import corr as OC

    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    N = g.disps.shape[0]
    t0 = time.perf_counter()
    OB.ba_extended(g.poses.clone(), g.disps.clone(), g.intrinsics.clone(), g.disps_sens, g.targets, g.weights,
                   g.eta, g.ii, g.jj, torch.ones(N), g.t0, g.t1, 1, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0,
                   False)
    t_iter = time.perf_counter() - t0
    sel = torch.arange(n_edges)
    f1 = g.fmaps[g.ii[sel]][None]
    f2 = g.fmaps[g.jj[sel]][None]
    pyr = OC.corr_pyramid(OC.corr_volume(f1, f2), 4)
    t1 = time.perf_counter()
    OC.corr_block_lookup(pyr, g.coords[sel][None].double(), 3)
    t_lk = (time.perf_counter() - t1) * g.ii.shape[0] / n_edges
    step = GN_ITERS * t_iter + t_lk
    return {"value": round(GN_ITERS / step, 4), "unit": "GN iters/s", "cores": threads, "kind": "port",
            "sample": f"1 GN iteration on the full graph ({t_iter:.2f} s) + lookup of {n_edges} of "
                      f"{g.ii.shape[0]} edges scaled ({t_lk:.2f} s per lookup); float64 oracle; "
                      f"cpu={platform.processor() or platform.machine()}"}


if __name__ == "__main__":
    main()
