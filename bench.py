#!/usr/bin/env python3
"""Frame-graph step throughput of the SLAM inner loop on one MI355X.

Workloads (vipe_amd.synth, seed 0): 64 keyframes, 384 edges, 64x80 maps (1/8
of 512x640), 128-channel float32 features, fx=fy=80, cx=40, cy=32.
``kf64_e384_64x80`` has SE3 poses, ``kf64_e384_64x80_sim3`` the same graph
with Sim3 poses (log-scale N(0, 0.02)).  One step = one correlation lookup
(4-level pyramid, radius 3, all 384 edges -> [384, 196, 64, 80]) + 8
Gauss-Newton iterations of dense BA (``slam_ext.ba_extended``, FP64
reduced-camera Cholesky).  The pyramid is built once before timing (its time
is reported as ``corr_build_ms``).  Inputs are resident in HBM when timing
starts.

The line carries the whole metric:
- ``value``: GN iterations/s of the plain FP32 step on the SE3 workload;
- ``mixed``: the step on the Sim3 workload with an FP16 volume built on FP16
  matrix cores (FP32 accumulation), FP32 Jacobians and Hessian, FP64 Cholesky
  and retraction; its speed-up over the FP32 step, with and without the
  pyramid build, and over an FP32 step on the same Sim3 workload
  (``fp32_sim3``);
- each run's error against the FP64 oracle: BA reprojection (px) and
  relative pose error after the 8 iterations, and the lookup's largest error
  on three sampled edges (first, middle, last);
- ``roofline`` (the lookup kernel, the step's dominant kernel),
  ``roofline_ba`` (one GN iteration) and ``roofline_adjoint`` (one GN
  iteration plus its implicit adjoint);
- ``cpu_baseline``: the reference's generic CPU path, restated in float32
  torch (oracle/generic_ba.py), on the box's host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode fp32|mixed]

``--mode mixed`` makes the mixed run the headline ``value``.
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
GN_ITERS = 8
SE3_WORKLOAD = "kf64_e384_64x80"
SIM3_WORKLOAD = "kf64_e384_64x80_sim3"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["fp32", "mixed"], default="fp32",
                    help="which run is the headline value (both always run)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 oracle comparisons")
    ap.add_argument("--only", choices=["fp32", "mixed"], default=None,
                    help="run one mode only (profiling); the line then lacks the comparison")
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


def adjoint_bytes(E, N, K, H, W):
    """Compulsory bytes of the implicit adjoint of one iteration: targets +
    weights read, their two gradients written, disps and dL/d(dz) read."""
    return 2 * E * 4 * H * W * 4 + N * H * W * 4 + K * H * W * 4


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


def _events():
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _ba(d, depth_active, g, iters):
    from vipe_amd.ext import slam_ext

    return slam_ext.ba_extended(d["poses"], d["disps"], d["intrinsics"], d["disps_sens"], d["targets"],
                                d["weights"], d["eta"], d["ii"], d["jj"], depth_active, g.t0, g.t1, iters, 1e-4,
                                0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, False)


def gpu_run(workload, mode, steps, warmup, build_variants=False):
    """Times the step of one (workload, mode) and returns its figures plus what
    the accuracy check needs (the lookup output of the sampled edges and the
    state after one step from the initial poses)."""
    from vipe_amd import synth
    from vipe_amd.corr import CorrBlock

    dev = torch.device("cuda", 0)
    s = torch.cuda.current_stream()
    g = synth.make_graph(workload, seed=0)
    N, H, W = g.disps.shape
    E = g.ii.shape[0]
    d = {k: (v.to(dev) if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}
    depth_active = torch.ones(N, device=dev)
    vol_dtype = torch.float32 if mode == "fp32" else torch.float16
    fm = d["fmaps"]
    f1 = fm[d["ii"]].contiguous()[None]
    f2 = fm[d["jj"]].contiguous()[None]
    if mode == "mixed":
        f1, f2 = f1.half(), f2.half()

    # pyramid build, timed after a warm build so that the allocator's cached
    # blocks serve the pyramid and its workspace
    corr = CorrBlock(f1, f2, num_levels=4, radius=3, volume_dtype=vol_dtype)
    del corr
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record(s)
    corr = CorrBlock(f1, f2, num_levels=4, radius=3, volume_dtype=vol_dtype)
    e1.record(s)
    torch.cuda.synchronize()
    build_ms = e0.elapsed_time(e1)
    variants = build_variant_report(f1, f2, vol_dtype) if build_variants else None
    del f1, f2
    coords = d["coords"][None]  # [1, E, H, W, 2]
    poses0, disps0 = d["poses"].clone(), d["disps"].clone()
    lk_events = []

    def step(record=False):
        if record:
            ea, eb = _events()
            ea.record(s)
        out = corr(coords)
        if record:
            eb.record(s)
            lk_events.append((ea, eb))
        _ba(d, depth_active, g, GN_ITERS)
        return out

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    d["poses"].copy_(poses0)
    d["disps"].copy_(disps0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0, ev1 = _events()
    ev0.record(s)
    for _ in range(steps):
        step(record=True)
    ev1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    lk_ms = sum(x.elapsed_time(y) for x, y in lk_events) / len(lk_events)
    elem = 4 if vol_dtype == torch.float32 else 2

    # BA alone (same inputs, separate region)
    d["poses"].copy_(poses0)
    d["disps"].copy_(disps0)
    torch.cuda.synchronize()
    eb0, eb1 = _events()
    eb0.record(s)
    nba = max(3, steps // 2)
    for _ in range(nba):
        _ba(d, depth_active, g, GN_ITERS)
    eb1.record(s)
    torch.cuda.synchronize()
    ba_ms_per_iter = eb0.elapsed_time(eb1) / (nba * GN_ITERS)

    # the state after one step from the initial poses, and the lookup of the
    # sampled edges (for the accuracy check, outside the timed regions)
    d["poses"].copy_(poses0)
    d["disps"].copy_(disps0)
    sample = torch.tensor([0, E // 2, E - 1])
    out = corr(coords)
    lk_sample = out[0, sample.to(dev)].float().cpu()
    del out
    _ba(d, depth_active, g, GN_ITERS)
    torch.cuda.synchronize()
    state = (d["poses"].cpu(), d["disps"].cpu())
    lk_kernel = (f"corr_lookup_tiled_kernel<{'float' if elem == 4 else '__half'}, 3>" if corr.layout == "tiled"
                 else f"corr_lookup_pyramid_kernel<{'float' if elem == 4 else '__half'}, 3, false>")
    del corr
    torch.cuda.empty_cache()
    return {
        "workload": workload, "mode": mode, "graph": g, "ms_per_step": gpu_ms / steps, "gpu_ms": gpu_ms,
        "wall": wall, "lookup_ms": lk_ms, "ba_ms_per_iter": ba_ms_per_iter, "build_ms": build_ms,
        "lookup_bytes": lookup_bytes(E, H, W, 4, elem), "lookup_kernel": lk_kernel, "vol_dtype": vol_dtype,
        "lk_sample_edges": sample, "lk_sample": lk_sample, "state": state, "build_variants": variants,
    }


def adjoint_run(workload, reps=10):
    """Device time of one GN iteration (forward) and of its implicit adjoint
    (vh_ba_adjoint), each bracketed by HIP events around its own call and
    averaged over ``reps`` pairs on the same inputs (the host work between the
    calls stays outside the brackets)."""
    from vipe_amd import synth
    from vipe_amd.ext import slam_ext

    dev = torch.device("cuda", 0)
    s = torch.cuda.current_stream()
    g = synth.make_graph(workload, seed=0, with_fmaps=False)
    N, H, W = g.disps.shape
    E = g.ii.shape[0]
    d = {k: (v.to(dev) if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}
    da = torch.ones(N, device=dev)
    p0, d0 = d["poses"].clone(), d["disps"].clone()
    gen = torch.Generator().manual_seed(1)
    gdx = torch.randn(g.t1 - g.t0, 6, generator=gen).to(dev)
    fwd, adj = [], []

    def pair(record):
        p, dd = p0.clone(), d0.clone()
        ea, eb, ec = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True), \
            torch.cuda.Event(enable_timing=True)
        ea.record(s)
        dx, dz, _, plan, ws = slam_ext._ba_extended(p, dd, d["intrinsics"].clone(), d["disps_sens"], d["targets"],
                                                    d["weights"], d["eta"], d["ii"], d["jj"], da, g.t0, g.t1, 1,
                                                    1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, False)
        eb.record(s)
        gdz = torch.ones_like(dz)
        eb2 = torch.cuda.Event(enable_timing=True)
        eb2.record(s)
        slam_ext.ba_adjoint(plan, ws, p0, d0, d["intrinsics"], d["targets"], d["weights"], da, 1e-4, False, gdx, gdz)
        ec.record(s)
        if record:
            fwd.append((ea, eb))
            adj.append((eb2, ec))
        return plan

    plan = pair(False)
    torch.cuda.synchronize()
    for _ in range(reps):
        pair(True)
    torch.cuda.synchronize()
    times = {"forward": sum(a.elapsed_time(b) for a, b in fwd) / reps,
             "adjoint": sum(a.elapsed_time(b) for a, b in adj) / reps}
    times["forward_adjoint"] = times["forward"] + times["adjoint"]
    alg = ba_bytes(E, N, H, W) + adjoint_bytes(E, N, plan.K, H, W)
    return times, alg


def accuracy(run, oracle_cache):
    """Errors of one run against the FP64 oracle: BA after one step (8
    iterations) from the initial state, and the lookup on the sampled edges."""
    from oracle import ba as OB
    from oracle import corr as OC
    from vipe_amd.synth import reproject

    g = run["graph"]
    key = run["workload"]
    if key not in oracle_cache:
        N = g.disps.shape[0]
        rp, rd = g.poses.clone(), g.disps.clone()
        OB.ba_extended(rp, rd, g.intrinsics.clone(), g.disps_sens, g.targets, g.weights, g.eta, g.ii, g.jj,
                       torch.ones(N), g.t0, g.t1, GN_ITERS, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, False)
        sel = run["lk_sample_edges"]
        f1 = g.fmaps[g.ii[sel]][None].double()
        f2 = g.fmaps[g.jj[sel]][None].double()
        ref_lk = OC.corr_block_lookup(OC.corr_pyramid(OC.corr_volume(f1, f2), 4), g.coords[sel][None].double(), 3)[0]
        oracle_cache[key] = (rp, rd, ref_lk)
    rp, rd, ref_lk = oracle_cache[key]
    p, dd = run["state"]
    ca, _ = reproject(p, dd, g.intrinsics, g.ii, g.jj)
    cb, _ = reproject(rp, rd, g.intrinsics, g.ii, g.jj)
    px = (ca - cb).abs().max().item()
    rel_pose = ((p.double() - rp.double()).norm(dim=1) / rp.double().norm(dim=1)).max().item()
    lk_err = (run["lk_sample"].double() - ref_lk).abs().max().item()
    return {"ba_px_err_vs_fp64": float(f"{px:.3e}"), "ba_rel_pose_err_vs_fp64": float(f"{rel_pose:.3e}"),
            "lookup_max_err_vs_fp64": float(f"{lk_err:.3e}"),
            "lookup_max_abs": float(f"{ref_lk.abs().max().item():.3e}")}


def summary(run):
    return {
        "workload": run["workload"],
        "ms_per_step": round(run["ms_per_step"], 4),
        "gn_iters_per_s": round(GN_ITERS / (run["ms_per_step"] * 1e-3), 2),
        "corr_lookups_per_s": round(1e3 / run["ms_per_step"], 2),
        "lookup_ms": round(run["lookup_ms"], 4),
        "ba_ms_per_gn_iter": round(run["ba_ms_per_iter"], 4),
        "corr_build_ms": round(run["build_ms"], 3),
        "volume_dtype": str(run["vol_dtype"]).replace("torch.", ""),
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(how the path would shard over several GPUs is described, not implemented).", file=sys.stderr)
        sys.exit(2)
    runs = {}
    if a.only in (None, "fp32"):
        runs["fp32"] = gpu_run(SE3_WORKLOAD, "fp32", a.steps, a.warmup, a.build_variants)
    if a.only in (None, "mixed"):
        runs["mixed"] = gpu_run(SIM3_WORKLOAD, "mixed", a.steps, a.warmup)
    if a.only is None:
        runs["fp32_sim3"] = gpu_run(SIM3_WORKLOAD, "fp32", a.steps, a.warmup)
    adj_times, adj_alg = adjoint_run(SE3_WORKLOAD)

    head = runs[a.mode] if a.mode in runs else next(iter(runs.values()))
    g = head["graph"]
    N, H, W = g.disps.shape
    E = g.ii.shape[0]
    lk_gbs = head["lookup_bytes"] / (head["lookup_ms"] * 1e-3)
    ba_alg = ba_bytes(E, N, H, W)
    ba_gbs = ba_alg / (head["ba_ms_per_iter"] * 1e-3)
    adj_gbs = adj_alg / (adj_times["forward_adjoint"] * 1e-3)
    result = {
        "metric": METRIC,
        "value": round(GN_ITERS / (head["ms_per_step"] * 1e-3), 2),
        "unit": "GN iters/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head["mode"] == "fp32" else "f16-volume/f32/f64",
        "data": "synthetic (seeded frame graph, random features)",
        "config": {"workload": f"{head['workload']}: 64 KF, 384 edges, 64x80, 128-ch, 4-level r=3 lookup + "
                               f"{GN_ITERS} GN iters (FP64 Cholesky)", "mode": head["mode"]},
    }
    result.update({k: v for k, v in summary(head).items() if k != "workload"})
    if "fp32" in runs and "mixed" in runs:
        f, m, fs = runs["fp32"], runs["mixed"], runs["fp32_sim3"]
        result["mixed"] = dict(summary(m), **{
            "dtype": "f16 volume (f16 MFMA build, f32 accumulation), f32 Jacobians/Hessian, f64 Cholesky + retraction",
            "speedup_vs_fp32": round(f["ms_per_step"] / m["ms_per_step"], 3),
            "speedup_vs_fp32_same_poses": round(fs["ms_per_step"] / m["ms_per_step"], 3),
            "speedup_vs_fp32_with_build": round((f["ms_per_step"] + f["build_ms"]) / (m["ms_per_step"] + m["build_ms"]),
                                                3),
        })
        result["fp32_sim3"] = summary(fs)
    if not a.no_accuracy:
        cache = {}
        for name, run in runs.items():
            acc = accuracy(run, cache)
            if run is head:
                result.update(acc)
            elif name in result:
                result[name].update(acc)
    if head["build_variants"] is not None:
        result["corr_build_variants"] = head["build_variants"]
    result["wall_s"] = round(head["wall"], 3)
    result["roofline"] = {
        "kernel": head["lookup_kernel"],
        "bound": "hbm",
        "achieved": round(lk_gbs / 1e9, 1),
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": round(lk_gbs / HBM_PEAK, 4),
        "traffic": profiled_traffic(head["lookup_kernel"]),
    }
    result["roofline_ba"] = {
        "kernel": "one BA GN iteration (all its kernels)",
        "bound": "hbm",
        "achieved": round(ba_gbs / 1e9, 1),
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": round(ba_gbs / HBM_PEAK, 4),
        "algorithmic_bytes": ba_alg,
    }
    result["roofline_adjoint"] = {
        "kernel": "one BA GN iteration + its implicit adjoint (vh_ba_adjoint), SE3 workload",
        "bound": "hbm",
        "ms": round(adj_times["forward_adjoint"], 4),
        "forward_ms": round(adj_times["forward"], 4),
        "adjoint_ms": round(adj_times["adjoint"], 4),
        "achieved": round(adj_gbs / 1e9, 1),
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": round(adj_gbs / HBM_PEAK, 4),
        "algorithmic_bytes": adj_alg,
    }
    if not a.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(runs.get("fp32", head)["graph"])
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
        e0, e1 = _events()
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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _generic_inputs(g):
    """The graph in the generic path's layout: targets / weights [E, HW, 2] and
    the per-frame damping whose 0.2 x + 1e-7 is the fused path's eta."""
    N, H, W = g.disps.shape
    E = g.ii.shape[0]
    tgt = g.targets.permute(0, 2, 3, 1).reshape(E, H * W, 2).contiguous()
    wt = g.weights.permute(0, 2, 3, 1).reshape(E, H * W, 2).contiguous()
    damp = torch.zeros(N, H, W)
    kx = torch.unique(torch.cat([torch.arange(g.t0, g.t1), g.ii]))
    damp[kx] = (g.eta - 1e-7) / 0.2
    return tgt, wt, damp


def cpu_baseline(g, sample_edges=8):
    """The reference's CPU path restated in float32 torch on the host cores
    (oracle/generic_ba.py: the generic Solver with scipy's sparse LU; the
    CorrBlock volume as a matmul, avg_pool2d pyramid and bilinear lookup).
    ``value``: GN iterations/s of the 64-KF step (one 4-level lookup over all
    384 edges + 8 GN iterations, pyramid built outside the step as on the
    GPU), from a bounded sample: one full GN iteration and the lookup of
    ``sample_edges`` edges scaled to 384.  ``kf8_e24_48x64``: the 8-KF step
    (lookup + 2 GN iterations, and its pyramid build) timed in full."""
    from oracle import corr as OC
    from oracle import generic_ba as G
    from vipe_amd import synth

    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    f32 = torch.float32

    def gn(graph, iters):
        tgt, wt, damp = _generic_inputs(graph)
        t0 = time.perf_counter()
        G.bundle_adjustment(graph.poses.clone(), graph.disps.clone(), graph.intrinsics.clone(), graph.disps_sens, tgt,
                            wt, damp, graph.ii, graph.jj, graph.t0, graph.t1, iters, 1e-4, 0.1, dtype=f32)
        return time.perf_counter() - t0

    def lookup(graph, sel):
        f1 = graph.fmaps[graph.ii[sel]][None]
        f2 = graph.fmaps[graph.jj[sel]][None]
        t0 = time.perf_counter()
        pyr = OC.corr_pyramid(OC.corr_volume(f1, f2, f32), 4)
        t1 = time.perf_counter()
        OC.corr_block_lookup(pyr, graph.coords[sel][None], 3, f32)
        return t1 - t0, time.perf_counter() - t1

    E = g.ii.shape[0]
    t_gn = gn(g, 1)
    _, t_lk = lookup(g, torch.arange(sample_edges))
    step = GN_ITERS * t_gn + t_lk * E / sample_edges
    g0 = synth.make_graph("kf8_e24_48x64", seed=0)
    b0, l0 = lookup(g0, torch.arange(g0.ii.shape[0]))
    t0_gn = gn(g0, 2)
    return {"value": round(GN_ITERS / step, 4), "unit": "GN iters/s", "cores": threads, "kind": "port",
            "sample": f"generic-path restatement, float32 torch + scipy spsolve on {_cpu_model()}: 1 GN iteration "
                      f"of the 64-KF graph ({t_gn:.2f} s) x {GN_ITERS} + lookup of {sample_edges} of {E} edges "
                      f"({t_lk:.2f} s) scaled to {E}",
            "kf8_e24_48x64": {"ms_per_step": round(1e3 * (l0 + t0_gn), 1), "gn_iters_per_s": round(2 / (l0 + t0_gn), 3),
                              "lookup_ms": round(1e3 * l0, 1), "gn_ms_per_iter": round(1e3 * t0_gn / 2, 1),
                              "corr_build_ms": round(1e3 * b0, 1),
                              "what": "8 KF, 24 edges, 48x64: 4-level r=3 lookup over all edges + 2 GN iterations, "
                                      "timed in full"}}


if __name__ == "__main__":
    main()
