"""rocprofv3 helpers for the kernel summaries.

    python3 -m vipe_amd.profiling copy [--mib 1024]
        A plain device-to-device copy of known size (3 launches).  Profiled with
        ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE``, it gives the factor between
        the counters' KiB and real bytes on this GPU.

    python3 -m vipe_amd.profiling stream [--mib 1024] [--iters 10]
        What plain streaming reaches on this GPU: torch's read-only sum and its
        device copy (read + write) of an FP32 buffer, timed with HIP events; the
        reference points for the HBM figures of the kernels.

    python3 -m vipe_amd.profiling build [--contraction f32|f16|bf16|fp8]
        Two correlation-pyramid builds of the bench workload (for counters).

    python3 -m vipe_amd.profiling spatial_corr [--dtype f32|f16|bf16] [--iters 5]
        Forward and backward of the spatial correlation sampler on the
        local-attention shape: [8, 256, 30, 54] inputs, kernel 1, a 15 x 15 patch.

    python3 -m vipe_amd.profiling scatter [--dtype f32|f16|bf16] [--iters 5]
        scatter_sum and scatter_max of [1, 384, 128 * 48 * 64] sources into 64
        slots along dim 1 by a 1-D index (the per-frame aggregation shape).

    python3 -m vipe_amd.profiling altcorr [--dtype f32|f16] [--iters 5] [--workload NAME]
        AltCorrBlock lookups (on-the-fly correlation, 4 levels, r = 3) of the
        bench workload's edges over its feature pyramid, one call per iteration
        after a warm-up.

    python3 -m vipe_amd.profiling altcorr_layer [--iters 3]
        CorrLayer (altcorr forward and backward, one level, r = 3) on 16 pairs of
        128-channel 64x80 FP32 feature maps, coordinates the identity plus
        N(0, 2 px) noise.

    python3 -m vipe_amd.profiling corr_index [--dtype f32|f16] [--edges 96] [--iters 5]
        CorrSampler's lookup (corr_index_forward, r = 3) into a plain level-0
        volume [edges, 64, 80, 64, 80] (random values), coordinates the
        identity plus N(0, 2 px) noise.

    python3 -m vipe_amd.profiling update_op [--edges 384] [--per-frame 6] [--iters 3]
        The DroidNet update operator (UpdateModule.forward with GraphAgg) on
        the bench shape: 384 edges at 64 x 80, six edges per frame (64 frames),
        random FP16 state and context, random parameters; one warm-up call,
        then `iters` calls.

    python3 -m vipe_amd.profiling ba_rig [--camera mei|pinhole] [--views 2] [--iters 8]
        The fused BA on the two-view rig version of the bench graph
        (synth.make_rig_graph: 64 frames x 2 views, 894 edges at 64 x 80):
        one warm-up call, then one call of `iters` Gauss-Newton iterations,
        timed with HIP events.

    python3 -m vipe_amd.profiling summarize --stats DIR [--fetch DIR --write DIR
        --cfetch DIR --cwrite DIR --copy-mib 1024] [--kernel SUBSTR ...]
        Reads the CSV output of ``rocprofv3 --output-format csv`` runs and prints
        a markdown summary: per-kernel calls / average / total time, and for the
        named kernels the HBM bytes per dispatch from FETCH_SIZE and WRITE_SIZE,
        each scaled by the copy calibration.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys
from collections import defaultdict


def _copy(mib: int) -> None:
    import torch

    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy: {n * 4} bytes read and {n * 4} bytes written per launch", flush=True)


def _stream(mib: int, iters: int) -> None:
    import torch

    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, fn, nbytes in (("read (sum)", lambda: a.sum(), n * 4), ("copy (read + write)", lambda: b.copy_(a),
                                                                    2 * n * 4)):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(iters):
            fn()
        e1.record()
        torch.cuda.synchronize()
        s = e0.elapsed_time(e1) * 1e-3 / iters
        print(f"{name}: {nbytes / s / 1e12:.2f} TB/s ({nbytes} bytes in {s * 1e6:.1f} us)", flush=True)


def _build(contraction: str, workload: str) -> None:
    """Two pyramid builds (one warm-up) of the bench workload's 384 edges with the
    given MFMA contraction type; output levels in f32 (f32 contraction) or f16."""
    import torch

    from . import synth
    from .corr import CorrBlock

    dev = torch.device("cuda", 0)
    g = synth.make_graph(workload, seed=0)
    fm = g.fmaps.to(dev)
    f1 = fm[g.ii.to(dev)].contiguous()[None]
    f2 = fm[g.jj.to(dev)].contiguous()[None]
    ctype = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16, "fp8": "fp8"}[contraction]
    out = torch.float32 if contraction == "f32" else torch.float16
    for _ in range(2):
        c = CorrBlock(f1, f2, num_levels=4, radius=3, volume_dtype=out, contraction_dtype=ctype)
        del c
        torch.cuda.synchronize()
    print(f"build: {g.ii.numel()} edges, contraction {contraction}, levels {out}", flush=True)


def _find(d: str, suffix: str) -> str:
    hits = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[0]


def kernel_stats(d: str):
    rows = []
    with open(_find(d, "kernel_stats.csv")) as f:
        for r in csv.DictReader(f):
            rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                         float(r["Percentage"])))
    return rows


def counter_per_kernel(d: str, counter: str):
    """Mean counter value per dispatch, grouped by kernel name."""
    acc = defaultdict(list)
    with open(_find(d, "counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


_DT = {"f32": "float32", "f16": "float16", "bf16": "bfloat16"}


def _spatial_corr(dtype: str, iters: int) -> None:
    import torch

    from .ext import corr_ext

    dt = getattr(torch, _DT[dtype])
    g = torch.Generator(device="cuda").manual_seed(0)
    q = torch.randn(8, 256, 30, 54, device="cuda", generator=g).to(dt)
    k = torch.randn(8, 256, 30, 54, device="cuda", generator=g).to(dt)
    args = (1, 1, 15, 15, 0, 0, 1, 1, 1, 1, 1, 1)
    go = torch.randn(8, 15, 15, 30, 54, device="cuda", generator=g).to(dt)
    for _ in range(iters):
        corr_ext.forward(q, k, *args)
        corr_ext.backward(q, k, go, *args)
    torch.cuda.synchronize()
    print(f"spatial_corr: {iters} forward + backward, [8, 256, 30, 54] {dtype}, kernel 1, patch 15", flush=True)


def _scatter(dtype: str, iters: int) -> None:
    import torch

    from .ext import scatter_ext

    dt = getattr(torch, _DT[dtype])
    g = torch.Generator(device="cuda").manual_seed(0)
    src = torch.randn(1, 384, 128 * 48 * 64, device="cuda", generator=g).to(dt)
    idx = torch.randint(0, 64, (384,), device="cuda", generator=g)
    for _ in range(iters):
        scatter_ext.scatter_sum(src, idx, 1, None, 64)
        scatter_ext.scatter_max(src, idx, 1, None, 64)
    torch.cuda.synchronize()
    print(f"scatter: {iters} x (sum, max), src [1, 384, 393216] {dtype} -> 64 slots", flush=True)


def _altcorr(dtype: str, iters: int, workload: str) -> None:
    import torch

    from . import synth
    from .corr import AltCorrBlock

    g = synth.make_graph(workload, seed=0)
    dt = getattr(torch, _DT[dtype])
    fm = g.fmaps.cuda().to(dt)[None]  # [1, N, C, H, W]
    ii, jj = g.ii.cuda(), g.jj.cuda()
    coords = g.coords.cuda()[None]  # [1, E, H, W, 2]
    blk = AltCorrBlock(fm, num_levels=4, radius=3)
    out = blk(coords, ii, jj)
    torch.cuda.synchronize()
    for _ in range(iters):
        out = blk(coords, ii, jj)
    torch.cuda.synchronize()
    E, H, W = coords.shape[1:4]
    print(f"altcorr: {iters} x lookup of {E} edges at {H}x{W}, C={fm.shape[2]} {dtype} -> {tuple(out.shape)}",
          flush=True)


def _altcorr_layer(iters: int) -> None:
    import torch

    from .corr import CorrLayer

    g = torch.Generator(device="cuda").manual_seed(0)
    B, H, W, C = 16, 64, 80, 128
    f1 = (torch.randn(B, H, W, C, device="cuda", generator=g) / C ** 0.5).requires_grad_()
    f2 = (torch.randn(B, H, W, C, device="cuda", generator=g) / C ** 0.5).requires_grad_()
    yy, xx = torch.meshgrid(torch.arange(H, device="cuda"), torch.arange(W, device="cuda"), indexing="ij")
    coords = torch.stack([xx, yy], -1).float()[None, None].repeat(B, 1, 1, 1, 1)
    coords = coords + 2.0 * torch.randn(coords.shape, device="cuda", generator=g)
    for _ in range(iters + 1):
        out = CorrLayer.apply(f1, f2, coords, 3)
        out.sum().backward()
    torch.cuda.synchronize()
    print(f"altcorr_layer: {iters + 1} x forward + backward, [{B}, {H}, {W}, {C}] f32 -> {tuple(out.shape)}",
          flush=True)


def _corr_index(dtype: str, edges: int, iters: int) -> None:
    import torch

    from .ext import droid_net_ext

    dt = getattr(torch, _DT[dtype])
    g = torch.Generator(device="cuda").manual_seed(0)
    H, W = 64, 80
    vol = torch.randn(edges, H, W, H, W, device="cuda", generator=g, dtype=dt)
    yy, xx = torch.meshgrid(torch.arange(H, device="cuda"), torch.arange(W, device="cuda"), indexing="ij")
    coords = torch.stack([xx, yy], 0).float()[None].repeat(edges, 1, 1, 1)
    coords = (coords + 2.0 * torch.randn(coords.shape, device="cuda", generator=g)).contiguous()
    for _ in range(iters + 1):
        (out,) = droid_net_ext.corr_index_forward(vol, coords, 3)
    torch.cuda.synchronize()
    print(f"corr_index: {iters + 1} x lookup of [{edges}, {H}, {W}, {H}, {W}] {dtype} -> {tuple(out.shape)}", flush=True)


def _update_op(edges: int, per_frame: int, iters: int) -> None:
    import torch

    from .update_op import UpdateModule

    torch.manual_seed(0)
    m = UpdateModule().cuda()
    g = torch.Generator(device="cuda").manual_seed(0)
    H, W = 64, 80
    net = torch.tanh(torch.randn(1, edges, 128, H, W, device="cuda", generator=g)).half()
    inp = torch.relu(torch.randn(1, edges, 128, H, W, device="cuda", generator=g)).half()
    corr = torch.randn(1, edges, 196, H, W, device="cuda", generator=g) * 0.5
    flow = (torch.randn(1, edges, 4, H, W, device="cuda", generator=g) * 4).clamp(-64, 64)
    ix = torch.arange(edges, device="cuda") // per_frame
    for _ in range(iters + 1):
        out = m(net, inp, corr, flow, ix)
    torch.cuda.synchronize()
    print(f"update_op: {iters + 1} x forward, {edges} edges ({per_frame} per frame) at {H}x{W} -> "
          f"net {tuple(out[0].shape)}, upmask {tuple(out[4].shape)}", flush=True)


def _feature_grad(workload: str, iters: int) -> None:
    """CorrBlock built from features that need gradients: the fused lookup and its
    backward to both feature maps (altcorr backward per level, NHWC pyramid and
    gradient combination kernels), timed with HIP events over ``iters`` calls."""
    import torch

    from . import synth
    from .corr import CorrBlock

    g = synth.make_graph(workload, seed=0)
    dev = torch.device("cuda", 0)
    f1 = g.fmaps[g.ii][None].to(dev).requires_grad_(True)
    f2 = g.fmaps[g.jj][None].to(dev).requires_grad_(True)
    coords = g.coords[None].to(dev)
    gen = torch.Generator(device=dev).manual_seed(0)
    go = None

    def step():
        nonlocal go
        blk = CorrBlock(f1, f2, num_levels=4, radius=3)
        out = blk(coords)
        if go is None:
            go = torch.randn(out.shape, device=dev, generator=gen)
        t0 = torch.cuda.Event(enable_timing=True)
        t0.record()
        gf1, gf2 = torch.autograd.grad(out, (f1, f2), go)
        t1 = torch.cuda.Event(enable_timing=True)
        t1.record()
        return t0, t1, gf1, gf2

    step()
    torch.cuda.synchronize()
    tb = 0.0
    for _ in range(iters):
        t0, t1, _, _ = step()
        torch.cuda.synchronize()
        tb += t0.elapsed_time(t1)
    E, C, H, W = g.fmaps[g.ii].shape
    print(f"feature_grad: {workload}: {E} edges, C={C}, {H}x{W}: backward to both feature maps "
          f"{tb / iters:.3f} ms per call ({iters} calls)", flush=True)


def _ba_rig(camera: str, views: int, iters: int) -> None:
    import torch

    from . import synth
    from .ext import slam_ext

    g = synth.make_rig_graph("kf64_e384_64x80", camera=camera, views=views)
    d = {k: (v.cuda() if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}
    act = torch.ones(d["disps"].shape[0], device="cuda")

    def run(n):
        p, dd = d["poses"].clone(), d["disps"].clone()
        return slam_ext.ba_rig(p, dd, d["intrinsics"], camera, d["rig"], d["disps_sens"], d["targets"],
                               d["weights"], d["eta"], d["pi"], d["pj"], d["qi"], d["qj"], act, g.t0, g.t1, n, 1e-4,
                               0.1, False, 0.05, True)

    run(1)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    _, _, en = run(iters)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e)
    print(f"ba_rig {camera}: {views} views, {g.pi.shape[0]} edges, {iters} GN iterations in {ms:.3f} ms "
          f"({ms / iters:.4f} ms per iteration); energy {en[0].item():.4g} -> {en[-1].item():.4g}", flush=True)


def _short(name: str, width: int = 90) -> str:
    return name if len(name) <= width else name[: width - 3] + "..."


def _summarize(a) -> None:
    out = []
    stats = kernel_stats(a.stats)
    total = sum(r[2] for r in stats)
    out.append("| kernel | calls | avg (us) | total (ms) | % |")
    out.append("|---|---:|---:|---:|---:|")
    for name, calls, tot, avg, pct in sorted(stats, key=lambda r: -r[2])[: a.top]:
        out.append(f"| `{_short(name)}` | {calls} | {avg / 1e3:.2f} | {tot / 1e6:.3f} | {pct:.1f} |")
    out.append(f"\nTotal kernel time: {total / 1e6:.3f} ms\n")
    if a.fetch and a.write and a.cfetch and a.cwrite:
        copy_bytes = a.copy_mib * 1024 * 1024
        cf = counter_per_kernel(a.cfetch, "FETCH_SIZE")
        cw = counter_per_kernel(a.cwrite, "WRITE_SIZE")
        ck = max(cf, key=cf.get)
        f_scale = copy_bytes / cf[ck]
        w_scale = copy_bytes / cw[max(cw, key=cw.get)]
        out.append(f"Copy calibration ({a.copy_mib} MiB copy, `{_short(ck, 60)}`): FETCH_SIZE reads "
                   f"{cf[ck]:.0f}, WRITE_SIZE {cw[max(cw, key=cw.get)]:.0f} per dispatch; "
                   f"bytes per counter unit: fetch {f_scale:.2f}, write {w_scale:.2f}.\n")
        fe = counter_per_kernel(a.fetch, "FETCH_SIZE")
        wr = counter_per_kernel(a.write, "WRITE_SIZE")
        out.append("| kernel | FETCH_SIZE/dispatch | WRITE_SIZE/dispatch | HBM bytes/dispatch (scaled) |")
        out.append("|---|---:|---:|---:|")
        lines = []
        for k in sorted(fe, key=lambda k: -fe[k]):
            if a.kernel and not any(s in k for s in a.kernel):
                continue
            byts = fe[k] * f_scale + wr.get(k, 0.0) * w_scale
            out.append(f"| `{_short(k)}` | {fe[k]:.0f} | {wr.get(k, 0.0):.0f} | {byts:.4e} |")
            lines.append(f"traffic_bytes_per_dispatch {kernel_short_name(k)}: {byts:.4e}")
        out.append("")
        out.extend(lines)
    print("\n".join(out))


def kernel_short_name(name: str) -> str:
    """`void vh::(anonymous namespace)::foo<float, 3>(args)` -> `foo<float, 3>`."""
    for prefix in ("void ", "vh::(anonymous namespace)::", "vh::ba::", "vh::"):
        name = name.replace(prefix, "")
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    st = sub.add_parser("stream")
    st.add_argument("--mib", type=int, default=1024)
    st.add_argument("--iters", type=int, default=10)
    bl = sub.add_parser("build")
    bl.add_argument("--contraction", choices=["f32", "f16", "bf16", "fp8"], default="f32")
    bl.add_argument("--workload", default="kf64_e384_64x80")
    for name in ("spatial_corr", "scatter"):
        w = sub.add_parser(name)
        w.add_argument("--dtype", choices=list(_DT), default="f32")
        w.add_argument("--iters", type=int, default=5)
    ac = sub.add_parser("altcorr")
    ac.add_argument("--dtype", choices=["f32", "f16"], default="f32")
    ac.add_argument("--iters", type=int, default=5)
    ac.add_argument("--workload", default="kf64_e384_64x80")
    al = sub.add_parser("altcorr_layer")
    al.add_argument("--iters", type=int, default=3)
    ci = sub.add_parser("corr_index")
    ci.add_argument("--dtype", choices=["f32", "f16"], default="f32")
    ci.add_argument("--edges", type=int, default=96)
    ci.add_argument("--iters", type=int, default=5)
    br = sub.add_parser("ba_rig")
    br.add_argument("--camera", choices=["mei", "pinhole"], default="mei")
    br.add_argument("--views", type=int, default=2)
    br.add_argument("--iters", type=int, default=8)
    fg = sub.add_parser("feature_grad")
    fg.add_argument("--workload", default="kf32_e192_48x64")
    fg.add_argument("--iters", type=int, default=5)
    uo = sub.add_parser("update_op")
    uo.add_argument("--edges", type=int, default=384)
    uo.add_argument("--per-frame", type=int, default=6)
    uo.add_argument("--iters", type=int, default=3)
    s = sub.add_parser("summarize")
    s.add_argument("--stats", required=True)
    s.add_argument("--fetch")
    s.add_argument("--write")
    s.add_argument("--cfetch")
    s.add_argument("--cwrite")
    s.add_argument("--copy-mib", type=int, default=1024)
    s.add_argument("--kernel", action="append", default=[])
    s.add_argument("--top", type=int, default=25)
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        _copy(a.mib)
    elif a.cmd == "stream":
        _stream(a.mib, a.iters)
    elif a.cmd == "build":
        _build(a.contraction, a.workload)
    elif a.cmd == "spatial_corr":
        _spatial_corr(a.dtype, a.iters)
    elif a.cmd == "scatter":
        _scatter(a.dtype, a.iters)
    elif a.cmd == "altcorr":
        _altcorr(a.dtype, a.iters, a.workload)
    elif a.cmd == "altcorr_layer":
        _altcorr_layer(a.iters)
    elif a.cmd == "corr_index":
        _corr_index(a.dtype, a.edges, a.iters)
    elif a.cmd == "ba_rig":
        _ba_rig(a.camera, a.views, a.iters)
    elif a.cmd == "feature_grad":
        _feature_grad(a.workload, a.iters)
    elif a.cmd == "update_op":
        _update_op(a.edges, a.per_frame, a.iters)
    else:
        _summarize(a)


if __name__ == "__main__":
    main()
