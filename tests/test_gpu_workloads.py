"""Parity at the sizes of the measured and documented workloads, and the
bookkeeping edge cases of the fused BA, on the GPU against the FP64 oracle:

- the 64-KF / 384-edge / 64x80 / 128-channel lookup of the measured step
  (tiled volume, 4 levels, radius 3) in FP32 and FP16, on the first, middle
  and last edge at every level;
- the 32-KF / 192-edge / 48x64 / 128-channel graph: the lookup's gradients to
  both feature maps on sampled edges (oracle autograd and central
  differences), and the implicit adjoint of one BA iteration (central
  differences of the oracle);
- the 16-clip batch at its real windows (8 to 96 frames) and resolutions,
  sampled clips against the oracle;
- stereo edges (ii == jj) passed straight to ba_extended
  (geom_kernels.cu:231-241);
- the depth-frame list of the plan against
  unique(cat(arange(t0, t1), ii), sorted) (geom_kernels.cu:1648-1652).
"""
import pytest
import torch

from oracle import ba as OB
from oracle import corr as OC
from vipe_amd import synth
from vipe_amd.synth import make_graph, reproject

pytestmark = pytest.mark.gpu

F64 = torch.float64


def _sample(E):
    return torch.tensor([0, E // 2, E - 1])


def _cuda(g):
    return {k: (v.cuda() if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}


@pytest.mark.parametrize("vol_dt,tol", [(torch.float32, 1e-5), (torch.float16, 2e-3)])
def test_measured_lookup_matches_oracle_on_sampled_edges(vol_dt, tol):
    """The whole 384-edge pyramid is built and looked up (the FP32 level 0 alone
    is 40 GB, addressed through 64-bit wave bases); edges 0, 192 and 383 are
    compared at each of the 4 levels, relative to the level's largest value.
    The FP16 oracle sees the features rounded to FP16, as the device does."""
    from vipe_amd.corr import CorrBlock

    g = make_graph("kf64_e384_64x80", seed=0)
    E = g.ii.shape[0]
    sel = _sample(E)
    f1 = g.fmaps[g.ii][None].cuda()
    f2 = g.fmaps[g.jj][None].cuda()
    if vol_dt == torch.float16:
        f1, f2 = f1.half(), f2.half()
    blk = CorrBlock(f1, f2, num_levels=4, radius=3, volume_dtype=vol_dt)
    del f1, f2
    out = blk(g.coords[None].cuda())
    assert out.shape == (1, E, 4 * 49, 64, 80)
    got = out[0, sel.cuda()].double().cpu()
    del out, blk
    torch.cuda.empty_cache()
    fa, fb = g.fmaps[g.ii[sel]][None], g.fmaps[g.jj[sel]][None]
    if vol_dt == torch.float16:
        fa, fb = fa.half(), fb.half()
    ref = OC.corr_block_lookup(OC.corr_pyramid(OC.corr_volume(fa.double(), fb.double()), 4),
                               g.coords[sel][None].double(), 3)[0]
    for lvl in range(4):
        a, b = got[:, 49 * lvl:49 * (lvl + 1)], ref[:, 49 * lvl:49 * (lvl + 1)]
        scale = b.abs().max().item()
        err = (a - b).abs().max().item()
        print(f"[lookup {vol_dt} level {lvl}] max err {err:.3e} (max |ref| {scale:.3e})")
        assert err <= tol * scale


def test_feature_gradients_at_32_keyframes():
    """32 KF / 192 edges / 48x64 / 128 channels: the backward of the lookup to
    the per-edge feature maps, with a random upstream gradient over all edges,
    equals FP64 autograd through the oracle on edges 0, 96 and 191 (the edges
    are independent) and agrees with central differences of the oracle along
    8 random directions supported on those edges."""
    from vipe_amd.corr import CorrBlock

    g = make_graph("kf32_e192_48x64", seed=0)
    E = g.ii.shape[0]
    sel = _sample(E)
    a1 = g.fmaps[g.ii][None].cuda().requires_grad_(True)
    a2 = g.fmaps[g.jj][None].cuda().requires_grad_(True)
    gen = torch.Generator(device="cuda").manual_seed(17)
    out = CorrBlock(a1, a2, num_levels=4, radius=3)(g.coords[None].cuda())
    go = torch.randn(out.shape, generator=gen, device="cuda")
    (out * go).sum().backward()
    d1 = a1.grad[0, sel.cuda()].double().cpu()
    d2 = a2.grad[0, sel.cuda()].double().cpu()
    go_s = go[:, sel.cuda()].double().cpu()
    del out, go, a1, a2
    torch.cuda.empty_cache()
    coords = g.coords[sel][None].double()

    def loss(x1, x2):
        o = OC.corr_block_lookup(OC.corr_pyramid(OC.corr_volume(x1, x2), 4), coords, 3)
        return (o * go_s).sum()

    f1 = g.fmaps[g.ii[sel]][None].double()
    f2 = g.fmaps[g.jj[sel]][None].double()
    r1, r2 = f1.clone().requires_grad_(True), f2.clone().requires_grad_(True)
    loss(r1, r2).backward()
    scale = max(r1.grad.abs().max().item(), r2.grad.abs().max().item())
    e1 = (d1 - r1.grad[0]).abs().max().item()
    e2 = (d2 - r2.grad[0]).abs().max().item()
    print(f"[feature grads 32 KF] max err fmap1 {e1:.3e} fmap2 {e2:.3e} (max |grad| {scale:.3e})")
    assert e1 < 1e-5 * scale and e2 < 1e-5 * scale
    hg = torch.Generator().manual_seed(23)
    eps = 1e-6
    errs = []
    for _ in range(8):
        v1 = torch.randn(f1.shape, generator=hg, dtype=F64)
        v2 = torch.randn(f2.shape, generator=hg, dtype=F64)
        with torch.no_grad():
            fd = (loss(f1 + eps * v1, f2 + eps * v2) - loss(f1 - eps * v1, f2 - eps * v2)).item() / (2 * eps)
        an = ((d1 * v1[0]).sum() + (d2 * v2[0]).sum()).item()
        gv = (d1.norm() ** 2 + d2.norm() ** 2).sqrt().item() * (v1.norm() ** 2 + v2.norm() ** 2).sqrt().item()
        errs.append(abs(fd - an) / gv)
    print(f"[feature grads 32 KF] max normalised directional error over 8 directions: {max(errs):.3e}")
    assert max(errs) < 1e-5


def test_ba_adjoint_at_32_keyframes():
    """32 KF / 192 edges / 48x64: dL/d(targets, weights) of one BA iteration
    through the implicit adjoint against central differences (eps = 1e-6) of
    the FP64 oracle along 12 random directions."""
    from vipe_amd.ba_layer import ba_iteration

    g = make_graph("kf32_e192_48x64", with_fmaps=False)
    N = g.disps.shape[0]
    da = torch.ones(N)
    ds = g.disps.clone() * 1.01
    ds[:, ::2] = 0.0
    lm, ep, alpha = 1e-4, 0.1, 0.05
    gen = torch.Generator().manual_seed(29)
    d = _cuda(g)
    tg = d["targets"].clone().requires_grad_(True)
    wt = d["weights"].clone().requires_grad_(True)
    dx, dz = ba_iteration(tg, wt, d["poses"], d["disps"], d["intrinsics"], ds.cuda(), d["eta"], d["ii"], d["jj"],
                          da.cuda(), g.t0, g.t1, lm, ep, False, alpha)
    Y = dx.detach().cpu().double() + 0.01 * torch.randn(dx.shape, generator=gen, dtype=F64)
    Z = dz.detach().cpu().double() + 0.01 * torch.randn(dz.shape, generator=gen, dtype=F64)
    loss = ((dx - Y.float().cuda()) ** 2).sum() + ((dz - Z.float().cuda()) ** 2).sum()
    loss.backward()
    gt_, gw_ = tg.grad.double().cpu(), wt.grad.double().cpu()

    def loss64(t64, w64):
        rdx, rdz, _ = OB.ba_extended(g.poses.clone(), g.disps.clone(), g.intrinsics.clone(), ds, t64, w64, g.eta,
                                     g.ii, g.jj, da, g.t0, g.t1, 1, lm, ep, False, alpha, False, 1e-6, 1e-6, 1.0,
                                     False)
        return (((rdx - Y) ** 2).sum() + ((rdz - Z) ** 2).sum()).item()

    t64, w64 = g.targets.double(), g.weights.double()
    eps = 1e-6
    errs = []
    for _ in range(12):
        vt = torch.randn(t64.shape, generator=gen, dtype=F64)
        vw = torch.randn(w64.shape, generator=gen, dtype=F64)
        fd = (loss64(t64 + eps * vt, w64 + eps * vw) - loss64(t64 - eps * vt, w64 - eps * vw)) / (2 * eps)
        an = ((gt_ * vt).sum() + (gw_ * vw).sum()).item()
        errs.append(abs(fd - an) / max(abs(fd), 1e-6))
    print(f"[BA adjoint 32 KF] max rel directional error over 12 directions: {max(errs):.3e}")
    assert max(errs) < 2e-3


def test_sixteen_clip_batch_matches_oracle_on_sampled_clips():
    """The 16-clip uneven batch (windows 8..96, out-degree 2..12, three
    resolutions), 4 GN iterations over 4 streams; the smallest clip and the two
    largest (96 frames, 940 and 1152 edges) against the FP64 oracle within
    1e-4 px reprojection and 1e-5 relative pose error."""
    from vipe_amd.batched import BatchedBA, Clip

    iters = 4
    gs = synth.make_batch(16, 0)
    clips = [Clip.from_graph(g, "cuda") for g in gs]
    b = BatchedBA(clips, iterations=iters, n_streams=4)
    b.run()
    torch.cuda.synchronize()
    sizes = [g.ii.shape[0] * g.disps[0].numel() for g in gs]
    order = sorted(range(16), key=lambda n: sizes[n])
    for n in (order[0], order[-2], order[-1]):
        g, c = gs[n], clips[n]
        N = g.disps.shape[0]
        rp, rd = g.poses.clone(), g.disps.clone()
        OB.ba_extended(rp, rd, g.intrinsics.clone(), g.disps_sens, g.targets, g.weights, g.eta, g.ii, g.jj,
                       torch.ones(N), g.t0, g.t1, iters, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, False)
        p, dd = c.poses.cpu(), c.disps.cpu()
        ca, _ = reproject(p, dd, g.intrinsics, g.ii, g.jj)
        cb, _ = reproject(rp, rd, g.intrinsics, g.ii, g.jj)
        px = (ca - cb).abs().max().item()
        rel = ((p.double() - rp.double()).norm(dim=1) / rp.double().norm(dim=1)).max().item()
        print(f"[batch clip {n}: {N} frames, {g.ii.shape[0]} edges, {tuple(g.disps.shape[1:])}] "
              f"px err {px:.3e}, relative pose err {rel:.3e}")
        assert px < 1e-4 and rel < 1e-5, n


@pytest.mark.parametrize("group", ["se3", "sim3"])
def test_stereo_edges_passed_straight_to_ba_extended(group):
    """Edges with ii == jj (one inside the window, one on the fixed frame 0)
    project through T = (t = (-0.1, 0, 0), R = I) and add depth terms but no
    pose terms; the device matches the oracle after 2 iterations."""
    from vipe_amd.ext import slam_ext

    g = make_graph("kf8_e24_48x64" if group == "se3" else "kf8_e24_48x64_sim3", with_fmaps=False)
    N, H, W = g.disps.shape
    st = torch.tensor([3, 0])
    gen = torch.Generator().manual_seed(31)
    vv, uu = torch.meshgrid(torch.arange(H, dtype=F64), torch.arange(W, dtype=F64), indexing="ij")
    tgt = []
    for k in st.tolist():
        # X = ((u-cx)/fx, (v-cy)/fy, 1, d) moved by t = (-0.1, 0, 0)
        fx, fy, cx, cy = g.intrinsics.double().tolist()
        u = uu + fx * (-0.1) * g.disps[k].double()
        tgt.append(torch.stack([u, vv], 0) + 0.3 * torch.randn(2, H, W, generator=gen, dtype=F64))
    g.ii = torch.cat([g.ii, st])
    g.jj = torch.cat([g.jj, st])
    g.targets = torch.cat([g.targets, torch.stack(tgt).float()]).contiguous()
    g.weights = torch.cat([g.weights, torch.rand(2, 2, H, W, generator=gen)]).contiguous()
    kx = torch.unique(torch.cat([torch.arange(g.t0, g.t1), g.ii]))
    assert kx.shape[0] == g.eta.shape[0]  # frame 0 already has outgoing edges
    d = _cuda(g)
    da = torch.ones(N)
    iters = 2
    p, dd = d["poses"].clone(), d["disps"].clone()
    dx, dz, en = slam_ext.ba_extended(p, dd, d["intrinsics"].clone(), d["disps_sens"], d["targets"], d["weights"],
                                      d["eta"], d["ii"], d["jj"], da.cuda(), g.t0, g.t1, iters, 1e-4, 0.1, False,
                                      0.05, False, 1e-6, 1e-6, 1.0, True)
    rp, rd = g.poses.clone(), g.disps.clone()
    rdx, rdz, ren = OB.ba_extended(rp, rd, g.intrinsics.clone(), g.disps_sens, g.targets, g.weights, g.eta, g.ii,
                                   g.jj, da, g.t0, g.t1, iters, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, True)
    torch.cuda.synchronize()
    assert (dx.cpu().double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert (dz.cpu().double() - rdz).abs().max().item() < 1e-4
    assert (p.cpu().double() - rp.double()).abs().max().item() < 1e-5
    assert (dd.cpu().double() - rd.double()).abs().max().item() < 1e-4
    assert torch.allclose(en.cpu().double(), ren, rtol=1e-4, atol=1e-6)
    # the stereo edges change the depth of their frames
    without = slam_ext.ba_extended(d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone(), d["disps_sens"],
                                   d["targets"][:-2].contiguous(), d["weights"][:-2].contiguous(), d["eta"],
                                   d["ii"][:-2], d["jj"][:-2], da.cuda(), g.t0, g.t1, 1, 1e-4, 0.1, False, 0.05,
                                   False, 1e-6, 1e-6, 1.0, False)[1]
    k3 = int((kx == 3).nonzero())
    assert (without[k3] - dz[k3]).abs().max().item() > 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_plan_depth_frames_equal_sorted_unique(seed):
    """The plan's depth frames are unique(cat(arange(t0, t1), ii)) in ascending
    order, bit for bit, with duplicate, stereo and out-of-window edges."""
    from vipe_amd.ext import slam_ext

    gen = torch.Generator().manual_seed(seed)
    N, t0, t1 = 40, 7, 19
    E = 90
    ii = torch.randint(0, N, (E,), generator=gen)
    jj = torch.randint(0, N, (E,), generator=gen)
    ii[:6] = ii[6:12]  # duplicated edges
    jj[:6] = jj[6:12]
    ii[12:15] = jj[12:15]  # stereo edges
    ii[15:20] = torch.tensor([0, 1, 2, 39, 38])  # sources outside the window
    plan = slam_ext.get_plan(ii.cuda(), jj.cuda(), N, 8, 8, t0, t1)
    expect = torch.unique(torch.cat([torch.arange(t0, t1), ii]), sorted=True)
    assert plan.K == expect.numel()
    assert torch.equal(plan.kx, expect)
