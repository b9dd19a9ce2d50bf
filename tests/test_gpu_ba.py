"""gfx950 dense BA vs the FP64 oracle, plus the reference tests' invariants
(tests/test_slam_ext_ba_extended.py, tests/test_graph_buffer_fused_ba.py)."""
import pytest
import torch

from oracle import ba as OB
from vipe_amd.synth import make_graph, reproject

pytestmark = pytest.mark.gpu


def _cuda(g):
    return {k: (v.cuda() if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}


def _run_pair(name, iters, motion_only=False, intr=False, limited=False, lm=1e-4, ep=0.1, alpha=0.05):
    from vipe_amd.ext import slam_ext

    g = make_graph(name, with_fmaps=False)
    N = g.disps.shape[0]
    da = torch.ones(N)
    if limited:
        da.zero_()
        da[g.t0:g.t1] = 1.0
    # some depth-sensor pixels so the prior term is exercised
    ds = g.disps.clone() * 1.01
    ds[:, ::2] = 0.0
    ref = dict(poses=g.poses.clone(), disps=g.disps.clone(), intr=g.intrinsics.clone())
    d = _cuda(g)
    p, dd, it = d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone()
    dx, dz, en = slam_ext.ba_extended(p, dd, it, ds.cuda(), d["targets"], d["weights"], d["eta"], d["ii"], d["jj"],
                                      da.cuda(), g.t0, g.t1, iters, lm, ep, motion_only, alpha, intr, 1e-6, 1e-6,
                                      0.125, True)
    rdx, rdz, ren = OB.ba_extended(ref["poses"], ref["disps"], ref["intr"], ds, g.targets, g.weights, g.eta, g.ii,
                                   g.jj, da, g.t0, g.t1, iters, lm, ep, motion_only, alpha, intr, 1e-6, 1e-6, 0.125,
                                   True)
    torch.cuda.synchronize()
    return g, (p.cpu(), dd.cpu(), it.cpu(), dx.cpu(), None if dz is None else dz.cpu(), en.cpu()), \
        (ref["poses"], ref["disps"], ref["intr"], rdx, rdz, ren)


def _reproj_err(g, poses_a, disps_a, poses_b, disps_b, intr):
    ca, _ = reproject(poses_a, disps_a, intr, g.ii, g.jj)
    cb, _ = reproject(poses_b, disps_b, intr, g.ii, g.jj)
    return (ca - cb).abs().max().item()


@pytest.mark.parametrize("iters", [1, 2])
def test_ba_matches_oracle_kf8(iters):
    g, mine, ref = _run_pair("kf8_e24_48x64", iters)
    p, d, it, dx, dz, en = mine
    rp, rd, rit, rdx, rdz, ren = ref
    assert (dx.double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert (dz.double() - rdz).abs().max().item() < 1e-4
    assert (p.double() - rp.double()).abs().max().item() < 1e-5
    assert (d.double() - rd.double()).abs().max().item() < 1e-4
    assert _reproj_err(g, p, d, rp, rd, g.intrinsics) < 1e-3
    assert torch.allclose(en.double(), ren, rtol=1e-4, atol=1e-6)


def test_ba_motion_only_leaves_disps_bit_exact():
    g, mine, ref = _run_pair("kf8_e24_48x64", 1, motion_only=True)
    p, d, it, dx, dz, en = mine
    assert dz is None
    assert torch.equal(d, g.disps)
    assert (p.double() - ref[0].double()).abs().max().item() < 1e-5


def test_ba_intrinsics_and_limited_disp():
    g, mine, ref = _run_pair("kf8_e24_48x64", 1, intr=True, limited=True)
    p, d, it, dx, dz, en = mine
    rp, rd, rit, rdx, rdz, ren = ref
    assert (it.double() - rit.double()).abs().max().item() < 1e-4
    assert (it[:2] - g.intrinsics[:2]).abs().max().item() > 0.0
    assert torch.equal(d[0], g.disps[0])  # frame 0 is outside [t0, t1): depth inactive
    assert (p.double() - rp.double()).abs().max().item() < 1e-5
    assert (d.double() - rd.double()).abs().max().item() < 1e-4


def _make_ref_test_inputs(device="cuda"):
    """tests/test_slam_ext_ba_extended.py:15-33 construction (CPU RNG, seed 2026)."""
    torch.manual_seed(2026)
    n_frames, height, width, n_edges = 4, 4, 5, 5
    poses = torch.zeros(n_frames, 7)
    poses[:, 6] = 1.0
    poses[:, 0] = torch.linspace(0.0, 0.03, n_frames)
    disps = torch.ones(n_frames, height, width) * 0.8
    disps_sens = torch.zeros_like(disps)
    intrinsics = torch.tensor([40.0, 40.0, 2.0, 2.0])
    targets = torch.rand(n_edges, 2, height, width) * 3.0
    weights = torch.rand(n_edges, 2, height, width)
    ii = torch.tensor([1, 2, 2, 3, 3])
    jj = torch.tensor([0, 0, 1, 1, 2])
    kx = torch.unique(torch.cat([torch.arange(1, 4), ii]))
    eta = torch.ones(n_frames, height, width)[kx].contiguous() * 0.1
    out = [poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj]
    return [t.to(device) for t in out]


def test_extended_matches_legacy_bit_exact():
    from vipe_amd.ext import slam_ext

    poses0, disps0, intr0, ds, tg, wt, eta, ii, jj = _make_ref_test_inputs()
    poses1, disps1 = poses0.clone(), disps0.clone()
    da = torch.ones(4, device="cuda")
    slam_ext.ba(poses0, disps0, intr0.clone(), ds, tg, wt, eta, ii, jj, 1, 4, 1, 1e-3, 0.1, False, 0.001)
    slam_ext.ba_extended(poses1, disps1, intr0.clone(), ds, tg, wt, eta, ii, jj, da, 1, 4, 1, 1e-3, 0.1, False, 0.001,
                         False, 1e-6, 1e-6, 0.125)
    assert torch.equal(poses1, poses0)
    assert torch.equal(disps1, disps0)


def test_extended_intrinsics_and_masked_frame_like_reference():
    from vipe_amd.ext import slam_ext

    poses, disps, intr, ds, tg, wt, eta, ii, jj = _make_ref_test_inputs()
    d0, i0 = disps.clone(), intr.clone()
    da = torch.ones(4, device="cuda")
    da[1] = 0.0
    slam_ext.ba_extended(poses, disps, intr, ds, tg, wt, eta, ii, jj, da, 1, 4, 1, 1e-3, 0.1, False, 0.001, True,
                         1e-6, 1e-6, 0.125)
    assert torch.isfinite(intr).all()
    assert (intr[:2] - i0[:2]).abs().max().item() > 0.0
    assert torch.equal(disps[1], d0[1])


def test_extended_energy_like_reference_and_oracle():
    from vipe_amd.ext import slam_ext

    poses, disps, intr, ds, tg, wt, eta, ii, jj = _make_ref_test_inputs()
    cpu = [t.cpu().clone() for t in (poses, disps, intr, ds, tg, wt, eta, ii, jj)]
    da = torch.ones(4, device="cuda")
    _, _, en = slam_ext.ba_extended(poses, disps, intr, ds, tg, wt, eta, ii, jj, da, 1, 4, 2, 1e-3, 0.1, False, 0.001,
                                    False, 1e-6, 1e-6, 0.125, True)
    assert tuple(en.shape) == (2,)
    assert torch.isfinite(en).all() and en[0].item() > 0.0
    rdx, rdz, ren = OB.ba_extended(cpu[0], cpu[1], cpu[2], cpu[3], cpu[4], cpu[5], cpu[6], cpu[7], cpu[8],
                                   torch.ones(4), 1, 4, 2, 1e-3, 0.1, False, 0.001, False, 1e-6, 1e-6, 0.125, True)
    assert torch.allclose(en.cpu().double(), ren, rtol=1e-4)
    assert (poses.cpu().double() - cpu[0].double()).abs().max().item() < 1e-5
    assert (disps.cpu().double() - cpu[1].double()).abs().max().item() < 1e-4


def test_failed_factorisation_gives_zero_update():
    """A non-positive-definite reduced system returns dx = 0 (geom_kernels.cu:1283-1285)."""
    from vipe_amd.ext import slam_ext

    poses, disps, intr, ds, tg, wt, eta, ii, jj = _make_ref_test_inputs()
    p0 = poses.clone()
    wt = torch.zeros_like(wt)
    dx, dz, _ = slam_ext.ba(poses, disps, intr, ds, tg, wt, eta, ii, jj, 1, 4, 1, 0.0, -1.0, True)
    assert torch.count_nonzero(dx) == 0
    assert torch.equal(poses, p0)


def test_motion_only_ignores_stale_workspace():
    """Regression: motion-only BA must not read Schur terms left in a reused
    workspace by an earlier full solve."""
    from vipe_amd.ext import slam_ext

    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    d = _cuda(g)
    N = g.disps.shape[0]
    da = torch.ones(N, device="cuda")
    for it in range(3):
        # full solve first (fills the Gram buffer), then motion-only on fresh copies
        slam_ext.ba_extended(d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone(), d["disps_sens"],
                             d["targets"], d["weights"], d["eta"], d["ii"], d["jj"], da, g.t0, g.t1, 1, 1e-4, 0.1,
                             False, 0.05, False, 1e-6, 1e-6, 1.0)
        p = d["poses"].clone()
        slam_ext.ba_extended(p, d["disps"].clone(), d["intrinsics"].clone(), d["disps_sens"], d["targets"],
                             d["weights"], d["eta"], d["ii"], d["jj"], da, g.t0, g.t1, 1, 1e-4, 0.1, True, 0.05,
                             False, 1e-6, 1e-6, 1.0)
        rp = g.poses.clone()
        OB.ba_extended(rp, g.disps.clone(), g.intrinsics.clone(), g.disps_sens, g.targets, g.weights, g.eta, g.ii,
                       g.jj, torch.ones(N), g.t0, g.t1, 1, 1e-4, 0.1, True, 0.05, False, 1e-6, 1e-6, 1.0)
        assert (p.cpu().double() - rp.double()).abs().max().item() < 1e-5


def test_ba_full_size_fp32_meets_accuracy_targets():
    """64 KF / 384 edges / 64x80, 2 GN iterations: FP32 device BA vs the FP64
    oracle within 1e-4 px reprojection and 1e-5 relative pose error."""
    g, mine, ref = _run_pair("kf64_e384_64x80", 2)
    p, d, it, dx, dz, en = mine
    rp, rd, rit, rdx, rdz, ren = ref
    px = _reproj_err(g, p, d, rp, rd, g.intrinsics)
    rel_pose = ((p.double() - rp.double()).norm(dim=1) / rp.double().norm(dim=1)).max().item()
    print(f"[full-size BA] reprojection err {px:.3e} px, relative pose err {rel_pose:.3e}")
    assert px < 1e-4
    assert rel_pose < 1e-5
    assert torch.allclose(en.double(), ren, rtol=1e-5)


@pytest.mark.parametrize("motion_only", [False, True])
def test_ba_implicit_adjoint_against_fp64_central_differences(motion_only):
    """dL/d(targets, weights) of one BA iteration through the implicit adjoint
    (forward Cholesky factor reused) against central differences (FP64 oracle,
    eps = 1e-6) on 64 random directions; L is the squared error of (dx, dz)
    to a random ground truth."""
    from vipe_amd.ba_layer import ba_iteration

    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    N = g.disps.shape[0]
    da = torch.ones(N)
    ds = g.disps.clone() * 1.01
    ds[:, ::2] = 0.0
    lm, ep, alpha = 1e-4, 0.1, 0.05
    gen = torch.Generator().manual_seed(11)
    d = _cuda(g)
    tg = d["targets"].clone().requires_grad_(True)
    wt = d["weights"].clone().requires_grad_(True)
    dx, dz = ba_iteration(tg, wt, d["poses"], d["disps"], d["intrinsics"], ds.cuda(), d["eta"], d["ii"], d["jj"],
                          da.cuda(), g.t0, g.t1, lm, ep, motion_only, alpha)
    Y = dx.detach().cpu().double() + 0.01 * torch.randn(dx.shape, generator=gen, dtype=torch.float64)
    loss = ((dx - Y.float().cuda()) ** 2).sum()
    if not motion_only:
        Z = dz.detach().cpu().double() + 0.01 * torch.randn(dz.shape, generator=gen, dtype=torch.float64)
        loss = loss + ((dz - Z.float().cuda()) ** 2).sum()
    loss.backward()
    gt_, gw_ = tg.grad.double().cpu(), wt.grad.double().cpu()

    def loss64(t64, w64):
        rdx, rdz, _ = OB.ba_extended(g.poses.clone(), g.disps.clone(), g.intrinsics.clone(), ds, t64, w64, g.eta,
                                     g.ii, g.jj, da, g.t0, g.t1, 1, lm, ep, motion_only, alpha, False, 1e-6, 1e-6,
                                     1.0, False)
        val = ((rdx - Y) ** 2).sum()
        if not motion_only:
            val = val + ((rdz - Z) ** 2).sum()
        return val.item()

    t64, w64 = g.targets.double(), g.weights.double()
    eps = 1e-6
    errs = []
    for k in range(64):
        vt = torch.randn(t64.shape, generator=gen, dtype=torch.float64)
        vw = torch.randn(w64.shape, generator=gen, dtype=torch.float64)
        fd = (loss64(t64 + eps * vt, w64 + eps * vw) - loss64(t64 - eps * vt, w64 - eps * vw)) / (2 * eps)
        an = ((gt_ * vt).sum() + (gw_ * vw).sum()).item()
        errs.append(abs(fd - an) / max(abs(fd), 1e-6))
    print(f"[BA adjoint motion_only={motion_only}] max rel directional error over 64 directions: {max(errs):.3e}")
    assert max(errs) < 2e-3


def _graph_with_edges(N, H, W, spec, seed, sim3=False, t0=1):
    """N frames of H x W pixels with the outgoing edges spec[i] of frame i, the
    synthetic targets / weights / damping of vipe_amd.synth."""
    from vipe_amd.synth import FrameGraph, _quat_exp

    gen = torch.Generator().manual_seed(seed)
    ii, jj = [], []
    for i, js in spec.items():
        ii.extend([i] * len(js))
        jj.extend(js)
    ii, jj = torch.tensor(ii), torch.tensor(jj)
    E = ii.shape[0]
    tau = torch.randn(N, 3, generator=gen, dtype=torch.float64) * 0.1
    phi = torch.randn(N, 3, generator=gen, dtype=torch.float64) * 0.05
    poses = torch.cat([tau, _quat_exp(phi)], -1)
    if sim3:
        poses = torch.cat([poses, torch.exp(torch.randn(N, 1, generator=gen, dtype=torch.float64) * 0.02)], -1)
    disps = 0.2 + 0.8 * torch.rand(N, H, W, generator=gen, dtype=torch.float64)
    intr = torch.tensor([20.0, 20.0, W / 2.0, H / 2.0], dtype=torch.float64)
    coords, _ = reproject(poses, disps, intr, ii, jj)
    targets = (coords + torch.randn(E, H, W, 2, generator=gen, dtype=torch.float64) * 0.5).permute(0, 3, 1, 2)
    weights = torch.rand(E, 2, H, W, generator=gen, dtype=torch.float64)
    kx = torch.unique(torch.cat([torch.arange(t0, N), ii]))
    eta = 0.2 * (0.05 + 0.1 * torch.rand(kx.shape[0], H, W, generator=gen, dtype=torch.float64)) + 1e-7
    return FrameGraph(poses=poses.float().contiguous(), disps=disps.float().contiguous(),
                      disps_sens=torch.zeros(N, H, W), intrinsics=intr.float(), ii=ii, jj=jj,
                      targets=targets.float().contiguous(), weights=weights.float().contiguous(),
                      eta=eta.float().contiguous(), fmaps=torch.zeros(0), coords=coords.float().contiguous(), t0=t0,
                      t1=N)


def _high_degree_graph(seed=21):
    """24 frames of 16 x 24 pixels.  Frame 1 has 20 outgoing edges (more than the
    18 whose E rows stay in LDS: the wide-frame path), frame 8 has 13 (15 rows
    of E: 21 tile pairs of the Schur Gram), frame 5 has 9 (11: 15 pairs),
    frames 10-23 one to three each."""
    N = 24
    spec = {1: list(range(2, 22)), 5: [0, 1, 2, 3, 4, 6, 7, 9, 12], 8: [0, 1, 2, 3, 4, 5, 6, 7, 9, 10, 11, 12, 13]}
    for i in range(10, N):
        spec[i] = [(i + 1) % N, (i + 2) % N, (i + N - 3) % N][: 1 + i % 3]
    return _graph_with_edges(N, 16, 24, spec, seed)


def _wide_graph(sim3=False, seed=41):
    """48 frames of 12 x 20 pixels: frame 1 with 40 outgoing edges (42 E-row
    blocks, 252 / 294 rows, 136 / 190 tile pairs: three and four Gram chunks),
    frame 30 with 33 including a stereo edge and edges to the fixed frame 0,
    frame 0 (outside the window) with 34, frame 8 with 13 and the others one
    to three."""
    N = 48
    spec = {1: list(range(2, 42)), 30: [30, 0] + list(range(31, 48)) + list(range(2, 16)),
            0: list(range(1, 35)), 8: [0, 1, 2, 3, 4, 5, 6, 7, 9, 10, 11, 12, 13]}
    for i in range(9, N):
        if i not in spec:
            spec[i] = [(i + 1) % N, (i + 2) % N, (i + N - 3) % N][: 1 + i % 3]
    return _graph_with_edges(N, 12, 20, spec, seed, sim3=sim3)


@pytest.mark.parametrize("tile", [0, 64, 128, 256])
def test_ba_high_degree_frames_match_oracle(tile):
    """Frames with 9, 13 and 20 outgoing edges: the Schur Gram's larger tile-pair
    counts per wave and its padding rows, and a wide frame, at every pixel tile."""
    from vipe_amd.ext import slam_ext

    g = _high_degree_graph()
    N, H, W = g.disps.shape
    da = torch.ones(N)
    ds = g.disps.clone() * 1.01
    ds[:, ::2] = 0.0
    iters, lm, ep, alpha = 2, 1e-4, 0.1, 0.05
    d = _cuda(g)
    p, dd = d["poses"].clone(), d["disps"].clone()
    plan = slam_ext.get_plan(d["ii"], d["jj"], N, H, W, g.t0, g.t1, tile)
    if tile:
        assert plan.pixel_tile == tile
    ws = torch.empty((plan.ws_bytes,), dtype=torch.uint8, device="cuda")
    dx = torch.empty((g.t1 - g.t0, 6), device="cuda")
    dz = torch.empty((plan.K, H * W), device="cuda")
    slam_ext.run_planned(plan, ws, p, dd, d["intrinsics"], ds.cuda(), d["targets"], d["weights"], d["eta"],
                         da.cuda(), iters, lm, ep, alpha, dx, dz)
    rp, rd = g.poses.clone(), g.disps.clone()
    rdx, rdz, _ = OB.ba_extended(rp, rd, g.intrinsics.clone(), ds, g.targets, g.weights, g.eta, g.ii, g.jj, da,
                                 g.t0, g.t1, iters, lm, ep, False, alpha, False, 1e-6, 1e-6, 1.0, False)
    torch.cuda.synchronize()
    assert (dx.cpu().double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert (p.cpu().double() - rp.double()).abs().max().item() < 1e-5
    assert (dd.cpu().double() - rd.double()).abs().max().item() < 1e-4
    assert _reproj_err(g, p.cpu(), dd.cpu(), rp, rd, g.intrinsics) < 1e-3


@pytest.mark.parametrize("sim3", [False, True])
@pytest.mark.parametrize("variant", ["plain", "intrinsics", "motion", "huber"])
def test_wide_frames_match_oracle(sim3, variant):
    """Frames with 40, 34 and 33 outgoing edges (one of them fixed, one with a
    stereo edge) through the wide-frame kernel, in each frame-kernel variant,
    against the FP64 oracle after 2 iterations."""
    from oracle import robust as R
    from vipe_amd.ext import slam_ext

    g = _wide_graph(sim3)
    N, H, W = g.disps.shape
    da = torch.ones(N)
    ds = g.disps.clone() * 1.01
    ds[:, ::2] = 0.0
    iters, lm, ep, alpha = 2, 1e-4, 0.1, 0.05
    intr, motion = variant == "intrinsics", variant == "motion"
    robust = (R.HUBER, 0.5, 1.0, 1e6) if variant == "huber" else None
    d = _cuda(g)
    p, dd, it = d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone()
    args = (ds.cuda(), d["targets"], d["weights"], d["eta"], d["ii"], d["jj"], da.cuda(),
            g.t0, g.t1, iters, lm, ep, motion, alpha, intr, 1e-6, 1e-6, 0.125, True)
    if robust:
        dx, dz, en = slam_ext.ba_extended_robust(p, dd, it, *args, robust)
    else:
        dx, dz, en = slam_ext.ba_extended(p, dd, it, *args)
    rp, rd, rit = g.poses.clone(), g.disps.clone(), g.intrinsics.clone()
    rdx, rdz, ren = OB.ba_extended(rp, rd, rit, ds, g.targets, g.weights, g.eta, g.ii, g.jj, da, g.t0, g.t1, iters,
                                   lm, ep, motion, alpha, intr, 1e-6, 1e-6, 0.125, True, robust=robust)
    torch.cuda.synchronize()
    print(f"[wide frames sim3={sim3} {variant}] dx err {(dx.cpu().double() - rdx).abs().max().item():.3e}")
    assert (dx.cpu().double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert (p.cpu().double() - rp.double()).abs().max().item() < 1e-5
    if motion:
        assert torch.equal(dd, d["disps"])
    else:
        assert (dz.cpu().double() - rdz).abs().max().item() < 1e-4
        assert (dd.cpu().double() - rd.double()).abs().max().item() < 1e-4
    if intr:
        assert (it.cpu().double() - rit.double()).abs().max().item() < 1e-4
    assert torch.allclose(en.cpu().double(), ren, rtol=1e-4, atol=1e-6)


def test_adjoint_refuses_wide_frames():
    from vipe_amd.ba_layer import ba_iteration

    g = _wide_graph()
    N = g.disps.shape[0]
    d = _cuda(g)
    tg = d["targets"].clone().requires_grad_(True)
    dx, dz = ba_iteration(tg, d["weights"], d["poses"], d["disps"], d["intrinsics"], d["disps_sens"], d["eta"],
                          d["ii"], d["jj"], torch.ones(N, device="cuda"), g.t0, g.t1, 1e-4, 0.1, False, 0.05)
    with pytest.raises(RuntimeError, match="outgoing edges"):
        (dx.sum() + dz.sum()).backward()


# ---- Sim3 poses ([N,8], 7-dof pose blocks) --------------------------------

@pytest.mark.parametrize("iters", [1, 2])
@pytest.mark.parametrize("intr", [False, True])
def test_sim3_ba_matches_oracle_kf8(iters, intr):
    g, mine, ref = _run_pair("kf8_e24_48x64_sim3", iters, intr=intr)
    p, d, it, dx, dz, en = mine
    rp, rd, rit, rdx, rdz, ren = ref
    assert dx.shape == (g.t1 - g.t0, 7) and rdx.shape == dx.shape
    assert (dx.double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert (dz.double() - rdz).abs().max().item() < 1e-4
    assert (p.double() - rp.double()).abs().max().item() < 1e-5
    assert (d.double() - rd.double()).abs().max().item() < 1e-4
    assert (it.double() - rit.double()).abs().max().item() < 1e-4
    assert _reproj_err(g, p, d, rp, rd, it) < 1e-3
    assert torch.allclose(en.double(), ren, rtol=1e-4, atol=1e-6)
    assert (p[g.t0:, 7] - g.poses[g.t0:, 7]).abs().max().item() > 0.0  # the scales move


def test_sim3_ba_motion_only_leaves_disps_bit_exact():
    g, mine, ref = _run_pair("kf8_e24_48x64_sim3", 2, motion_only=True)
    p, d, it, dx, dz, en = mine
    assert dz is None and torch.equal(d, g.disps)
    assert (p.double() - ref[0].double()).abs().max().item() < 1e-5


@pytest.mark.parametrize("name", ["kf64_e384_64x80", "kf64_e384_64x80_sim3"])
def test_ba_full_size_8_iterations_meets_accuracy_targets(name):
    """64 KF / 384 edges / 64x80, the 8 GN iterations of the measured step, SE3
    and Sim3 poses: FP32 device BA vs the FP64 oracle within 1e-4 px
    reprojection and 1e-5 relative pose error."""
    g, mine, ref = _run_pair(name, 8)
    p, d, it, dx, dz, en = mine
    rp, rd, rit, rdx, rdz, ren = ref
    px = _reproj_err(g, p, d, rp, rd, g.intrinsics)
    rel_pose = ((p.double() - rp.double()).norm(dim=1) / rp.double().norm(dim=1)).max().item()
    print(f"[full-size BA {name}, 8 iterations] reprojection err {px:.3e} px, relative pose err {rel_pose:.3e}")
    assert px < 1e-4
    assert rel_pose < 1e-5
    assert torch.allclose(en.double(), ren, rtol=1e-5)


def test_sim3_ba_high_degree_frames():
    """Sim3 frames with 9, 13 and 20 outgoing edges (the 20-edge frame on the
    wide-frame path) match the oracle."""
    from vipe_amd.ext import slam_ext

    g = _high_degree_graph()
    N, H, W = g.disps.shape
    s = torch.exp(torch.randn(N, 1, generator=torch.Generator().manual_seed(3)) * 0.02)
    g.poses = torch.cat([g.poses, s], -1).contiguous()
    d = _cuda(g)
    da = torch.ones(N, device="cuda")
    iters = 2
    p, dd = d["poses"].clone(), d["disps"].clone()
    dx, dz, _ = slam_ext.ba_extended(p, dd, d["intrinsics"].clone(), d["disps_sens"], d["targets"], d["weights"],
                                     d["eta"], d["ii"], d["jj"], da, g.t0, g.t1, iters, 1e-4, 0.1, False, 0.05, False,
                                     1e-6, 1e-6, 1.0)
    rp, rd = g.poses.clone(), g.disps.clone()
    rdx, _, _ = OB.ba_extended(rp, rd, g.intrinsics.clone(), g.disps_sens, g.targets, g.weights, g.eta, g.ii, g.jj,
                               torch.ones(N), g.t0, g.t1, iters, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, False)
    torch.cuda.synchronize()
    assert (dx.cpu().double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert (p.cpu().double() - rp.double()).abs().max().item() < 1e-5
    assert (dd.cpu().double() - rd.double()).abs().max().item() < 1e-4


LONG_WINDOW = dict(n_frames=640, n_edges=2554 + 256, H=8, W=12, intr=(10.0, 10.0, 6.0, 4.0), window=2,
                   n_random=256, sim3=False)


@pytest.mark.parametrize("sim3", [False, True])
def test_long_window_matches_oracle(sim3):
    """640 keyframes optimised over [1, 640) (639 poses: a 3834 / 4473-wide
    reduced system, 60 / 70 Cholesky blocks), 2810 edges, 2 GN iterations,
    against the FP64 oracle; a window of 1024 poses is refused."""
    from vipe_amd.ext import slam_ext
    from vipe_amd.synth import make_graph_from

    g = make_graph_from(dict(LONG_WINDOW, sim3=sim3), seed=5, with_fmaps=False)
    N = g.disps.shape[0]
    d = _cuda(g)
    da = torch.ones(N)
    iters = 2
    p, dd = d["poses"].clone(), d["disps"].clone()
    dx, dz, en = slam_ext.ba_extended(p, dd, d["intrinsics"].clone(), d["disps_sens"], d["targets"], d["weights"],
                                      d["eta"], d["ii"], d["jj"], da.cuda(), g.t0, g.t1, iters, 1e-4, 0.1, False, 0.05,
                                      False, 1e-6, 1e-6, 1.0, True)
    rp, rd = g.poses.clone(), g.disps.clone()
    rdx, rdz, ren = OB.ba_extended(rp, rd, g.intrinsics.clone(), g.disps_sens, g.targets, g.weights, g.eta, g.ii,
                                   g.jj, da, g.t0, g.t1, iters, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, True)
    torch.cuda.synchronize()
    assert dx.shape == (N - 1, 7 if sim3 else 6)
    px = _reproj_err(g, p.cpu(), dd.cpu(), rp, rd, g.intrinsics)
    rel_pose = ((p.cpu().double() - rp.double()).norm(dim=1) / rp.double().norm(dim=1)).max().item()
    print(f"[long window, sim3={sim3}] dx err {(dx.cpu().double() - rdx).abs().max().item():.3e}, "
          f"px err {px:.3e}, relative pose err {rel_pose:.3e}")
    assert (dx.cpu().double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert px < 1e-4 and rel_pose < 1e-5
    assert torch.allclose(en.cpu().double(), ren, rtol=1e-5)
    with pytest.raises(RuntimeError, match="1024 poses in the window"):
        slam_ext.get_plan(torch.tensor([0, 1]), torch.tensor([1, 0]), 1100, 4, 4, 1, 1025)
