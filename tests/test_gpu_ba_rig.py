"""gfx950 fused BA on multi-view rigs and the Mei camera (slam_ext.ba_rig /
vh_ba_run_rig) against the FP64 oracle (oracle/ba_rig.py).  Tolerances as the
single-view BA tests: dx 1e-5 relative, dz and disparities 1e-4, poses 1e-5,
energy 1e-4 relative."""
import pytest
import torch

from oracle import ba_rig as OR
from oracle import lie as L

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _se3(g, n, rot, trans):
    xi = torch.cat([torch.randn(n, 3, generator=g, dtype=F64) * trans, torch.randn(n, 3, generator=g, dtype=F64) * rot],
                   -1)
    return L.expm(3, xi)


def rig_graph(cam, N=12, V=2, H=24, W=32, window=2, t0=1, hub_edges=0, seed=0):
    """A rig buffer: N frames seen by V views (view 1 a 0.2 baseline to the
    right, rotated 0.1 rad about y), same-view edges between frames |i-j| <=
    window, an edge from each view of every window frame to the next view of
    the same frame, and `hub_edges` extra edges leaving view 0 of frame 3 (a
    wide depth frame above 18).  Targets are the FP64 projections + N(0, 0.5
    px), weights U[0, 1]."""
    g = torch.Generator().manual_seed(seed)
    poses = _se3(g, N, 0.05, 0.1).float()
    rig = torch.stack([L.expm(3, torch.tensor([[0.01, 0.0, -0.02, 0.0, 0.01, 0.0]], dtype=F64))[0],
                       L.expm(3, torch.tensor([[0.2, 0.0, 0.0, 0.0, 0.1, 0.0]], dtype=F64))[0]][:V]).float()
    if cam == OR.MEI:
        intr = torch.tensor([[26.0, 26.0, W / 2.0, H / 2.0, 0.7], [25.0, 25.5, W / 2.0 + 0.5, H / 2.0 - 0.5, 0.5]])
    else:
        intr = torch.tensor([[30.0, 30.0, W / 2.0, H / 2.0], [29.0, 29.5, W / 2.0 + 0.5, H / 2.0 - 0.5]])
    intr = intr[:V]
    pi, pj, qi, qj = [], [], [], []
    for i in range(N):
        for d in range(-window, window + 1):
            j = i + d
            if d == 0 or j < 0 or j >= N:
                continue
            for q in range(V):
                pi.append(i); pj.append(j); qi.append(q); qj.append(q)
        if t0 <= i and V > 1:
            for q in range(V):
                pi.append(i); pj.append(i); qi.append(q); qj.append((q + 1) % V)
    for k in range(hub_edges):
        pi.append(3); qi.append(0); pj.append(k % N if k % N != 3 else (k + 1) % N); qj.append(k % V)
    pi, pj, qi, qj = (torch.tensor(x) for x in (pi, pj, qi, qj))
    disps = (0.3 + 0.7 * torch.rand(N * V, H, W, generator=g)).float()
    E = pi.shape[0]
    T = OR.edge_terms(poses, disps, intr, cam, rig, torch.zeros(E, 2, H, W), torch.ones(E, 2, H, W), pi, pj, qi,
                      qj, V)
    proj = torch.stack([T["u"], T["v"]], 1).reshape(E, 2, H, W)
    targets = (proj + 0.5 * torch.randn(E, 2, H, W, generator=g, dtype=F64)).float()
    weights = torch.rand(E, 2, H, W, generator=g)
    disps_sens = disps * 1.02
    disps_sens[::3] = 0.0
    kx, _ = OR.bookkeeping(pi, qi, V, t0, N)
    eta = (torch.rand(kx.shape[0], H, W, generator=g) * 1e-3 + 1e-7).float()
    return dict(poses=poses, disps=disps, intr=intr, rig=rig, pi=pi, pj=pj, qi=qi, qj=qj, targets=targets,
                weights=weights, disps_sens=disps_sens, eta=eta, t0=t0, t1=N, N=N, V=V, kx=kx)


def _run(d, cam, iters, motion_only=False, robust=None, lm=1e-4, ep=0.1, active=None):
    from vipe_amd.ext import slam_ext

    name = "mei" if cam == OR.MEI else "pinhole"
    act = torch.ones(d["N"] * d["V"]) if active is None else active
    c = {k: (v.cuda() if torch.is_tensor(v) else v) for k, v in d.items()}
    p, dd = c["poses"].clone(), c["disps"].clone()
    print(f"ba_rig {name} N={d['N']} V={d['V']} E={d['pi'].shape[0]} iters={iters} motion={motion_only} "
          f"robust={robust}", flush=True)
    dx, dz, en = slam_ext.ba_rig(p, dd, c["intr"], name, c["rig"], c["disps_sens"], c["targets"], c["weights"],
                                 c["eta"], c["pi"], c["pj"], c["qi"], c["qj"], act.cuda(), d["t0"], d["t1"], iters,
                                 lm, ep, motion_only, 0.05, True, robust)
    torch.cuda.synchronize()
    rp, rd = d["poses"].clone(), d["disps"].clone()
    rob = None if robust is None else tuple(robust)
    rdx, rdz, ren = OR.ba_rig(rp, rd, d["intr"], cam, d["rig"], d["disps_sens"], d["targets"], d["weights"],
                              d["eta"], d["pi"], d["pj"], d["qi"], d["qj"], act, d["t0"], d["t1"], iters, lm, ep,
                              motion_only, 0.05, True, robust=rob)
    return (p.cpu(), dd.cpu(), dx.cpu(), None if dz is None else dz.cpu(), en.cpu()), (rp, rd, rdx, rdz, ren)


def _compare(mine, ref, motion_only=False):
    p, d, dx, dz, en = mine
    rp, rd, rdx, rdz, ren = ref
    edx = (dx.double() - rdx).abs().max().item()
    ep_ = (p.double() - rp.double()).abs().max().item()
    print(f"  dx err {edx:.2e} (max {rdx.abs().max().item():.2e}), pose err {ep_:.2e}", flush=True)
    assert edx < 1e-5 * max(1.0, rdx.abs().max().item())
    assert ep_ < 1e-5
    assert torch.allclose(en.double(), ren, rtol=1e-4, atol=1e-6)
    if motion_only:
        assert dz is None
        return
    edz = (dz.double() - rdz).abs().max().item()
    edd = (d.double() - rd.double()).abs().max().item()
    print(f"  dz err {edz:.2e} (max {rdz.abs().max().item():.2e}), disp err {edd:.2e}", flush=True)
    assert edz < 1e-4 and edd < 1e-4
    assert rdz.abs().max().item() > 1e-3 and rdx.abs().max().item() > 1e-4  # non-trivial steps


@pytest.mark.parametrize("cam", [OR.PINHOLE, OR.MEI], ids=["pinhole", "mei"])
@pytest.mark.parametrize("iters", [1, 3])
def test_rig_ba_matches_oracle(cam, iters):
    d = rig_graph(cam)
    mine, ref = _run(d, cam, iters)
    _compare(mine, ref)


@pytest.mark.parametrize("cam", [OR.PINHOLE, OR.MEI], ids=["pinhole", "mei"])
def test_rig_ba_wide_frame_matches_oracle(cam):
    """View 0 of frame 3 has 4 + 24 outgoing edges: the wide-frame kernel."""
    d = rig_graph(cam, N=10, H=12, W=16, hub_edges=24, seed=3)
    mine, ref = _run(d, cam, 2)
    _compare(mine, ref)


def test_rig_ba_motion_only_and_robust():
    d = rig_graph(OR.MEI, N=8, H=12, W=16, seed=4)
    mine, ref = _run(d, OR.MEI, 2, motion_only=True)
    assert torch.equal(mine[1], d["disps"])
    _compare(mine, ref, motion_only=True)
    from vipe_amd.ba_kernel import HuberRobustKernel

    # targets with outliers: the Huber weights change the step
    d2 = rig_graph(OR.PINHOLE, N=8, H=12, W=16, seed=5)
    d2["targets"][::3, :, ::4] += 6.0
    mine, ref = _run(d2, OR.PINHOLE, 2, robust=HuberRobustKernel(1.5).fused_args())
    _compare(mine, ref)


def test_rig_ba_depth_active_mask():
    d = rig_graph(OR.MEI, N=8, H=12, W=16, seed=6)
    act = torch.ones(d["N"] * d["V"])
    act[: 2 * d["V"]] = 0.0
    mine, ref = _run(d, OR.MEI, 1, active=act)
    _compare(mine, ref)


def test_one_view_identity_rig_matches_single_view_entry_point():
    """One view, an identity rig and the pinhole model: the same step as
    slam_ext.ba_extended on the same graph (no stereo edges)."""
    from vipe_amd.ext import slam_ext
    from vipe_amd.synth import make_graph

    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    sel = g.ii != g.jj
    ii, jj = g.ii[sel].cuda(), g.jj[sel].cuda()
    tg, wt = g.targets[sel].contiguous().cuda(), g.weights[sel].contiguous().cuda()
    N = g.disps.shape[0]
    ds = (g.disps * 1.01).cuda()
    act = torch.ones(N).cuda()
    kx = torch.unique(torch.cat([torch.arange(g.t0, g.t1), ii.cpu()]))
    eta = (torch.rand(kx.shape[0], *g.disps.shape[1:], generator=torch.Generator().manual_seed(1)) * 1e-3).cuda()
    intr = g.intrinsics.cuda()
    p1, d1 = g.poses.cuda(), g.disps.cuda().clone()
    p2, d2 = g.poses.cuda(), g.disps.cuda().clone()
    r1 = slam_ext.ba_extended(p1, d1, intr.clone(), ds, tg, wt, eta, ii, jj, act, g.t0, g.t1, 2, 1e-4, 0.1, False,
                              0.05, False, 1e-6, 1e-6, 1.0, True)
    zero = torch.zeros_like(ii)
    rig = torch.tensor([[0.0, 0, 0, 0, 0, 0, 1]]).cuda()
    r2 = slam_ext.ba_rig(p2, d2, intr[None].contiguous(), "pinhole", rig, ds, tg, wt, eta, ii, jj, zero, zero, act,
                         g.t0, g.t1, 2, 1e-4, 0.1, False, 0.05, True)
    torch.cuda.synchronize()
    for a, b in zip(r1, r2):
        err = (a - b).abs().max().item()
        print(f"  single-view vs rig: {err:.3e}", flush=True)
        assert err <= 1e-6 * max(1.0, a.abs().max().item())
    assert (p1 - p2).abs().max().item() <= 1e-6 and (d1 - d2).abs().max().item() <= 1e-6


def test_rig_entry_point_errors():
    from vipe_amd import _lib
    from vipe_amd.ext import slam_ext

    d = rig_graph(OR.PINHOLE, N=6, H=6, W=8)
    # an edge from a view of a frame onto itself
    with pytest.raises(RuntimeError, match="onto itself"):
        slam_ext.get_rig_plan(torch.tensor([1]), torch.tensor([1]), torch.tensor([0]), torch.tensor([0]), 6, 2, 6, 8,
                              1, 6)
    # a view index outside [0, V)
    with pytest.raises(RuntimeError, match="views"):
        slam_ext.get_rig_plan(torch.tensor([1]), torch.tensor([2]), torch.tensor([0]), torch.tensor([2]), 6, 2, 6, 8,
                              1, 6)
    # a rig plan refused by the single-view entry point
    plan = slam_ext.get_rig_plan(d["pi"], d["pj"], d["qi"], d["qj"], d["N"], d["V"], 6, 8, d["t0"], d["t1"])
    lib = _lib.load()
    x = torch.zeros(16, device="cuda")
    rc = lib.vh_ba_run(plan.h, _lib.ptr(x), _lib.ptr(x), _lib.ptr(x), _lib.ptr(x), _lib.ptr(x), _lib.ptr(x),
                       _lib.ptr(x), _lib.ptr(x), 1, 1e-4, 0.1, 0.05, 1e-6, 1e-6, 1.0, 0, _lib.ptr(x), _lib.ptr(x), None,
                       _lib.ptr(x), None)
    assert rc != 0 and "vh_ba_run_rig" in _lib.last_error()
    # Sim3 and intrinsics optimisation are refused on rig plans
    rc = lib.vh_ba_run_rig(plan.h, _lib.ptr(x), _lib.ptr(x), _lib.ptr(x), 0, _lib.ptr(x), _lib.ptr(x), _lib.ptr(x),
                           _lib.ptr(x), _lib.ptr(x), _lib.ptr(x), 1, 1e-4, 0.1, 0.05, 8, 0, 0.0, 0.0, 0.0,
                           _lib.ptr(x), _lib.ptr(x), None, _lib.ptr(x), None)
    assert rc != 0 and "SE3" in _lib.last_error()
    with pytest.raises(ValueError, match="camera"):
        slam_ext.ba_rig(*[None] * 3, "fisheye", *[None] * 19)


@pytest.mark.parametrize("camera", ["pinhole", "mei"])
def test_synthetic_rig_graph_kf8_matches_oracle(camera):
    """The rig version of the 8-keyframe graph (two views, 62 edges at 48 x 64)."""
    from vipe_amd.synth import make_rig_graph

    g = make_rig_graph("kf8_e24_48x64", camera=camera)
    N, V = g.poses.shape[0], 2
    ds = g.disps * 1.01
    ds[::2] = 0.0
    d = dict(poses=g.poses, disps=g.disps, intr=g.intrinsics, rig=g.rig, pi=g.pi, pj=g.pj, qi=g.qi, qj=g.qj,
             targets=g.targets, weights=g.weights, disps_sens=ds, eta=g.eta, t0=g.t0, t1=g.t1, N=N, V=V)
    mine, ref = _run(d, OR.MEI if camera == "mei" else OR.PINHOLE, 2)
    _compare(mine, ref)
