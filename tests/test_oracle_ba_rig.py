"""Pins oracle/ba_rig.py (host only): the one-view identity-rig pinhole case
against oracle/ba.py, finite-difference Jacobians for both camera models under
a non-identity rig, the Mei model's closed forms, and one step against a
brute-force dense Gauss-Newton solve."""
import pytest
import torch

from oracle import ba as OB
from oracle import ba_rig as OR
from oracle import lie as L

F64 = torch.float64


def _se3(g, n, rot=0.05, trans=0.1):
    xi = torch.cat([torch.randn(n, 3, generator=g, dtype=F64) * trans, torch.randn(n, 3, generator=g, dtype=F64) * rot],
                   -1)
    return L.expm(3, xi)


def rig_graph(cam=OR.MEI, seed=0, H=5, W=6, N=4):
    """Four frames seen by two views, same-view edges between frames, one edge
    within a frame (view 0 -> view 1) and an edge across views."""
    g = torch.Generator().manual_seed(seed)
    V = 2
    poses = _se3(g, N).float()
    rig = torch.stack([L.expm(3, torch.tensor([[0.02, -0.01, 0.03, 0.01, -0.02, 0.015]], dtype=F64))[0],
                       L.expm(3, torch.tensor([[-0.2, 0.01, 0.0, 0.0, 0.3, 0.0]], dtype=F64))[0]]).float()
    K = [4.0, 4.5, W / 2.0, H / 2.0] + ([0.6] if cam == OR.MEI else [])
    intr = torch.tensor([K, [4.2, 4.0, W / 2.0 - 0.3, H / 2.0 + 0.2] + ([0.4] if cam == OR.MEI else [])])
    pi = torch.tensor([0, 1, 2, 3, 1, 2, 0])
    pj = torch.tensor([1, 2, 1, 2, 1, 0, 3])
    qi = torch.tensor([0, 1, 0, 1, 0, 1, 1])
    qj = torch.tensor([0, 1, 0, 1, 1, 0, 1])
    disps = (0.5 + torch.rand(N * V, H, W, generator=g)).float()
    E = pi.shape[0]
    T = OR.edge_terms(poses, disps, intr, cam, rig, torch.zeros(E, 2, H, W), torch.ones(E, 2, H, W), pi, pj, qi, qj, V)
    proj = torch.stack([T["u"], T["v"]], 1).reshape(E, 2, H, W)
    targets = (proj + 0.3 * torch.randn(E, 2, H, W, generator=g, dtype=F64)).float()
    weights = torch.rand(E, 2, H, W, generator=g)
    disps_sens = disps * 1.03
    disps_sens[::3] = 0.0
    return dict(poses=poses, disps=disps, intr=intr, cam=cam, rig=rig, pi=pi, pj=pj, qi=qi, qj=qj, V=V,
                targets=targets, weights=weights, disps_sens=disps_sens, H=H, W=W, N=N)


def test_one_view_identity_rig_equals_fused_oracle():
    from vipe_amd.synth import make_graph

    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    H, W = 6, 8
    sel = g.ii != g.jj
    ii, jj = g.ii[sel], g.jj[sel]
    disps = g.disps[:, :H, :W].contiguous()
    targets = g.targets[sel][:, :, :H, :W].contiguous()
    weights = g.weights[sel][:, :, :H, :W].contiguous()
    intr = torch.tensor([60.0, 60.0, W / 2.0, H / 2.0])
    ds = disps * 1.02
    ds[::2] = 0.0
    kx, _ = OR.bookkeeping(ii, torch.zeros_like(ii), 1, g.t0, g.t1)
    eta = torch.rand(kx.shape[0], H, W, generator=torch.Generator().manual_seed(3)) * 0.01
    act = torch.ones(disps.shape[0])
    act[1] = 0.0
    a = dict(poses=g.poses.clone(), disps=disps.clone())
    b = dict(poses=g.poses.clone(), disps=disps.clone())
    ra = OB.ba_extended(a["poses"], a["disps"], intr.clone(), ds, targets, weights, eta, ii, jj, act, g.t0, g.t1, 2,
                        1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, True)
    ident = torch.tensor([[0.0, 0, 0, 0, 0, 0, 1]])
    zero = torch.zeros_like(ii)
    rb = OR.ba_rig(b["poses"], b["disps"], intr[None], OR.PINHOLE, ident, ds, targets, weights, eta, ii, jj, zero,
                   zero, act, g.t0, g.t1, 2, 1e-4, 0.1, False, 0.05, True)
    for x, y in zip(ra, rb):
        assert torch.allclose(x, y, rtol=1e-11, atol=1e-13)
    assert torch.equal(a["poses"], b["poses"]) and torch.equal(a["disps"], b["disps"])


@pytest.mark.parametrize("cam", [OR.PINHOLE, OR.MEI])
def test_jacobians_match_finite_differences(cam):
    d = rig_graph(cam)
    V, E = d["V"], d["pi"].shape[0]
    base = dict(disps=d["disps"].double(), intr=d["intr"].double(), cam=cam, rig=d["rig"].double(),
                targets=d["targets"], weights=d["weights"], pi=d["pi"], pj=d["pj"], qi=d["qi"], qj=d["qj"], V=V)

    def proj(poses, disps=None):
        T = OR.edge_terms(poses, base["disps"] if disps is None else disps, base["intr"], cam, base["rig"],
                          base["targets"], base["weights"], base["pi"], base["pj"], base["qi"], base["qj"], V)
        return torch.stack([T["u"], T["v"]], -1), T

    poses = d["poses"].double()
    _, T0 = proj(poses)
    eps = 1e-6
    checked = 0
    for f in range(d["N"]):
        for k in range(6):
            xi = torch.zeros(1, 6, dtype=F64)
            xi[0, k] = eps
            pp, pm = poses.clone(), poses.clone()
            pp[f] = L.mul(3, L.expm(3, xi), poses[f][None])[0]
            pm[f] = L.mul(3, L.expm(3, -xi), poses[f][None])[0]
            fd = (proj(pp)[0] - proj(pm)[0]) / (2 * eps)  # [E,HW,2]
            for e in range(E):
                on_i, on_j = int(d["pi"][e]) == f, int(d["pj"][e]) == f
                if not (on_i or on_j):
                    assert fd[e].abs().max() < 1e-9
                    continue
                ana = torch.zeros_like(fd[e])
                if on_i:
                    ana += torch.stack([T0["Jiu"][e][:, k], T0["Jiv"][e][:, k]], -1)
                if on_j:
                    ana += torch.stack([T0["Jju"][e][:, k], T0["Jjv"][e][:, k]], -1)
                ok = T0["z"][e] > OR.MIN_DEPTH + 1e-3
                err = (fd[e] - ana)[ok].abs().max()
                assert err < 1e-6 * (1 + ana[ok].abs().max()), (f, k, e, float(err))
                if on_i and on_j:  # an edge within one frame: the pose columns cancel
                    assert ana.abs().max() < 1e-12
                checked += 1
    assert checked > 20
    # disparity of each source depth frame
    for e in range(E):
        dframe = int(d["pi"][e]) * V + int(d["qi"][e])
        dp, dm = base["disps"].clone(), base["disps"].clone()
        dp[dframe] += eps
        dm[dframe] -= eps
        fd = (proj(poses, dp)[0][e] - proj(poses, dm)[0][e]) / (2 * eps)
        ana = torch.stack([T0["Jzu"][e], T0["Jzv"][e]], -1)
        ok = T0["z"][e] > OR.MIN_DEPTH + 1e-3
        assert (fd - ana)[ok].abs().max() < 1e-6 * (1 + ana.abs().max())


def test_mei_model_closed_forms():
    g = torch.Generator().manual_seed(5)
    u = torch.rand(200, generator=g, dtype=F64) * 64
    v = torch.rand(200, generator=g, dtype=F64) * 48
    Kp = torch.tensor([50.0, 52.0, 31.0, 23.5], dtype=F64)
    # k1 = 0 is the pinhole model
    K0 = torch.cat([Kp, torch.zeros(1, dtype=F64)])
    for a, b in zip(OR.iproj(OR.MEI, K0, u, v), OR.iproj(OR.PINHOLE, Kp, u, v)):
        assert torch.allclose(a, b, rtol=1e-14, atol=1e-14)
    X, Y, Z = torch.randn(3, 200, generator=g, dtype=F64)
    Z = Z.abs() + 0.5
    for a, b in zip(OR.project(OR.MEI, K0, X, Y, Z), OR.project(OR.PINHOLE, Kp, X, Y, Z)):
        assert torch.allclose(a.double(), b.double(), rtol=1e-13, atol=1e-13)
    for k1 in (0.3, 0.8):
        K = torch.cat([Kp, torch.tensor([k1], dtype=F64)])
        # proj(iproj(u, v)) = (u, v): the unit-depth ray maps back to its pixel, at any scale
        X0, Y0 = OR.iproj(OR.MEI, K, u, v)
        for s in (1.0, 3.7):
            pu, pv, _, _, ok = OR.project(OR.MEI, K, s * X0, s * Y0, s * torch.ones_like(X0))
            assert bool(ok.all())
            assert torch.allclose(pu, u, atol=1e-10) and torch.allclose(pv, v, atol=1e-10)
        # the projection Jacobian against central differences
        _, _, gu, gv, _ = OR.project(OR.MEI, K, X, Y, Z)
        eps = 1e-6
        for c in range(3):
            dp = [X, Y, Z]
            dm = [X, Y, Z]
            dp = [t + eps if i == c else t for i, t in enumerate(dp)]
            dm = [t - eps if i == c else t for i, t in enumerate(dm)]
            up, vp = OR.project(OR.MEI, K, *dp)[:2]
            um, vm = OR.project(OR.MEI, K, *dm)[:2]
            assert torch.allclose((up - um) / (2 * eps), gu[:, c], atol=1e-6)
            assert torch.allclose((vp - vm) / (2 * eps), gv[:, c], atol=1e-6)


@pytest.mark.parametrize("cam", [OR.PINHOLE, OR.MEI])
def test_schur_step_equals_dense_gauss_newton(cam):
    d = rig_graph(cam, seed=1)
    V, H, W = d["V"], d["H"], d["W"]
    t0, t1 = 1, 4
    kx, _ = OR.bookkeeping(d["pi"], d["qi"], V, t0, t1)
    eta = torch.rand(kx.shape[0], H, W, generator=torch.Generator().manual_seed(2)) * 0.01 + 1e-3
    act = torch.ones(d["N"] * V)
    poses, disps = d["poses"].clone(), d["disps"].clone()
    args = (d["intr"], cam, d["rig"], d["disps_sens"], d["targets"], d["weights"], eta, d["pi"], d["pj"], d["qi"],
            d["qj"])
    ddx, ddz = OR.dense_gn_step_rig(d["poses"], d["disps"], *args, t0, t1, 0.0, 1e-4, 0.05)
    dx, dz, _ = OR.ba_rig(poses, disps, *args, act, t0, t1, 1, 0.0, 1e-4, False, 0.05, skip_quirk=False)
    assert torch.allclose(dx, ddx, rtol=1e-8, atol=1e-10), float((dx - ddx).abs().max())
    assert torch.allclose(dz, ddz, rtol=1e-8, atol=1e-10), float((dz - ddz).abs().max())
    # the step is not trivial
    assert float(dx.abs().max()) > 1e-4 and float(dz.abs().max()) > 1e-4


def test_back_substitution_skip_and_motion_only():
    d = rig_graph(OR.MEI, seed=2)
    V, H, W = d["V"], d["H"], d["W"]
    t0, t1 = 0, 4
    kx, _ = OR.bookkeeping(d["pi"], d["qi"], V, t0, t1)
    eta = torch.full((kx.shape[0], H, W), 1e-3)
    act = torch.ones(d["N"] * V)
    args = (d["intr"], OR.MEI, d["rig"], d["disps_sens"], d["targets"], d["weights"], eta, d["pi"], d["pj"],
            d["qi"], d["qj"], act, t0, t1, 1, 1e-4, 0.1)
    p1, d1 = d["poses"].clone(), d["disps"].clone()
    p2, d2 = d["poses"].clone(), d["disps"].clone()
    dx1, dz1, _ = OR.ba_rig(p1, d1, *args, False, 0.05, skip_quirk=True)
    dx2, dz2, _ = OR.ba_rig(p2, d2, *args, False, 0.05, skip_quirk=False)
    assert torch.equal(dx1, dx2)  # the skip changes only the depth update
    assert not torch.allclose(dz1, dz2)
    # motion only: disparities untouched, no depth update
    p3, d3 = d["poses"].clone(), d["disps"].clone()
    dx3, dz3, en = OR.ba_rig(p3, d3, *args, True, 0.05, compute_energy=True)
    assert dz3 is None and torch.equal(d3, d["disps"]) and float(en[0]) > 0
    assert not torch.equal(p3, d["poses"])


@pytest.mark.parametrize("camera", ["pinhole", "mei"])
def test_synthetic_rig_graph_targets_are_the_projections(camera):
    """vipe_amd.synth.make_rig_graph with zero noise: its targets are the
    oracle's projections of the same points (its own float64 chain of rig,
    pose and camera maps)."""
    from vipe_amd.synth import make_rig_graph

    g = make_rig_graph("kf8_e24_48x64", camera=camera, views=2, noise=0.0)
    cam = OR.MEI if camera == "mei" else OR.PINHOLE
    M = g.pi.shape[0]
    H, W = g.disps.shape[1:]
    T = OR.edge_terms(g.poses, g.disps, g.intrinsics, cam, g.rig, torch.zeros(M, 2, H, W), torch.ones(M, 2, H, W),
                      g.pi, g.pj, g.qi, g.qj, 2)
    ok = T["z"] > 0.3
    assert float(ok.double().mean()) > 0.9
    proj = torch.stack([T["u"], T["v"]], 1)
    err = (g.targets.double().reshape(M, 2, H * W) - proj).abs().permute(0, 2, 1)[ok].max()
    assert float(err) < 1e-3, float(err)
    # every view of the window's frames, and the cross-view edges q -> q + 1
    assert int(((g.pi == g.pj) & (g.qj == (g.qi + 1) % 2)).sum()) == 2 * (g.t1 - g.t0)
