"""Pins oracle/generic_ba.py (restatement of the reference's generic BA path):
finite differences of its residuals, the marginalised step against a
brute-force dense Gauss-Newton solve over every free unknown, its edge
rules (Z < 0.1 clamp and mask, steps above 10 dropped, the damping
placement with limited_disp), and the reference test's fused-vs-generic
agreement (tests/test_graph_buffer_fused_ba.py:203-249) against the fused
oracle."""
import pytest
import torch

from oracle import ba as OB
from oracle import generic_ba as G
from oracle import lie as L
from vipe_amd.synth import make_graph

F64 = torch.float64


def _small(H=6, W=8, seed=0):
    g = make_graph("kf8_e24_48x64", seed=seed, with_fmaps=False)
    E = g.ii.shape[0]
    d = dict(poses=g.poses.double().clone(), disps=g.disps[:, :H, :W].double().contiguous(),
             intr=torch.tensor([60.0, 60.0, W / 2.0, H / 2.0], dtype=F64),
             target=g.targets[:, :, :H, :W].permute(0, 2, 3, 1).reshape(E, H * W, 2).double().contiguous(),
             weight=g.weights[:, :, :H, :W].permute(0, 2, 3, 1).reshape(E, H * W, 2).double().contiguous(),
             ii=g.ii, jj=g.jj)
    d["disps_sens"] = d["disps"] * 1.02
    d["disps_sens"][::3] = 0.0
    d["damping"] = 0.1 + 0.05 * torch.rand(8, H, W, generator=torch.Generator().manual_seed(seed), dtype=F64)
    return d


def test_jacobians_match_finite_differences():
    d = _small()
    ii, jj, E = d["ii"], d["jj"], d["ii"].shape[0]
    args = (d["intr"], d["target"], 0.001 * d["weight"], ii, jj, ii)
    T0 = G.edge_terms(d["poses"], d["disps"], *args, intrinsics_factor=2.0, with_focal=True)
    ok = (T0["w"][..., 0] > 0)
    eps = 1e-6
    checked = 0
    for f in torch.unique(ii).tolist()[:4]:
        for k in range(6):
            xi = torch.zeros(6, dtype=F64)
            xi[k] = eps

            def moved(sgn):
                p = d["poses"].clone()
                p[f] = L.SE3.mul(L.SE3.exp((sgn * xi)[None]), p[f][None])[0]
                return p

            fd = (G.edge_terms(moved(1.0), d["disps"], *args, intrinsics_factor=2.0)["r"]
                  - G.edge_terms(moved(-1.0), d["disps"], *args, intrinsics_factor=2.0)["r"]) / (2 * eps)
            for side, mask in (("Ji", (ii == f) & (jj != f)), ("Jj", (jj == f) & (ii != f))):
                if mask.any():
                    err = ((fd - T0[side][..., k]).abs() * ok[..., None])[mask].max()
                    assert err < 1e-5, (f, side, k, err)
                    checked += 1
    assert checked >= 40
    dp = G.edge_terms(d["poses"], d["disps"] + eps, *args, intrinsics_factor=2.0)["r"]
    dm = G.edge_terms(d["poses"], d["disps"] - eps, *args, intrinsics_factor=2.0)["r"]
    assert (((dp - dm) / (2 * eps) - T0["Jz"]).abs() * ok[..., None]).max() < 1e-5
    ip, im = d["intr"].clone(), d["intr"].clone()
    ip[:2] += eps
    im[:2] -= eps
    fp = G.edge_terms(d["poses"], d["disps"], ip, *args[1:], intrinsics_factor=2.0)["r"]
    fm = G.edge_terms(d["poses"], d["disps"], im, *args[1:], intrinsics_factor=2.0)["r"]
    assert (((fp - fm) / (2 * eps) - T0["Jf"]).abs() * ok[..., None]).max() < 1e-5


def _dense_step(d, t0, t1, lm, ep, alpha=0.001):
    """Brute force: stack the explicit Jacobian of every residual over the free
    poses (source poses outside [t0, t1) fixed) and every source frame's
    disparities, add the damping and the sensor prior, solve densely."""
    ii, jj = d["ii"], d["jj"]
    E = ii.shape[0]
    N, H, W = d["disps"].shape
    HW = H * W
    T = G.edge_terms(d["poses"], d["disps"], d["intr"], d["target"], 0.001 * d["weight"], ii, jj, ii)
    pi_u = torch.unique(ii)
    fixed = set(pi_u[(pi_u < t0) | (pi_u >= t1)].tolist())
    poses = sorted((set(ii.tolist()) | set(jj.tolist())) - fixed)
    ps = {p: k for k, p in enumerate(poses)}
    frames = torch.unique(ii).tolist()
    fs = {f: k for k, f in enumerate(frames)}
    nvar = 6 * len(poses) + len(frames) * HW
    rows, res, wts = [], [], []
    for e in range(E):
        for c in range(2):
            J = torch.zeros(HW, nvar, dtype=F64)
            for side, p in (("Ji", int(ii[e])), ("Jj", int(jj[e]))):
                if p in ps:
                    J[:, 6 * ps[p]:6 * ps[p] + 6] += T[side][e, :, c]
            col = 6 * len(poses) + fs[int(ii[e])] * HW
            J[torch.arange(HW), col + torch.arange(HW)] = T["Jz"][e, :, c]
            rows.append(J)
            res.append(T["r"][e, :, c])
            wts.append(T["w"][e, :, c])
    J, r, w = torch.cat(rows), torch.cat(res), torch.cat(wts)
    Hm = J.T @ (w[:, None] * J)
    g = -J.T @ (w * r)
    n6 = 6 * len(poses)
    idx = torch.arange(n6)
    Hm[idx, idx] += ep + lm * Hm[idx, idx]
    dz = d["disps"].reshape(N, HW)
    ds = d["disps_sens"].reshape(N, HW)
    for f, k in fs.items():
        sl = slice(n6 + k * HW, n6 + (k + 1) * HW)
        dd = torch.arange(sl.start, sl.stop)
        Hm[dd, dd] += 0.2 * d["damping"].reshape(N, HW)[f] + 2e-7
        if ds[f].sum() > 0:
            Hm[dd, dd] += alpha
            g[sl] -= alpha * (dz[f] - ds[f])
    x = torch.linalg.solve(Hm, g)
    return poses, frames, x[:n6].reshape(-1, 6), x[n6:].reshape(-1, HW)


def test_marginalised_step_equals_dense_gauss_newton():
    d = _small()
    N = d["poses"].shape[0]
    p, dd, it = d["poses"].clone(), d["disps"].clone(), d["intr"].clone()
    G.gn_step(p, dd, it, d["disps_sens"], d["target"], d["weight"], d["damping"], d["ii"], d["jj"], 1, N, 1e-3, 0.1,
              False, False, False, dtype=F64)
    poses, frames, xp, xz = _dense_step(d, 1, N, 1e-3, 0.1)
    want_p = d["poses"].clone()
    for k, f in enumerate(poses):
        want_p[f] = L.SE3.mul(L.SE3.exp(xp[k][None]), d["poses"][f][None])[0]
    assert (p - want_p).abs().max() < 1e-10
    want_d = d["disps"].clone().reshape(N, -1)
    want_d[frames] += xz
    assert (dd.reshape(N, -1) - want_d).abs().max() < 1e-10
    assert (xz.abs().max() > 1e-6) and (xp.abs().max() > 1e-6)


def test_depth_below_0_1_projects_with_unit_depth_and_no_weight():
    d = _small()
    ii, jj = d["ii"], d["jj"]
    t, R = G._relative(d["poses"], ii, jj)
    e = int(torch.argmax(t[:, 2].abs() * (ii != jj)))
    X0 = torch.tensor([(0 - 4.0) / 60.0, (0 - 3.0) / 60.0, 1.0], dtype=F64)
    # choose pixel (0, 0)'s disparity so that the point lands at Z = 0.05
    disp = (0.05 - (R[e] @ X0)[2]) / t[e, 2]
    d["disps"][int(ii[e]), 0, 0] = disp
    T = G.edge_terms(d["poses"], d["disps"], d["intr"], d["target"], 0.001 * d["weight"], ii, jj, ii)
    assert abs(T["z"][e, 0].item() - 0.05) < 1e-9
    X1 = R[e] @ X0 + t[e] * disp
    want = torch.stack([60.0 * X1[0] + 4.0, 60.0 * X1[1] + 3.0])  # Z := 1
    assert torch.allclose(T["r"][e, 0] + d["target"][e, 0], want)
    assert torch.equal(T["w"][e, 0], torch.zeros(2, dtype=F64))
    assert (T["w"][e, 1:] > 0).all()


@pytest.mark.parametrize("offset,moves", [(5.0, True), (20.0, False)])
def test_disparity_steps_above_10_are_dropped(offset, moves):
    """A strong sensor prior pulls a frame by ~offset: a step above 10 is dropped
    (DenseDispRetractor, retractor.py:41-44)."""
    d = _small()
    N = d["poses"].shape[0]
    d["disps_sens"].zero_()
    f = int(d["ii"][0])
    d["disps_sens"][f] = d["disps"][f] + offset
    dd = d["disps"].clone()
    G.gn_step(d["poses"].clone(), dd, d["intr"].clone(), d["disps_sens"], d["target"], d["weight"], d["damping"],
              d["ii"], d["jj"], 1, N, 1e-3, 0.1, False, False, False, alpha=1e6, dtype=F64)
    step = (dd[f] - d["disps"][f]).abs().max().item()
    assert (abs(step - offset) < 1e-3) if moves else step == 0.0


def test_limited_disp_damping_lands_on_the_next_free_frame():
    """With limited_disp the fixed source frame 0 (its edge leaves the window)
    keeps its damping entry, which searchsorted places on frame 1
    (matrix.py:187-191): the step equals the dense solve with frame 1 damped
    by both frames' vectors and frame 0 fixed."""
    d = _small()
    N = d["poses"].shape[0]
    assert 0 in d["ii"].tolist()
    dd = d["disps"].clone()
    G.gn_step(d["poses"].clone(), dd, d["intr"].clone(), d["disps_sens"], d["target"], d["weight"], d["damping"],
              d["ii"], d["jj"], 1, N, 1e-3, 0.1, False, True, False, dtype=F64)
    assert torch.equal(dd[0], d["disps"][0])
    # frame 0's disparities fixed: drop its source edges from the disparity
    # unknowns by giving the brute force frame 0 an enormous damping instead
    d2 = {k: (v.clone() if torch.is_tensor(v) else v) for k, v in d.items()}
    d2["damping"][0] = 1e30
    d2["damping"][1] = d["damping"][0] + d["damping"][1] + 1e-7 / 0.2
    _, frames, _, xz = _dense_step(d2, 1, N, 1e-3, 0.1)
    k1 = frames.index(1)
    # (the 1e30 stand-in for a fixed frame leaves ~1e-8 of coupling)
    assert (dd[1].reshape(-1) - (d["disps"][1].reshape(-1) + xz[k1])).abs().max() < 1e-7
    d2["damping"][1] = d["damping"][1]
    _, _, _, xz_own = _dense_step(d2, 1, N, 1e-3, 0.1)
    assert (dd[1].reshape(-1) - (d["disps"][1].reshape(-1) + xz_own[k1])).abs().max() > 1e-4


def _fixture(tracks):
    """The reference test's graph (tests/test_graph_buffer_fused_ba.py:70-139)
    with the sparse-track term merged into one target and weight, which gives
    the same normal equations as the generic path's second term."""
    from tests.test_graph_ba import make_buffer, make_inputs

    buf = make_buffer("cpu", tracks=tracks)
    target, weight, damping, ii, jj = make_inputs(buf)
    if tracks:
        st, sw = buf.sparse_tracks.compute_dense_disp_target_weight(ii * 0, buf.tstamp[ii], jj * 0, buf.tstamp[jj],
                                                                    (32, 40), (4, 5))
        st, sw = st.flatten(1, 2), sw.flatten(1, 2)
        cw = weight + sw
        target = torch.where(cw > 0, (weight * target + sw * st) / cw.clamp_min(1e-12), target)
        weight = cw
    return buf, target, weight, damping, ii, jj


@pytest.mark.parametrize("tracks,limited,intr,motion", [(True, True, True, False), (False, False, False, True)])
def test_generic_matches_fused_within_reference_tolerance(tracks, limited, intr, motion):
    """tests/test_graph_buffer_fused_ba.py:203-249: one iteration of each path on
    the same near-converged graph agrees within 2e-4 (poses, disps, focal)."""
    buf, target, weight, damping, ii, jj = _fixture(tracks)
    poses = buf.poses.double().clone()
    disps = buf.disps[:, 0].double().clone()
    focal = buf.intrinsics[0, :4].double().clone()
    G.bundle_adjustment(poses, disps, focal, buf.disps_sens[:, 0].double(), target.double(), weight.double(),
                        damping.double(), ii, jj, 1, 5, 1, 1e-3, 0.1, motion_only=motion, limited_disp=limited,
                        optimize_intrinsics=intr, intrinsics_factor=8.0, dtype=F64)
    # fused oracle on the reference's fused preparation
    m = ii != jj
    fi, fj = ii[m], jj[m]
    ft = target[m].view(-1, 4, 5, 2).permute(0, 3, 1, 2).contiguous()
    fw = weight[m].view(-1, 4, 5, 2).permute(0, 3, 1, 2).contiguous()
    kx = torch.unique(torch.cat([torch.arange(1, 5), fi]))
    eta = 0.2 * damping[kx] + 1e-7
    da = torch.ones(6, dtype=F64)
    if limited:
        da.zero_()
        da[1:5] = 1.0
    rp, rd = buf.poses.double().clone(), buf.disps[:, 0].double().clone()
    ri = buf.intrinsics[0, :4].double() / 8.0
    OB.ba_extended(rp, rd, ri, buf.disps_sens[:, 0], ft, fw, eta, fi, fj, da, 1, 5, 1, 1e-3, 0.1, motion, 0.001, intr,
                   1e-6, 1e-6, 0.125, False)
    rd = rd.clamp(min=0.001)
    torch.testing.assert_close(poses[:5], rp[:5], atol=2e-4, rtol=2e-4)
    torch.testing.assert_close(disps[:5], rd[:5], atol=2e-4, rtol=2e-4)
    torch.testing.assert_close(focal, ri * 8.0, atol=2e-4, rtol=2e-4)


def test_float32_restatement_tracks_float64():
    d = _small()
    N = d["poses"].shape[0]
    outs = []
    for dt in (torch.float32, F64):
        p, dd, it = d["poses"].clone().to(dt), d["disps"].clone().to(dt), d["intr"].clone().to(dt)
        G.bundle_adjustment(p, dd, it, d["disps_sens"].to(dt), d["target"].to(dt), d["weight"].to(dt),
                            d["damping"].to(dt), d["ii"], d["jj"], 1, N, 2, 1e-3, 0.1, dtype=dt)
        outs.append((p.double(), dd.double()))
    assert (outs[0][0] - outs[1][0]).abs().max() < 1e-4
    assert (outs[0][1] - outs[1][1]).abs().max() < 1e-3
