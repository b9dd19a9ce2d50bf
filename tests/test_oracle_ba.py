"""Pins oracle/ba.py: finite-difference Jacobians, the Schur path against a
brute-force dense Gauss-Newton solve, the back-substitution quirk against the
dense system, the energy against an independent reprojection, and the
reference tests' invariants."""
import pytest
import torch

from oracle import ba as OB
from oracle import lie as L
from vipe_amd.synth import make_graph, reproject

F64 = torch.float64


def _small(name="kf8_e24_48x64", H=6, W=8):
    g = make_graph(name, with_fmaps=False)
    d = {k: (v.clone() if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}
    for k in ("disps", "disps_sens"):
        d[k] = d[k][:, :H, :W].contiguous()
    for k in ("targets", "weights"):
        d[k] = d[k][:, :, :H, :W].contiguous()
    d["eta"] = d["eta"][:, :H, :W].contiguous()
    d["intrinsics"] = torch.tensor([60.0, 60.0, W / 2.0, H / 2.0])
    d["disps_sens"] = d["disps"] * 1.02
    d["disps_sens"][:, ::2] = 0.0
    return d


SE3_SIM3 = ["kf8_e24_48x64", "kf8_e24_48x64_sim3"]


@pytest.mark.parametrize("name", SE3_SIM3)
def test_jacobians_match_finite_differences(name):
    d = _small(name)
    poses = d["poses"].double()
    D = OB.pose_dof(poses)
    gid = 4 if D == 7 else 3
    T0 = OB.edge_terms(poses, d["disps"].double(), d["intrinsics"].double(), d["targets"], d["weights"], d["ii"],
                       d["jj"], intrinsics_scale=0.125)
    eps = 1e-6
    E = d["ii"].shape[0]
    sel = ~(d["ii"] == d["jj"])
    for k in range(D):
        xi = torch.zeros(D, dtype=F64)
        xi[k] = eps
        def proj_with(which, sgn):
            p = poses.clone()
            idx = d["jj"] if which == "j" else d["ii"]
            for e in range(E):
                if which == "i" and d["ii"][e] == d["jj"][e]:
                    continue
            # perturb every frame on one side by exp(+-xi) (edges touch each frame once per side)
            for f in torch.unique(idx).tolist():
                p[f] = L.mul(gid, L.expm(gid, (sgn * xi)[None]), p[f][None])[0]
            return p
        for which, Jname in (("j", "J"), ("i", "Ji")):
            pp = proj_with(which, 1.0)
            pm = proj_with(which, -1.0)
            # only edges whose other endpoint is unperturbed isolate the derivative
            Tp = OB.edge_terms(pp, d["disps"].double(), d["intrinsics"].double(), d["targets"], d["weights"],
                               d["ii"], d["jj"])
            Tm = OB.edge_terms(pm, d["disps"].double(), d["intrinsics"].double(), d["targets"], d["weights"],
                               d["ii"], d["jj"])
            fdu = (Tp.proj_u - Tm.proj_u) / (2 * eps)
            fdv = (Tp.proj_v - Tm.proj_v) / (2 * eps)
            other = d["ii"] if which == "j" else d["jj"]
            moved = torch.unique(d["jj"] if which == "j" else d["ii"])
            ok = sel & ~torch.isin(other, moved)
            if ok.sum() == 0:
                continue
            valid = (T0.z >= OB.MIN_DEPTH)[ok]
            if which == "j":
                au, av = T0.Ju[ok][..., k], T0.Jv[ok][..., k]
            else:
                At = T0.Adj[ok].transpose(1, 2)
                au = -(At[:, None] @ T0.Ju[ok][..., None])[..., 0][..., k]
                av = -(At[:, None] @ T0.Jv[ok][..., None])[..., 0][..., k]
            assert ((au - fdu[ok]).abs() * valid).max() < 1e-5
            assert ((av - fdv[ok]).abs() * valid).max() < 1e-5
    # depth and focal derivatives
    dd = d["disps"].double()
    Tp = OB.edge_terms(poses, dd + eps, d["intrinsics"].double(), d["targets"], d["weights"], d["ii"], d["jj"])
    Tm = OB.edge_terms(poses, dd - eps, d["intrinsics"].double(), d["targets"], d["weights"], d["ii"], d["jj"])
    v = T0.z >= OB.MIN_DEPTH
    assert (((Tp.proj_u - Tm.proj_u) / (2 * eps) - T0.Jzu).abs() * v).max() < 1e-5
    assert (((Tp.proj_v - Tm.proj_v) / (2 * eps) - T0.Jzv).abs() * v).max() < 1e-5
    s = 0.125
    ip = d["intrinsics"].double().clone(); ip[:2] += eps * s
    im = d["intrinsics"].double().clone(); im[:2] -= eps * s
    Tp = OB.edge_terms(poses, dd, ip, d["targets"], d["weights"], d["ii"], d["jj"])
    Tm = OB.edge_terms(poses, dd, im, d["targets"], d["weights"], d["ii"], d["jj"])
    assert (((Tp.proj_u - Tm.proj_u) / (2 * eps) - T0.Jfu).abs() * v).max() < 1e-5
    assert (((Tp.proj_v - Tm.proj_v) / (2 * eps) - T0.Jfv).abs() * v).max() < 1e-5


def _args(d):
    return (d["disps_sens"], d["targets"], d["weights"], d["eta"], d["ii"], d["jj"])


@pytest.mark.parametrize("name", SE3_SIM3)
def test_schur_path_equals_dense_gauss_newton(name):
    d = _small(name)
    N = d["poses"].shape[0]
    da = torch.ones(N)
    p, dd, it = d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone()
    dx, dz, _ = OB.ba_extended(p, dd, it, *_args(d)[:4], d["ii"], d["jj"], da, 1, N, 1, 0.0, 0.1, False, 0.05,
                               False, 1e-6, 1e-6, 1.0, False, skip_quirk=False)
    rdx, rdz = OB.dense_gn_step(d["poses"], d["disps"], d["intrinsics"], *_args(d)[:4], d["ii"], d["jj"], da, 1, N,
                                0.0, 0.1, 0.05)
    assert (dx - rdx).abs().max() < 1e-10
    assert (dz - rdz).abs().max() < 1e-10


@pytest.mark.parametrize("name", SE3_SIM3)
def test_back_substitution_skips_pose_t0_like_the_reference(name):
    """dz with the reference's EvT quirk equals the dense back-substitution with the
    columns of local pose 0 dropped (geom_kernels.cu:1186)."""
    d = _small(name)
    D = OB.pose_dof(d["poses"])
    N = d["poses"].shape[0]
    da = torch.ones(N)
    p, dd, it = d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone()
    dx, dz, _ = OB.ba_extended(p, dd, it, *_args(d)[:4], d["ii"], d["jj"], da, 1, N, 1, 0.0, 0.1, False, 0.05,
                               False, 1e-6, 1e-6, 1.0, False, skip_quirk=True)
    rdx, rdz, Hm, gvec = OB.dense_gn_step(d["poses"], d["disps"], d["intrinsics"], *_args(d)[:4], d["ii"], d["jj"],
                                          da, 1, N, 0.0, 0.1, 0.05, return_system=True)
    P = N - 1
    Hzz = torch.diagonal(Hm[D * P:, D * P:])
    Hzp = Hm[D * P:, D:D * P]  # drop local pose 0
    dz_q = (gvec[D * P:] - Hzp @ rdx[1:].reshape(-1)) / Hzz
    assert (dz.reshape(-1) - dz_q).abs().max() < 1e-10
    assert (dz.reshape(-1) - rdz.reshape(-1)).abs().max() > 1e-6  # the quirk is visible


def test_energy_equals_independent_reprojection():
    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    N = g.poses.shape[0]
    _, _, en = OB.ba_extended(g.poses.clone(), g.disps.clone(), g.intrinsics.clone(), g.disps_sens, g.targets,
                              g.weights, g.eta, g.ii, g.jj, torch.ones(N), g.t0, g.t1, 1, 1e-4, 0.1, False, 0.05, False,
                              1e-6, 1e-6, 1.0, True)
    coords, z = reproject(g.poses, g.disps, g.intrinsics, g.ii, g.jj)
    r = g.targets.double().permute(0, 2, 3, 1) - coords
    w = 0.001 * g.weights.double().permute(0, 2, 3, 1) * (z >= 0.25)[..., None]
    ref = (w * r * r).sum()
    assert abs(en[0].item() - ref.item()) < 1e-6 * ref.item()


def test_invariants_motion_only_inactive_frames_and_legacy_entry():
    d = _small()
    N = d["poses"].shape[0]
    dd = d["disps"].clone()
    OB.ba_extended(d["poses"].clone(), dd, d["intrinsics"].clone(), *_args(d)[:4], d["ii"], d["jj"], torch.ones(N),
                   1, N, 1, 1e-4, 0.1, True, 0.05, False, 1e-6, 1e-6, 1.0, False)
    assert torch.equal(dd, d["disps"])
    da = torch.ones(N)
    da[2] = 0.0
    dd = d["disps"].clone()
    OB.ba_extended(d["poses"].clone(), dd, d["intrinsics"].clone(), *_args(d)[:4], d["ii"], d["jj"], da, 1, N, 1,
                   1e-4, 0.1, False, 0.05, True, 1e-6, 1e-6, 0.125, False)
    assert torch.equal(dd[2], d["disps"][2])
    p1, p2 = d["poses"].clone(), d["poses"].clone()
    OB.ba(p1, d["disps"].clone(), d["intrinsics"].clone(), *_args(d)[:4], d["ii"], d["jj"], 1, N, 1, 1e-4, 0.1, False)
    OB.ba_extended(p2, d["disps"].clone(), d["intrinsics"].clone(), *_args(d)[:4], d["ii"], d["jj"], torch.ones(N), 1,
                   N, 1, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, False)
    assert torch.equal(p1, p2)


def test_failed_factorisation_returns_zero_update():
    d = _small()
    N = d["poses"].shape[0]
    p = d["poses"].clone()
    dx, _, _ = OB.ba_extended(p, d["disps"].clone(), d["intrinsics"].clone(), d["disps_sens"], d["targets"],
                              torch.zeros_like(d["weights"]), d["eta"], d["ii"], d["jj"], torch.ones(N), 1, N, 1, 0.0,
                              -1.0, True, 0.05, False, 1e-6, 1e-6, 1.0, False)
    assert torch.count_nonzero(dx) == 0
    assert torch.equal(p, d["poses"])


def test_sim3_step_retracts_by_the_group_product_and_observes_scale():
    """Sim3 poses: the update is T <- Exp(dx) T as 4x4 matrices (sim3.h Exp and
    product), the log-scale unknowns receive a non-zero update, and a frame's
    dx has 7 entries."""
    d = _small("kf8_e24_48x64_sim3")
    N = d["poses"].shape[0]
    p = d["poses"].clone()
    dx, dz, _ = OB.ba_extended(p, d["disps"].clone(), d["intrinsics"].clone(), *_args(d)[:4], d["ii"], d["jj"],
                               torch.ones(N), 1, N, 1, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, False)
    assert dx.shape == (N - 1, 7) and dx[:, 6].abs().max() > 1e-6
    want = L.Sim3.matrix4(L.Sim3.exp(dx)) @ L.Sim3.matrix4(d["poses"][1:].double())
    got = L.Sim3.matrix4(p[1:].double())
    assert (got - want).abs().max() < 1e-6  # poses are stored in float32
    assert torch.equal(p[0], d["poses"][0])


def test_sim3_poses_with_unit_scale_keep_the_se3_residuals():
    """A Sim3 pose with s = 1 is the SE3 pose: identical residuals, j-side
    Jacobians equal in the first six columns and zero in the scale column."""
    d = _small()
    p7 = d["poses"].double()
    p8 = torch.cat([p7, torch.ones(p7.shape[0], 1, dtype=F64)], -1)
    a = OB.edge_terms(p7, d["disps"], d["intrinsics"], d["targets"], d["weights"], d["ii"], d["jj"])
    b = OB.edge_terms(p8, d["disps"], d["intrinsics"], d["targets"], d["weights"], d["ii"], d["jj"])
    assert torch.equal(a.ru, b.ru) and torch.equal(a.rv, b.rv)
    assert torch.allclose(b.Ju[..., :6], a.Ju) and torch.count_nonzero(b.Ju[..., 6]) == 0
    assert torch.allclose(b.Adj[:, :6, :6], a.Adj)
