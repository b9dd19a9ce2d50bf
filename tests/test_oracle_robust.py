"""Pins oracle/robust.py and the robust weighting of oracle/ba.py (host only).

Each IRLS weight equals rho'(r) / (rho''(0) r) of its loss, checked by central
differences of the loss written from its own definition:
  Huber      rho = r^2/2 (|r| <= k), k|r| - k^2/2 otherwise
  Tukey      rho = c^2/6 (1 - (1 - (r/c)^2)^3) (|r| <= c), c^2/6 otherwise
  GNC-TLS    the graduated surrogate of truncated least squares (Yang et al.
             2020): r^2 below mu/(mu+1) c^2; 2 c |r| sqrt(mu (mu+1)) - mu (c^2 + r^2)
             in the band; c^2 above (mu+1)/mu c^2
and the Schur path with robust weights equals a brute-force dense solve."""
import math

import pytest
import torch

from oracle import ba as OB
from oracle import robust as R
from vipe_amd.synth import make_graph, reproject

F64 = torch.float64


def _rho_huber(r, k):
    a = abs(r)
    return 0.5 * r * r if a <= k else k * a - 0.5 * k * k


def _rho_tukey(r, c):
    return c * c / 6.0 * (1.0 - (1.0 - (r / c) ** 2) ** 3) if abs(r) <= c else c * c / 6.0


def _rho_gnc(r, c, mu):
    r2, c2 = r * r, c * c
    lo, hi = mu / (mu + 1.0) * c2, (mu + 1.0) / mu * c2
    if r2 <= lo:
        return r2
    if r2 >= hi:
        return c2
    return 2.0 * c * abs(r) * math.sqrt(mu * (mu + 1.0)) - mu * (c2 + r2)


def _fd_weight(rho, r, scale, h=1e-6):
    return (rho(r + h) - rho(r - h)) / (2 * h) / (scale * r)


@pytest.mark.parametrize("kind,c,mu,rho,scale", [
    (R.HUBER, 1.3, 1.0, lambda r: _rho_huber(r, 1.3), 1.0),
    (R.TUKEY, 2.0, 1.0, lambda r: _rho_tukey(r, 2.0), 1.0),
    (R.GNC_TLS, 3.0, 1.0, lambda r: _rho_gnc(r, 3.0, 1.0), 2.0),
    (R.GNC_TLS, 1.5, 7.3, lambda r: _rho_gnc(r, 1.5, 7.3), 2.0),
])
def test_weights_are_the_losses_irls_weights(kind, c, mu, rho, scale):
    rs = [x for x in torch.linspace(-2.2 * c, 2.2 * c, 157).tolist() if abs(x) > 1e-3]
    w = R.weight(torch.tensor(rs, dtype=F64), kind, c, mu, 1e6)
    for r, wi in zip(rs, w.tolist()):
        assert abs(wi - _fd_weight(rho, r, scale)) < 1e-5, (r, wi)


def test_gnc_limits_and_schedule():
    r = torch.linspace(-5, 5, 1001, dtype=F64)
    c = 2.0
    hard = R.weight(r, R.GNC_TLS, c, 1e6, 1e6)
    assert torch.equal(hard, (r * r < c * c).to(F64))
    # large mu approaches the indicator away from the threshold
    near = R.weight(r, R.GNC_TLS, c, 1e4, 1e6)
    away = (r.abs() - c).abs() > 0.05
    assert (near[away] - hard[away]).abs().max() < 1e-6
    # continuity: weight 1 at the inner edge, 0 at the outer edge of the band
    mu = 2.5
    lo, hi = math.sqrt(mu / (mu + 1) * c * c), math.sqrt((mu + 1) / mu * c * c)
    edge = R.weight(torch.tensor([lo * (1 + 1e-9), hi * (1 - 1e-9)], dtype=F64), R.GNC_TLS, c, mu, 1e6)
    assert abs(edge[0] - 1.0) < 1e-6 and abs(edge[1]) < 1e-6
    assert R.gnc_schedule(1.0, 1.4, 2.0, 4) == [1.0, 1.4, 1.4 * 1.4, 2.0]
    # Huber with a threshold above every residual is plain least squares
    assert torch.equal(R.weight(r, R.HUBER, 1e9), torch.ones_like(r))


def _small(H=6, W=8):
    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    d = {k: (v.clone() if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}
    for k in ("disps", "disps_sens"):
        d[k] = d[k][:, :H, :W].contiguous()
    for k in ("targets", "weights"):
        d[k] = d[k][:, :, :H, :W].contiguous()
    d["eta"] = d["eta"][:, :H, :W].contiguous()
    d["intrinsics"] = torch.tensor([60.0, 60.0, W / 2.0, H / 2.0])
    d["disps_sens"] = d["disps"] * 1.02
    d["disps_sens"][:, ::2] = 0.0
    # outliers: a few targets far off
    d["targets"][3, 0, 1:3, 2:5] += 25.0
    d["targets"][7, 1, 4, :] -= 40.0
    return d


@pytest.mark.parametrize("robust", [(R.HUBER, 1.0, 1.0, 1e6), (R.TUKEY, 4.0, 1.0, 1e6),
                                    (R.GNC_TLS, 3.0, 1.4, 1e6), (R.GNC_TLS, 3.0, 1e6, 1e6)])
def test_robust_schur_path_equals_dense_gauss_newton(robust):
    d = _small()
    N = d["poses"].shape[0]
    da = torch.ones(N)
    args = (d["disps_sens"], d["targets"], d["weights"], d["eta"])
    p, dd, it = d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone()
    dx, dz, _ = OB.ba_extended(p, dd, it, *args, d["ii"], d["jj"], da, 1, N, 1, 0.0, 0.1, False, 0.05,
                               False, 1e-6, 1e-6, 1.0, False, skip_quirk=False, robust=robust)
    rdx, rdz = OB.dense_gn_step(d["poses"], d["disps"], d["intrinsics"], *args, d["ii"], d["jj"], da, 1, N,
                                0.0, 0.1, 0.05, robust=robust)
    assert (dx - rdx).abs().max() < 1e-10
    assert (dz - rdz).abs().max() < 1e-10
    # and the robust weights change the step
    ldx, _, _ = OB.ba_extended(d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone(), *args, d["ii"],
                               d["jj"], da, 1, N, 1, 0.0, 0.1, False, 0.05, False, 1e-6, 1e-6, 1.0, False,
                               skip_quirk=False)
    assert (dx - ldx).abs().max() > 1e-6


def test_robust_energy_equals_independent_reprojection():
    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    N = g.poses.shape[0]
    rob = (R.TUKEY, 1.5, 1.0, 1e6)
    _, _, en = OB.ba_extended(g.poses.clone(), g.disps.clone(), g.intrinsics.clone(), g.disps_sens, g.targets,
                              g.weights, g.eta, g.ii, g.jj, torch.ones(N), g.t0, g.t1, 1, 1e-4, 0.1, False, 0.05, False,
                              1e-6, 1e-6, 1.0, True, robust=rob)
    coords, z = reproject(g.poses, g.disps, g.intrinsics, g.ii, g.jj)
    r = g.targets.double().permute(0, 2, 3, 1) - coords
    w = 0.001 * g.weights.double().permute(0, 2, 3, 1) * (z >= 0.25)[..., None]
    w = w * R.weight(r, *rob)
    ref = (w * r * r).sum()
    assert abs(en[0].item() - ref.item()) < 1e-6 * ref.item()
