"""gfx950 robust kernels: the weight kernel and the fused BA with robust
weights vs the FP64 oracle (oracle/robust.py, oracle/ba.py).  Targets carry
planted outliers so that every branch of each kernel is exercised.  FP32
tolerances as for the plain BA (tests/test_gpu_ba.py)."""
import pytest
import torch

from oracle import ba as OB
from oracle import robust as R
from vipe_amd.synth import make_graph

pytestmark = pytest.mark.gpu

KINDS = [(R.HUBER, 0.5, 1.0, 1e6), (R.TUKEY, 2.0, 1.0, 1e6), (R.GNC_TLS, 1.5, 1.4, 1e6),
         (R.GNC_TLS, 1.5, 1e6, 1e6)]


@pytest.mark.parametrize("kind", KINDS)
def test_weight_kernel_matches_oracle(kind):
    from vipe_amd import ba_kernel

    r = torch.linspace(-8, 8, 100003)
    k = {R.HUBER: ba_kernel.HuberRobustKernel(kind[1]), R.TUKEY: ba_kernel.TukeyRobustKernel(kind[1])}.get(kind[0])
    if k is None:
        k = ba_kernel.TLSGncRobustKernel(threshold=kind[1], mu_max=kind[3])
        k.set_mu(kind[2])
    w = k.apply(r.cuda()).cpu().double()
    ref = R.weight(r.double(), *kind)
    assert (w - ref).abs().max().item() < 2e-6
    assert k.fused_args() == (kind[0], kind[1], kind[2] if kind[0] == R.GNC_TLS else 1.0,
                              kind[3] if kind[0] == R.GNC_TLS else 1.0)


def _outlier_graph():
    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    gen = torch.Generator().manual_seed(3)
    mask = torch.rand(g.targets.shape, generator=gen) < 0.05
    g.targets = torch.where(mask, g.targets + 30.0 * torch.randn(g.targets.shape, generator=gen), g.targets)
    return g


@pytest.mark.parametrize("kind", KINDS)
def test_fused_ba_with_robust_kernel_matches_oracle(kind):
    from vipe_amd.ext import slam_ext

    g = _outlier_graph()
    N = g.disps.shape[0]
    da = torch.ones(N)
    ds = g.disps.clone() * 1.01
    ds[:, ::2] = 0.0
    p, d, it = g.poses.clone().cuda(), g.disps.clone().cuda(), g.intrinsics.clone().cuda()
    dx, dz, en = slam_ext.ba_extended_robust(p, d, it, ds.cuda(), g.targets.cuda(), g.weights.cuda(), g.eta.cuda(),
                                             g.ii.cuda(), g.jj.cuda(), da.cuda(), g.t0, g.t1, 2, 1e-4, 0.1, False,
                                             0.05, False, 1e-6, 1e-6, 0.125, True, kind)
    rp, rd, rit = g.poses.clone(), g.disps.clone(), g.intrinsics.clone()
    rdx, rdz, ren = OB.ba_extended(rp, rd, rit, ds, g.targets, g.weights, g.eta, g.ii, g.jj, da, g.t0, g.t1, 2,
                                   1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6, 0.125, True, robust=kind)
    torch.cuda.synchronize()
    assert (dx.cpu().double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert (dz.cpu().double() - rdz).abs().max().item() < 1e-4
    assert (p.cpu().double() - rp.double()).abs().max().item() < 1e-5
    assert (d.cpu().double() - rd.double()).abs().max().item() < 1e-4
    assert torch.allclose(en.cpu().double(), ren, rtol=1e-4, atol=1e-6)
    # the robust weights matter: plain least squares takes another step
    ldx, _, _ = OB.ba_extended(g.poses.clone(), g.disps.clone(), g.intrinsics.clone(), ds, g.targets, g.weights,
                               g.eta, g.ii, g.jj, da, g.t0, g.t1, 1, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6,
                               0.125, False)
    rdx1, _, _ = OB.ba_extended(g.poses.clone(), g.disps.clone(), g.intrinsics.clone(), ds, g.targets, g.weights,
                                g.eta, g.ii, g.jj, da, g.t0, g.t1, 1, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6,
                                0.125, False, robust=kind)
    assert (ldx - rdx1).abs().max().item() > 1e-5


def test_huber_above_every_residual_is_plain_least_squares():
    from vipe_amd.ext import slam_ext

    g = make_graph("kf8_e24_48x64", with_fmaps=False)
    N = g.disps.shape[0]
    outs = []
    for robust in (None, (R.HUBER, 1e9, 1.0, 1.0)):
        p, d = g.poses.clone().cuda(), g.disps.clone().cuda()
        dx, dz, _ = slam_ext.ba_extended_robust(p, d, g.intrinsics.clone().cuda(), g.disps_sens.cuda(),
                                                g.targets.cuda(), g.weights.cuda(), g.eta.cuda(), g.ii.cuda(),
                                                g.jj.cuda(), torch.ones(N).cuda(), g.t0, g.t1, 2, 1e-4, 0.1, False,
                                                0.05, False, 1e-6, 1e-6, 0.125, False, robust)
        outs.append((p.cpu(), d.cpu(), dx.cpu(), dz.cpu()))
    for a, b in zip(*outs):
        assert (a - b).abs().max().item() < 1e-6


def test_graph_caller_runs_the_gnc_schedule():
    """GraphBuffer-style caller with ba.fused_robust and GNC-TLS vs the oracle
    running the same mu schedule."""
    import test_graph_ba as TG
    from vipe_amd import graph_ba

    cfg = dict(fused_robust=True, robust_kernel="gnc_tls", robust_kernel_threshold=0.02, gnc_mu_init=1.0,
               gnc_mu_step=2.0, gnc_mu_max=1e6, gnc_n_mu_steps=3, gnc_gn_iters_per_mu=2)
    buf = TG.make_buffer("cuda")
    buf.ba_config.update(cfg)
    inp = TG.make_inputs(buf)
    ref = TG.make_buffer("cpu")
    TG._call(buf, inp)
    target, weight, damping, ii, jj = [t.cpu() for t in inp]
    m = ii != jj
    ii, jj, target, weight = ii[m], jj[m], target[m], weight[m]
    target = target.view(-1, 4, 5, 2).permute(0, 3, 1, 2).contiguous()
    weight = weight.view(-1, 4, 5, 2).permute(0, 3, 1, 2).contiguous()
    kx = torch.unique(torch.cat([torch.arange(1, 5), ii]))
    eta = 0.2 * damping[kx] + 1e-7
    rp, rd, ri = ref.poses.clone(), ref.disps[:, 0].clone(), ref.intrinsics[0, :4] / 8.0
    for mu in R.gnc_schedule(1.0, 2.0, 1e6, 3):
        OB.ba_extended(rp, rd, ri, ref.disps_sens[:, 0], target, weight, eta, ii, jj, torch.ones(6), 1, 5, 2, 1e-3,
                       0.1, False, 0.001, False, 1e-6, 1e-6, 0.125, False, robust=(R.GNC_TLS, 0.02, mu, 1e6))
    rd = rd.clamp(min=0.001)
    torch.testing.assert_close(buf.poses[:5].cpu().double(), rp[:5].double(), atol=2e-5, rtol=1e-5)
    torch.testing.assert_close(buf.disps[:5, 0].cpu().double(), rd[:5].double(), atol=2e-4, rtol=2e-4)


@pytest.mark.parametrize("intr,motion", [(True, False), (False, True), (True, True)])
def test_robust_with_intrinsics_and_motion_only(intr, motion):
    """The other robust frame-kernel instantiations: intrinsics optimisation and
    motion-only, with Tukey weights and planted outliers."""
    from vipe_amd.ext import slam_ext

    g = _outlier_graph()
    N = g.disps.shape[0]
    da = torch.ones(N)
    kind = (R.TUKEY, 2.0, 1.0, 1e6)
    p, d, it = g.poses.clone().cuda(), g.disps.clone().cuda(), g.intrinsics.clone().cuda()
    dx, dz, en = slam_ext.ba_extended_robust(p, d, it, g.disps_sens.cuda(), g.targets.cuda(), g.weights.cuda(),
                                             g.eta.cuda(), g.ii.cuda(), g.jj.cuda(), da.cuda(), g.t0, g.t1, 2, 1e-4,
                                             0.1, motion, 0.05, intr, 1e-6, 1e-6, 0.125, True, kind)
    rp, rd, rit = g.poses.clone(), g.disps.clone(), g.intrinsics.clone()
    rdx, rdz, ren = OB.ba_extended(rp, rd, rit, g.disps_sens, g.targets, g.weights, g.eta, g.ii, g.jj, da, g.t0,
                                   g.t1, 2, 1e-4, 0.1, motion, 0.05, intr, 1e-6, 1e-6, 0.125, True, robust=kind)
    torch.cuda.synchronize()
    assert (dx.cpu().double() - rdx).abs().max().item() < 1e-5 * max(1.0, rdx.abs().max().item())
    assert (p.cpu().double() - rp.double()).abs().max().item() < 1e-5
    if motion:
        assert dz is None and torch.equal(d.cpu(), g.disps)
    else:
        assert (d.cpu().double() - rd.double()).abs().max().item() < 1e-4
    if intr:
        assert (it.cpu().double() - rit.double()).abs().max().item() < 1e-4
    assert torch.allclose(en.cpu().double(), ren, rtol=1e-4, atol=1e-6)
