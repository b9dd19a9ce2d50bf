"""Caller-side fused BA (GraphBuffer._fused_ba restatement): the reference's
error messages (tests/test_graph_buffer_fused_ba.py:277-324) on CPU, and the
fixture of that test file solved on the GPU against the FP64 oracle."""
from types import SimpleNamespace

import pytest
import torch

from oracle import ba as OB
from vipe_amd import graph_ba
from vipe_amd.synth import reproject


class FixtureSparseTracks:
    """tests/test_graph_buffer_fused_ba.py:24-50: a constant-velocity track target."""

    def __init__(self, enabled):
        self.enabled = enabled

    def compute_dense_disp_target_weight(self, source_view_inds, source_frame_inds, target_view_inds,
                                         target_frame_inds, image_size, dense_disp_size):
        h, w = dense_disp_size
        yy, xx = torch.meshgrid(torch.arange(h, device=source_frame_inds.device, dtype=torch.float32),
                                torch.arange(w, device=source_frame_inds.device, dtype=torch.float32), indexing="ij")
        coords = torch.stack([xx, yy], dim=-1)
        dt = (target_frame_inds - source_frame_inds).to(torch.float32).view(-1, 1, 1, 1)
        target = coords.unsqueeze(0) + torch.cat([0.03 * dt, -0.02 * dt], dim=-1)
        weight = torch.full_like(target, 0.07)
        weight = weight * (source_frame_inds != target_frame_inds).to(torch.float32).view(-1, 1, 1, 1)
        return target, weight


def make_buffer(device, tracks=False, robust=None):
    """tests/test_graph_buffer_fused_ba.py:70-106 construction (CPU RNG)."""
    torch.manual_seed(20260519)
    height, width, n_frames, size = 32, 40, 5, 6
    h8, w8 = height // 8, width // 8
    poses = torch.zeros(size, 7)
    poses[:, 6] = 1.0
    poses[:n_frames, 0] = torch.linspace(0.0, 0.08, n_frames)
    poses[:n_frames, 1] = torch.tensor([0.0, 0.01, -0.005, 0.015, 0.02])
    poses[:n_frames, 2] = torch.linspace(0.0, 0.025, n_frames)
    yy, xx = torch.meshgrid(torch.arange(h8, dtype=torch.float32), torch.arange(w8, dtype=torch.float32), indexing="ij")
    base = 0.72 + 0.015 * xx + 0.02 * yy
    disps = torch.full((size, 1, h8, w8), 0.8)
    for f in range(n_frames):
        disps[f, 0] = base + 0.025 * f
    disps_sens = torch.zeros(size, 1, h8, w8)
    disps_sens[2, 0] = disps[2, 0] * 1.015
    disps_sens[4, 0] = disps[4, 0] * 0.985
    intr = torch.zeros(1, 4)
    intr[0, :4] = torch.tensor([38.0, 38.0, 19.5, 15.5])
    buf = SimpleNamespace(
        poses=poses.to(device), disps=disps.to(device), disps_sens=disps_sens.to(device), intrinsics=intr.to(device),
        tstamp=torch.arange(size).to(device), rig=torch.tensor([[0, 0, 0, 0, 0, 0, 1.0]]).to(device), n_views=1,
        camera_type="pinhole", height=height, width=width, device=device,
        ba_config={"dense_disp_alpha": 0.001, "intrinsics_damping_scale": 1.0, "robust_kernel": robust},
        sparse_tracks=FixtureSparseTracks(tracks))
    return buf


def make_inputs(buf):
    """tests/test_graph_buffer_fused_ba.py:109-139 (targets from the reprojection)."""
    ii = torch.tensor([1, 2, 3, 4, 0, 3])
    jj = torch.tensor([0, 0, 1, 2, 2, 3])
    h8, w8 = buf.height // 8, buf.width // 8
    coords, _ = reproject(buf.poses.cpu(), buf.disps.cpu()[:, 0], buf.intrinsics.cpu()[0] / 8.0, ii, jj)
    target = coords.float().reshape(6, h8 * w8, 2)
    yy, xx = torch.meshgrid(torch.arange(h8, dtype=torch.float32), torch.arange(w8, dtype=torch.float32), indexing="ij")
    delta = torch.stack([0.01 * torch.sin(0.7 * xx + 0.2 * yy), -0.012 * torch.cos(0.3 * xx - 0.5 * yy)], -1)
    edge = torch.linspace(-0.015, 0.02, 6).view(-1, 1, 1)
    target = target + delta.reshape(1, -1, 2) + torch.cat([edge, -0.5 * edge], dim=-1)
    weight = (0.2 + 0.03 * (xx + yy).reshape(1, -1, 1)).repeat(6, 1, 2).contiguous()
    weight[-1].zero_()
    damping = 0.08 + 0.01 * torch.arange(6, dtype=torch.float32).view(-1, 1, 1) + 0.002 * xx + 0.001 * yy
    dev = buf.poses.device
    return target.contiguous().to(dev), weight.to(dev), damping.contiguous().to(dev), ii.to(dev), jj.to(dev)


def _call(buf, inp, **kw):
    target, weight, damping, ii, jj = inp
    args = dict(t0=1, t1=5, n_iters=1, pose_damping=1e-3, pose_ep=0.1, motion_only=False, limited_disp=False,
                optimize_intrinsics=False, optimize_rig_rotation=False, verbose=False)
    args.update(kw)
    return graph_ba.bundle_adjustment(buf, target.clone(), weight.clone(), damping.clone(), ii.clone(), jj.clone(),
                                      **args)


def test_errors_instead_of_falling_back():
    buf = make_buffer("cpu")
    inp = make_inputs(buf)
    with pytest.raises(RuntimeError, match="does not support optimizing rig rotation"):
        _call(buf, inp, optimize_rig_rotation=True)
    with pytest.raises(RuntimeError, match="the active optimization range is empty"):
        _call(buf, inp, t0=3, t1=3)
    buf2 = make_buffer("cpu", robust="huber")
    with pytest.raises(RuntimeError, match="robust kernels are not supported"):
        _call(buf2, inp)
    buf3 = make_buffer("cpu")
    buf3.n_views = 2
    with pytest.raises(RuntimeError, match="single-view pinhole-camera"):
        _call(buf3, inp)


@pytest.mark.gpu
@pytest.mark.parametrize("tracks,limited,intr,motion", [(True, True, True, False), (False, False, False, True),
                                                        (False, False, False, False)])
def test_fixture_matches_oracle(tracks, limited, intr, motion):
    buf = make_buffer("cuda", tracks=tracks)
    inp = make_inputs(buf)
    d0 = buf.disps.clone()
    ref = make_buffer("cpu", tracks=tracks)
    _call(buf, inp, limited_disp=limited, optimize_intrinsics=intr, motion_only=motion)
    # independent preparation for the oracle, per buffer.py:408-467
    target, weight, damping, ii, jj = [t.cpu() for t in inp]
    m = ii != jj
    ii, jj, target, weight = ii[m], jj[m], target[m], weight[m]
    if tracks:
        st, sw = ref.sparse_tracks.compute_dense_disp_target_weight(ii * 0, ref.tstamp[ii], jj * 0, ref.tstamp[jj],
                                                                    (32, 40), (4, 5))
        st, sw = st.flatten(1, 2), sw.flatten(1, 2)
        cw = weight + sw
        target = torch.where(cw > 0, (weight * target + sw * st) / cw.clamp_min(1e-12), target)
        weight = cw
    target = target.view(-1, 4, 5, 2).permute(0, 3, 1, 2).contiguous()
    weight = weight.view(-1, 4, 5, 2).permute(0, 3, 1, 2).contiguous()
    kx = torch.unique(torch.cat([torch.arange(1, 5), ii]))
    eta = 0.2 * damping[kx] + 1e-7
    da = torch.ones(6)
    if limited:
        da.zero_()
        da[1:5] = 1.0
    intrinsics = ref.intrinsics[0, :4] / 8.0
    disps = ref.disps[:, 0].clone()
    OB.ba_extended(ref.poses, disps, intrinsics, ref.disps_sens[:, 0], target, weight, eta, ii, jj, da, 1, 5, 1,
                   1e-3, 0.1, motion, 0.001, intr, 1e-6, 1e-6, 0.125, False)
    disps = disps.clamp(min=0.001)
    torch.testing.assert_close(buf.poses[:5].cpu().double(), ref.poses[:5].double(), atol=2e-5, rtol=1e-5)
    torch.testing.assert_close(buf.disps[:5, 0].cpu().double(), disps[:5].double(), atol=2e-4, rtol=2e-4)
    if intr:
        torch.testing.assert_close(buf.intrinsics[0, :4].cpu().double(), (intrinsics * 8.0).double(),
                                   atol=2e-4, rtol=2e-4)
    if motion:
        assert torch.equal(buf.disps, d0)


def test_robust_kernels_run_on_the_fused_path_when_enabled(monkeypatch):
    """ba.fused_robust: Huber / Tukey run n_iters iterations with their weights;
    GNC-TLS runs gnc_n_mu_steps x gnc_gn_iters_per_mu over the mu schedule
    (buffer.py:686-695).  The fused calls are recorded, not run."""
    from oracle import robust as R

    calls = []

    def fake(*a, robust=None, **kw):
        calls.append((a[8], robust))
        return torch.zeros(0)

    monkeypatch.setattr(graph_ba, "fused_ba", fake)
    buf = make_buffer("cpu", robust="tukey")
    buf.ba_config.update(fused_robust=True, robust_kernel_threshold=2.5)
    inp = make_inputs(buf)
    _call(buf, inp, n_iters=3)
    assert calls == [(3, (2, 2.5, 1.0, 1.0))]
    calls.clear()
    buf.ba_config.update(robust_kernel="gnc_tls", gnc_mu_init=1.0, gnc_mu_step=2.0, gnc_mu_max=5.0,
                         gnc_n_mu_steps=4, gnc_gn_iters_per_mu=2)
    _call(buf, inp, n_iters=9)
    assert [c[0] for c in calls] == [2, 2, 2, 2]
    assert [c[1][2] for c in calls] == R.gnc_schedule(1.0, 2.0, 5.0, 4)
    assert all(c[1][:2] == (3, 2.5) and c[1][3] == 5.0 for c in calls)
    with pytest.raises(ValueError, match="Unknown robust_kernel"):
        buf.ba_config.update(robust_kernel="cauchy")
        _call(buf, inp)


@pytest.mark.gpu
@pytest.mark.parametrize("tracks,limited,intr,motion", [(True, True, True, False), (False, False, False, True),
                                                        (False, False, False, False)])
def test_fused_gpu_matches_generic_path(tracks, limited, intr, motion):
    """tests/test_graph_buffer_fused_ba.py:203-249: the fused path (here the HIP
    kernels) and the generic path (the float64 restatement of core/solver.py's
    Solver) agree within 2e-4 after one iteration on the same graph."""
    from oracle import generic_ba as G

    buf = make_buffer("cuda", tracks=tracks)
    inp = make_inputs(buf)
    _call(buf, inp, limited_disp=limited, optimize_intrinsics=intr, motion_only=motion)
    ref = make_buffer("cpu", tracks=tracks)
    target, weight, damping, ii, jj = [t.cpu() for t in inp]
    if tracks:
        st, sw = ref.sparse_tracks.compute_dense_disp_target_weight(ii * 0, ref.tstamp[ii], jj * 0, ref.tstamp[jj],
                                                                    (32, 40), (4, 5))
        st, sw = st.flatten(1, 2), sw.flatten(1, 2)
        cw = weight + sw
        target = torch.where(cw > 0, (weight * target + sw * st) / cw.clamp_min(1e-12), target)
        weight = cw
    f64 = torch.float64
    poses = ref.poses.double().clone()
    disps = ref.disps[:, 0].double().clone()
    focal = ref.intrinsics[0, :4].double().clone()
    G.bundle_adjustment(poses, disps, focal, ref.disps_sens[:, 0].double(), target.double(), weight.double(),
                        damping.double(), ii, jj, 1, 5, 1, 1e-3, 0.1, motion_only=motion, limited_disp=limited,
                        optimize_intrinsics=intr, intrinsics_factor=8.0, dtype=f64)
    torch.testing.assert_close(buf.poses[:5].cpu().double(), poses[:5], atol=2e-4, rtol=2e-4)
    torch.testing.assert_close(buf.disps[:5, 0].cpu().double(), disps[:5], atol=2e-4, rtol=2e-4)
    torch.testing.assert_close(buf.intrinsics[0, :4].cpu().double(), focal, atol=2e-4, rtol=2e-4)


def make_rig_buffer(device, camera="mei", views=2, seed=7):
    """A two-view buffer (Mei or pinhole, the second view 0.15 to the right),
    6 frames of 4 x 5 maps, with ``ba.fused_rig`` enabled."""
    g = torch.Generator().manual_seed(seed)
    buf = make_buffer("cpu")
    n, h8, w8 = 6, 4, 5
    from oracle import lie as L

    xi = torch.zeros(n, 6, dtype=torch.float64)
    xi[:, :3] = torch.randn(n, 3, generator=g, dtype=torch.float64) * 0.03
    xi[:, 3:] = torch.randn(n, 3, generator=g, dtype=torch.float64) * 0.02
    buf.poses = L.expm(3, xi).float()
    rig = [[0.0, 0, 0, 0, 0, 0, 1], [0.15, 0, 0, 0, 0.0499792, 0, 0.99875026]][:views]
    buf.rig = torch.tensor(rig)
    buf.n_views = views
    buf.camera_type = camera
    k = [38.0, 38.0, 19.5, 15.5] + ([0.6] if camera == "mei" else [])
    buf.intrinsics = torch.tensor([k, [37.0, 38.5, 20.0, 15.0] + ([0.4] if camera == "mei" else [])][:views])
    buf.disps = (0.5 + 0.5 * torch.rand(n, views, h8, w8, generator=g)).contiguous()
    buf.disps_sens = buf.disps * 1.01
    buf.disps_sens[::2] = 0.0
    buf.cross_view_idx = None
    buf.ba_config["fused_rig"] = True
    for k_ in ("poses", "disps", "disps_sens", "intrinsics", "rig", "tstamp"):
        setattr(buf, k_, getattr(buf, k_).to(device))
    return buf


def rig_inputs(buf, seed=8):
    """Edges 1->0, 2->1, 3->2, 4->3, 0->2 and the stereo edges 2->2, 3->3,
    expanded per view; targets from the FP64 projection + noise."""
    from oracle import ba_rig as OR

    g = torch.Generator().manual_seed(seed)
    ii = torch.tensor([1, 2, 3, 4, 0, 2, 3])
    jj = torch.tensor([0, 1, 2, 3, 2, 2, 3])
    cpu = SimpleNamespace(n_views=buf.n_views, cross_view_idx=None)
    pi, qi, di, pj, qj, dj = graph_ba.expand_edge_multiview(cpu, ii, jj)
    cam = OR.MEI if buf.camera_type == "mei" else OR.PINHOLE
    intr = buf.intrinsics.cpu().clone()
    intr[:, :4] /= 8.0
    V, (n, _, h8, w8) = buf.n_views, buf.disps.shape
    M = pi.shape[0]
    T = OR.edge_terms(buf.poses.cpu(), buf.disps.cpu().reshape(n * V, h8, w8), intr, cam, buf.rig.cpu(),
                      torch.zeros(M, 2, h8, w8), torch.ones(M, 2, h8, w8), pi, pj, qi, qj, V)
    target = (torch.stack([T["u"], T["v"]], -1) + 0.2 * torch.randn(M, h8 * w8, 2, generator=g,
                                                                   dtype=torch.float64)).float()
    weight = torch.rand(M, h8 * w8, 2, generator=g)
    damping = 0.05 + 0.01 * torch.rand(n * V, h8, w8, generator=g)
    dev = buf.poses.device
    return target.to(dev), weight.to(dev), damping.to(dev), ii.to(dev), jj.to(dev)


def test_multiview_expansion_and_rig_errors():
    buf = SimpleNamespace(n_views=3, cross_view_idx=None)
    pi, qi, di, pj, qj, dj = graph_ba.expand_edge_multiview(buf, torch.tensor([2, 4]), torch.tensor([1, 4]))
    assert pi.tolist() == [2, 2, 2, 4, 4, 4] and qi.tolist() == [0, 1, 2, 0, 1, 2]
    assert pj.tolist() == [1, 1, 1, 4, 4, 4] and qj.tolist() == [0, 1, 2, 1, 2, 0]
    assert di.tolist() == [6, 7, 8, 12, 13, 14] and dj.tolist() == [3, 4, 5, 13, 14, 12]
    rb = make_rig_buffer("cpu")
    inp = rig_inputs(rb)
    with pytest.raises(RuntimeError, match="does not optimize intrinsics"):
        _call(rb, inp, optimize_intrinsics=True)
    rb.camera_type = "panorama"
    with pytest.raises(RuntimeError, match="no projection model"):
        _call(rb, inp)
    rb.camera_type = "mei"
    rb.ba_config["fused_rig"] = False
    with pytest.raises(RuntimeError, match="single-view pinhole-camera"):
        _call(rb, inp)
    with pytest.raises(RuntimeError, match="expanded graph edges"):
        rb.ba_config["fused_rig"] = True
        t, w, d, ii, jj = inp
        _call(rb, (t[:3], w[:3], d, ii, jj))


@pytest.mark.gpu
@pytest.mark.parametrize("camera,limited", [("mei", False), ("pinhole", True)])
def test_rig_buffer_on_the_fused_path_matches_oracle(camera, limited):
    """ba.fused_rig: a two-view buffer through bundle_adjustment against the
    FP64 oracle given an independent preparation of the same step."""
    from oracle import ba_rig as OR

    buf = make_rig_buffer("cuda", camera)
    inp = rig_inputs(buf)
    ref = make_rig_buffer("cpu", camera)
    _call(buf, inp, n_iters=2, limited_disp=limited)
    target, weight, damping, ii, jj = [t.cpu() for t in inp]
    pi, qi, _, pj, qj, _ = graph_ba.expand_edge_multiview(SimpleNamespace(n_views=2, cross_view_idx=None), ii, jj)
    V, (n, _, h8, w8) = 2, ref.disps.shape
    tg = target.view(-1, h8, w8, 2).permute(0, 3, 1, 2).contiguous()
    wt = weight.view(-1, h8, w8, 2).permute(0, 3, 1, 2).contiguous()
    kx, _ = OR.bookkeeping(pi, qi, V, 1, 5)
    eta = 0.2 * damping[kx] + 1e-7
    da = torch.ones(n * V)
    if limited:
        da.zero_()
        da[V:5 * V] = 1.0
    intr = ref.intrinsics.clone()
    intr[:, :4] /= 8.0
    disps = ref.disps.reshape(n * V, h8, w8).clone()
    OR.ba_rig(ref.poses, disps, intr, OR.MEI if camera == "mei" else OR.PINHOLE, ref.rig, ref.disps_sens.reshape(
        n * V, h8, w8), tg, wt, eta, pi, pj, qi, qj, da, 1, 5, 2, 1e-3, 0.1, False, 0.001)
    disps = disps.clamp(min=0.001).reshape(n, V, h8, w8)
    torch.testing.assert_close(buf.poses.cpu().double(), ref.poses.double(), atol=2e-5, rtol=1e-5)
    torch.testing.assert_close(buf.disps.cpu().double(), disps.double(), atol=2e-4, rtol=2e-4)
    assert (buf.poses.cpu() - make_rig_buffer("cpu", camera).poses).abs().max() > 1e-5
