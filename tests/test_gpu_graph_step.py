"""gfx950 factor-graph update glue vs the oracle: the fused motion features,
the target / weight epilogue, and one FactorGraph.update step end to end with a
deterministic stand-in for the update network."""
from types import SimpleNamespace

import pytest
import torch

from oracle import ba as OB
from oracle import graph_step as OG
from vipe_amd.synth import make_graph

pytestmark = pytest.mark.gpu


def test_motion_features_match_oracle():
    from vipe_amd import graph_step

    g = make_graph("kf32_e192_48x64", seed=3, with_fmaps=False)
    gen = torch.Generator().manual_seed(1)
    target = torch.randn(g.ii.shape[0], 48, 64, 2, generator=gen) * 30 + 20
    c1, m = graph_step.motion_features(g.poses.cuda(), g.disps.cuda(), g.intrinsics.cuda(), g.ii.cuda(), g.jj.cuda(),
                                       target.cuda())
    rc, rm = OG.motion_features(g.poses, g.disps, g.intrinsics, g.ii, g.jj, target)
    c1, m = c1.cpu().double(), m.cpu().double()
    scale = rc.abs().clamp(min=1.0)
    assert ((c1 - rc).abs() / scale).max().item() < 2e-5
    assert (m - rm).abs().max().item() < 2e-5 * 64
    assert (m.abs() <= 64).all() and (m.abs() == 64).any()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16])
def test_targets_epilogue_matches_oracle(dtype):
    from vipe_amd import graph_step

    gen = torch.Generator().manual_seed(2)
    E, H, W = 12, 9, 13
    c1 = torch.randn(E, H, W, 2, generator=gen) * 10
    delta = torch.randn(E, H, W, 2, generator=gen).to(dtype)
    wn = torch.rand(E, H, W, 2, generator=gen).to(dtype)
    mask = (torch.rand(4, H, W, generator=gen) < 0.3).to(torch.uint8)
    frame = torch.randint(0, 4, (E,), generator=gen)
    t, w, tba, wba = graph_step.update_targets(c1.cuda(), delta.cuda(), wn.cuda(), mask.cuda(), frame.cuda())
    rt, rw = OG.targets(c1, delta, wn, mask, frame)
    assert torch.equal(t.cpu(), (c1 + delta.float()))                 # one float32 add, as torch does it
    assert torch.equal(w.cpu().double(), rw)
    assert torch.equal(tba.cpu(), t.cpu().permute(0, 3, 1, 2))
    assert torch.equal(wba.cpu(), w.cpu().permute(0, 3, 1, 2))


def _fake_update(f_net, inp, corr, motn, ix):
    """Deterministic stand-in for the update module: delta from the residual flow,
    weights from the rigid flow, damping per unique frame."""
    delta = (0.25 * torch.tanh(motn[:, :, 2:4])).permute(0, 1, 3, 4, 2).contiguous()
    weight = torch.sigmoid(0.1 * motn[:, :, 0:2]).permute(0, 1, 3, 4, 2).contiguous()
    K = int(ix.max().item()) + 1
    damping = torch.full((1, K) + tuple(motn.shape[-2:]), 0.05, device=motn.device)
    return f_net + 1, delta, weight, damping, None


def test_update_step_matches_oracle_composition():
    import test_graph_ba as TG
    from vipe_amd import graph_step

    buf = TG.make_buffer("cuda")
    ref = TG.make_buffer("cpu")
    ii = torch.tensor([1, 2, 3, 4, 0, 3, 2])
    jj = torch.tensor([0, 0, 1, 2, 2, 4, 3])
    E = ii.shape[0]
    coords, _ = OG.reproject(ref.poses, ref.disps[:, 0], ref.intrinsics[0, :4] / 8.0, ii, jj), None
    target0 = (coords + 0.3).float()
    graph = SimpleNamespace(buffer=buf, ii=ii.cuda(), jj=jj.cuda(), target=target0[None].cuda(),
                            weight=torch.zeros(1, E, 4, 5, 2, device="cuda"), damping=torch.full((6, 4, 5), 0.1,
                                                                                                 device="cuda"),
                            f_net=torch.zeros(1), inp=None, age=0)
    corr_calls = []
    graph_step.update(graph, _fake_update, lambda c: corr_calls.append(c.shape) or None, t0=1, t1=5, itrs=2)
    assert corr_calls == [(1, E, 4, 5, 2)] and graph.age == 1 and graph.f_net.item() == 1.0
    # the same step composed from the oracle pieces on the CPU
    c1, motn = OG.motion_features(ref.poses, ref.disps[:, 0], ref.intrinsics[0, :4] / 8.0, ii, jj, target0)
    _, delta, weight, _, _ = _fake_update(torch.zeros(1), None, None, motn[None].float(), torch.unique(ii, return_inverse=True)[1])
    t, w = OG.targets(c1, delta[0].double(), weight[0].double())
    damping = torch.full((6, 4, 5), 0.1)
    damping[torch.unique(ii)] = 0.05
    tb = t.float().permute(0, 3, 1, 2).contiguous()
    wb = w.float().permute(0, 3, 1, 2).contiguous()
    kx = torch.unique(torch.cat([torch.arange(1, 5), ii]))
    eta = 0.2 * damping[kx] + 1e-7
    disps = ref.disps[:, 0].clone()
    OB.ba_extended(ref.poses, disps, ref.intrinsics[0, :4] / 8.0, ref.disps_sens[:, 0], tb, wb, eta, ii, jj,
                   torch.ones(6), 1, 5, 2, 1e-3, 0.1, False, 0.001, False, 1e-6, 1e-6, 0.125, False)
    disps = disps.clamp(min=0.001)
    assert (graph.target[0].cpu().double() - t).abs().max().item() < 1e-4
    torch.testing.assert_close(buf.poses[:5].cpu().double(), ref.poses[:5].double(), atol=2e-5, rtol=1e-5)
    torch.testing.assert_close(buf.disps[:5, 0].cpu().double(), disps[:5].double(), atol=2e-4, rtol=2e-4)


def test_update_batch_matches_oracle_composition():
    """update_batch with two source-frame batches (batch = 2) and two steps vs
    the same rounds composed from the oracle pieces."""
    import test_graph_ba as TG
    from vipe_amd import graph_step

    buf = TG.make_buffer("cuda")
    ref = TG.make_buffer("cpu")
    ii = torch.tensor([1, 2, 3, 4, 0, 3, 2, 0])
    jj = torch.tensor([0, 0, 1, 2, 2, 4, 3, 4])
    E = ii.shape[0]
    intr8 = ref.intrinsics[0, :4] / 8.0
    target0 = (OG.reproject(ref.poses, ref.disps[:, 0], intr8, ii, jj) - 0.2).float()
    graph = SimpleNamespace(buffer=buf, ii=ii.cuda(), jj=jj.cuda(), target=target0[None].cuda().clone(),
                            weight=torch.zeros(1, E, 4, 5, 2, device="cuda"),
                            damping=torch.full((6, 4, 5), 0.1, device="cuda"), f_net=torch.zeros(1, E, 1, device="cuda"),
                            inps=torch.zeros(6, 1, device="cuda"), n_frames=5)
    seen = []

    def alt_corr(c, a, b):
        seen.append(tuple(a.tolist()))
        return None

    def op(f_net, inp, corr, motn, ix):
        n, d, w, dmp, u = _fake_update(f_net, inp, corr, motn, ix)
        return n, d, w, dmp, u

    graph_step.update_batch(graph, op, alt_corr, itrs=2, steps=2, batch=2)
    assert len(seen) == 2 * 3   # batches [0,2), [2,4), [4,6) in each of two steps
    # oracle composition
    rt = target0.double().clone()
    damping = torch.full((6, 4, 5), 0.1)
    for _ in range(2):
        c1, motn = OG.motion_features(ref.poses, ref.disps[:, 0], intr8, ii, jj, rt)
        w_all = torch.zeros(E, 4, 5, 2, dtype=torch.float64)
        for i0 in range(0, 5, 2):
            v = (ii >= i0) & (ii < i0 + 2)
            dis, dix = torch.unique(ii[v], return_inverse=True)
            _, delta, weight, _, _ = _fake_update(torch.zeros(1), None, None, motn[v][None].float(), dix)
            t, w = OG.targets(c1[v], delta[0].double(), weight[0].double())
            rt[v], w_all[v] = t, w
            damping[dis] = 0.05
        tb = rt.float().permute(0, 3, 1, 2).contiguous()
        wb = w_all.float().permute(0, 3, 1, 2).contiguous()
        kx = torch.unique(torch.cat([torch.arange(1, 5), ii]))
        disps = ref.disps[:, 0]
        OB.ba_extended(ref.poses, disps, intr8, ref.disps_sens[:, 0], tb, wb, 0.2 * damping[kx] + 1e-7, ii, jj,
                       torch.ones(6), 1, 5, 2, 1e-5, 1e-2, False, 0.001, False, 1e-6, 1e-6, 0.125, False)
        ref.disps.clamp_(min=0.001)
    assert (graph.target[0].cpu().double() - rt).abs().max().item() < 2e-4
    torch.testing.assert_close(buf.poses[:5].cpu().double(), ref.poses[:5].double(), atol=5e-5, rtol=1e-5)
    torch.testing.assert_close(buf.disps[:5, 0].cpu().double(), ref.disps[:5, 0].double(), atol=5e-4, rtol=5e-4)
