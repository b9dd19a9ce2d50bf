"""gfx950 lietorch ops vs the FP64 oracle (all 4 groups, f32 and f64, forward and backward)."""
import pytest
import torch

from oracle import lie as L

pytestmark = pytest.mark.gpu

GIDS = [1, 2, 3, 4]
TOL = {torch.float64: 1e-10, torch.float32: 5e-5}


def _inputs(gid, n, dt, seed):
    g = torch.Generator().manual_seed(seed)
    G = L.group(gid)
    a = (torch.randn(n, G.K, generator=g, dtype=torch.float64) * 0.7)
    X = L.rand_group(gid, n, g)
    Y = L.rand_group(gid, n, g)
    p3 = torch.randn(n, 3, generator=g, dtype=torch.float64)
    p4 = torch.randn(n, 4, generator=g, dtype=torch.float64)
    gN = torch.randn(n, G.N, generator=g, dtype=torch.float64)
    gN[:, G.K:] = 0
    gK = torch.randn(n, G.K, generator=g, dtype=torch.float64)
    return {k: v.to(dt) for k, v in dict(a=a, X=X, Y=Y, p3=p3, p4=p4, gN=gN, gK=gK).items()}


def _close(mine, ref, dt, what):
    err = (mine.double().cpu() - ref).abs().max().item()
    scale = max(1.0, ref.abs().max().item())
    assert err <= TOL[dt] * scale, f"{what}: max err {err:.3e}"


@pytest.mark.parametrize("gid", GIDS)
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_forward_ops_match_oracle(gid, dt):
    from vipe_amd.ext import lietorch_ext as E

    n = 1000
    h = _inputs(gid, n, dt, 10 + gid)
    d = {k: v.cuda() for k, v in h.items()}
    D = {k: v.double() for k, v in h.items()}
    _close(E.expm(gid, d["a"]), L.expm(gid, D["a"]), dt, "expm")
    _close(E.logm(gid, d["X"]), L.logm(gid, D["X"]), dt, "logm")
    _close(E.inv(gid, d["X"]), L.inv(gid, D["X"]), dt, "inv")
    _close(E.mul(gid, d["X"], d["Y"]), L.mul(gid, D["X"], D["Y"]), dt, "mul")
    _close(E.adj(gid, d["X"], d["gK"]), L.adj(gid, D["X"], D["gK"]), dt, "adj")
    _close(E.adjT(gid, d["X"], d["gK"]), L.adjT(gid, D["X"], D["gK"]), dt, "adjT")
    _close(E.act(gid, d["X"], d["p3"]), L.act(gid, D["X"], D["p3"]), dt, "act")
    _close(E.act4(gid, d["X"], d["p4"]), L.act4(gid, D["X"], D["p4"]), dt, "act4")
    _close(E.as_matrix(gid, d["X"]), L.as_matrix(gid, D["X"]), dt, "as_matrix")
    _close(E.projector(gid, d["X"]), L.projector(gid, D["X"]), dt, "projector")
    _close(E.Jinv(gid, d["X"], d["gK"]), L.Jinv(gid, D["X"], D["gK"]), dt, "Jinv")


@pytest.mark.parametrize("gid", GIDS)
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_backward_ops_match_oracle(gid, dt):
    from vipe_amd.ext import lietorch_ext as E

    n = 1000
    h = _inputs(gid, n, dt, 20 + gid)
    d = {k: v.cuda() for k, v in h.items()}
    D = {k: v.double() for k, v in h.items()}
    _close(E.expm_backward(gid, d["gN"], d["a"])[0], L.expm_backward(gid, D["gN"], D["a"])[0], dt, "expm_b")
    _close(E.logm_backward(gid, d["gK"], d["X"])[0], L.logm_backward(gid, D["gK"], D["X"])[0], dt, "logm_b")
    _close(E.inv_backward(gid, d["gN"], d["X"])[0], L.inv_backward(gid, D["gN"], D["X"])[0], dt, "inv_b")
    for m, r in zip(E.mul_backward(gid, d["gN"], d["X"], d["Y"]), L.mul_backward(gid, D["gN"], D["X"], D["Y"])):
        _close(m, r, dt, "mul_b")
    for m, r in zip(E.adj_backward(gid, d["gK"], d["X"], d["gK"]), L.adj_backward(gid, D["gK"], D["X"], D["gK"])):
        _close(m, r, dt, "adj_b")
    for m, r in zip(E.adjT_backward(gid, d["gK"], d["X"], d["gK"]), L.adjT_backward(gid, D["gK"], D["X"], D["gK"])):
        _close(m, r, dt, "adjT_b")
    for m, r in zip(E.act_backward(gid, d["p3"], d["X"], d["p3"]), L.act_backward(gid, D["p3"], D["X"], D["p3"])):
        _close(m, r, dt, "act_b")
    for m, r in zip(E.act4_backward(gid, d["p4"], d["X"], d["p4"]), L.act4_backward(gid, D["p4"], D["X"], D["p4"])):
        _close(m, r, dt, "act4_b")


def test_device_and_host_paths_agree():
    from vipe_amd.ext import lietorch_ext as E

    h = _inputs(3, 257, torch.float64, 7)
    on_gpu = E.expm(3, h["a"].cuda()).cpu()
    on_cpu = E.expm(3, h["a"])
    assert (on_gpu - on_cpu).abs().max().item() < 1e-12


def test_unknown_group_matches_reference_conventions():
    from vipe_amd.ext import lietorch_ext as E

    a = torch.zeros(4, 6, device="cuda")
    assert E.expm(9, a) is None
    X = torch.zeros(4, 7, device="cuda")
    assert torch.equal(E.inv(9, X), torch.zeros_like(X))
