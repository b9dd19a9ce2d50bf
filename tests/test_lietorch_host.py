"""lietorch_ext on CPU tensors runs the library's host loop (the reference's
lietorch_cpu.cpp path); checked against the FP64 oracle for every op."""
import pytest
import torch

from oracle import lie as L

GIDS = [1, 2, 3, 4]


@pytest.mark.parametrize("gid", GIDS)
@pytest.mark.parametrize("dt,tol", [(torch.float64, 1e-10), (torch.float32, 5e-5)])
def test_host_ops_match_oracle(gid, dt, tol):
    from vipe_amd.ext import lietorch_ext as E

    g = torch.Generator().manual_seed(gid)
    G = L.group(gid)
    n = 64
    a = (torch.randn(n, G.K, generator=g, dtype=torch.float64) * 0.7)
    X, Y = L.rand_group(gid, n, g), L.rand_group(gid, n, g)
    p3, p4 = torch.randn(n, 3, generator=g, dtype=torch.float64), torch.randn(n, 4, generator=g, dtype=torch.float64)
    gN = torch.randn(n, G.N, generator=g, dtype=torch.float64)
    gN[:, G.K:] = 0
    gK = torch.randn(n, G.K, generator=g, dtype=torch.float64)
    c = lambda t: t.to(dt)  # noqa: E731

    def close(m, r):
        assert (m.double() - r).abs().max().item() <= tol * max(1.0, r.abs().max().item())

    close(E.expm(gid, c(a)), L.expm(gid, c(a).double()))
    close(E.logm(gid, c(X)), L.logm(gid, c(X).double()))
    close(E.inv(gid, c(X)), L.inv(gid, c(X).double()))
    close(E.mul(gid, c(X), c(Y)), L.mul(gid, c(X).double(), c(Y).double()))
    close(E.adj(gid, c(X), c(gK)), L.adj(gid, c(X).double(), c(gK).double()))
    close(E.adjT(gid, c(X), c(gK)), L.adjT(gid, c(X).double(), c(gK).double()))
    close(E.act(gid, c(X), c(p3)), L.act(gid, c(X).double(), c(p3).double()))
    close(E.act4(gid, c(X), c(p4)), L.act4(gid, c(X).double(), c(p4).double()))
    close(E.as_matrix(gid, c(X)), L.as_matrix(gid, c(X).double()))
    close(E.projector(gid, c(X)), L.projector(gid, c(X).double()))
    close(E.Jinv(gid, c(X), c(gK)), L.Jinv(gid, c(X).double(), c(gK).double()))
    close(E.expm_backward(gid, c(gN), c(a))[0], L.expm_backward(gid, c(gN).double(), c(a).double())[0])
    close(E.logm_backward(gid, c(gK), c(X))[0], L.logm_backward(gid, c(gK).double(), c(X).double())[0])
    close(E.inv_backward(gid, c(gN), c(X))[0], L.inv_backward(gid, c(gN).double(), c(X).double())[0])
    for m, r in zip(E.mul_backward(gid, c(gN), c(X), c(Y)), L.mul_backward(gid, c(gN).double(), c(X).double(), c(Y).double())):
        close(m, r)
    for m, r in zip(E.act4_backward(gid, c(p4), c(X), c(p4)), L.act4_backward(gid, c(p4).double(), c(X).double(), c(p4).double())):
        close(m, r)


def test_contiguity_error_like_reference():
    from vipe_amd.ext import lietorch_ext as E

    a = torch.zeros(6, 8)[:, ::2]
    with pytest.raises(RuntimeError, match="must be contiguous"):
        E.expm(1, a[:, :3])


def test_unknown_group_and_dtype_conventions():
    from vipe_amd.ext import lietorch_ext as E

    assert E.expm(7, torch.zeros(3, 6)) is None
    assert E.logm(3, torch.zeros(3, 7, dtype=torch.float16)) is None
    X = torch.zeros(3, 7)
    assert torch.equal(E.mul(7, X, X), torch.zeros_like(X))
