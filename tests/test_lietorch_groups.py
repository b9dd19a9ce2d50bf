"""The group classes of vipe_amd.lietorch (interface of vipe/ext/lietorch/groups.py)
on CPU tensors, i.e. on the library's host loop.  Values are checked against
the FP64 oracle and group identities; gradients against central differences."""
import pytest
import torch

from oracle import lie as L

GROUPS = ["SO3", "RxSO3", "SE3", "Sim3"]


def _cls(name):
    from vipe_amd import lietorch as LT

    return getattr(LT, name)


def _rand(cls, shape, seed, scale=0.5):
    g = torch.Generator().manual_seed(seed)
    a = torch.randn(tuple(shape) + (cls.manifold_dim,), generator=g, dtype=torch.float64) * scale
    return cls.exp(a), a


@pytest.mark.parametrize("name", GROUPS)
def test_exp_log_and_layout_match_oracle(name):
    cls = _cls(name)
    X, a = _rand(cls, (3, 5), 1)
    assert X.shape == (3, 5) and X.data.shape == (3, 5, cls.embedded_dim)
    assert X.tangent_shape == (3, 5, cls.manifold_dim)
    ref = L.expm(cls.group_id, a.reshape(-1, cls.manifold_dim)).reshape(3, 5, -1)
    assert torch.allclose(X.data, ref, atol=1e-12)
    assert torch.allclose(X.log(), a, atol=1e-10)
    q = X.quaternion()
    assert q.shape == (3, 5, 4) and torch.allclose(q.norm(dim=-1), torch.ones(3, 5, dtype=torch.float64))


@pytest.mark.parametrize("name", GROUPS)
def test_group_identities(name):
    cls = _cls(name)
    X, _ = _rand(cls, (4,), 2)
    Y, _ = _rand(cls, (4,), 3)
    I = cls.IdentityLike(X)
    assert torch.allclose((X * X.inv()).data, I.data, atol=1e-12)
    # associativity and the matrix homomorphism
    assert torch.allclose((X * Y).matrix(), X.matrix() @ Y.matrix(), atol=1e-12)
    # retr(a) = Exp(a) X
    a = torch.randn(4, cls.manifold_dim, dtype=torch.float64) * 0.3
    assert torch.allclose(X.retr(a).data, (cls.exp(a) * X).data, atol=1e-12)
    # adjoint: X Exp(a) X^-1 = Exp(Adj(X) a); adjT is its transpose
    lhs = (X * cls.exp(a) * X.inv()).log()
    assert torch.allclose(lhs, X.adj(a), atol=1e-9)
    b = torch.randn(4, cls.manifold_dim, dtype=torch.float64)
    assert torch.allclose((X.adj(a) * b).sum(-1), (a * X.adjT(b)).sum(-1), atol=1e-10)


@pytest.mark.parametrize("name", GROUPS)
def test_action_matrix_and_translation(name):
    cls = _cls(name)
    X, _ = _rand(cls, (2, 3), 4)
    p = torch.randn(2, 3, 3, dtype=torch.float64)
    M = X.matrix()
    ph = torch.cat([p, torch.ones(2, 3, 1, dtype=torch.float64)], -1)
    assert torch.allclose(X.act(p), (M @ ph[..., None])[..., :3, 0], atol=1e-12)
    assert torch.allclose(X * ph, (M @ ph[..., None])[..., 0], atol=1e-12)
    # the autograd form of matrix() (actions on the basis) equals the kernel's
    Xg = cls(X.data.clone().requires_grad_(True))
    assert torch.allclose(Xg.matrix(), M, atol=1e-12)
    assert torch.allclose(X.translation(), M[..., :, 3], atol=1e-12)
    with pytest.raises(ValueError):
        X.act(torch.zeros(2, 3, 5, dtype=torch.float64))


def test_broadcasting_over_batch_axes():
    SE3 = _cls("SE3")
    X, _ = _rand(SE3, (2, 1), 5)
    Y, _ = _rand(SE3, (1, 3), 6)
    Z = X * Y
    assert Z.shape == (2, 3)
    for i in range(2):
        for j in range(3):
            assert torch.allclose(Z.data[i, j], (X[i, 0] * Y[0, j]).data, atol=1e-12)
    p = torch.randn(5, 3, dtype=torch.float64)
    assert (X[0, 0].view((1,)) * p).shape == (5, 3)
    with pytest.raises(RuntimeError):
        X * _rand(SE3, (3, 2), 7)[0]


def test_constructors_conversions_and_containers():
    from vipe_amd.lietorch import SE3, SO3, RxSO3, Sim3, cat, stack

    I = SE3.Identity(2, 3)
    assert I.shape == (2, 3) and torch.equal(I.data[1, 2], SE3.id_elem)
    assert SE3.Identity([4]).shape == (4,) and SE3.Identity((4,), dtype=torch.float64).dtype == torch.float64
    R, _ = _rand(SO3, (3,), 8)
    T = SE3(R)
    assert torch.equal(T.data[:, :3], torch.zeros(3, 3, dtype=torch.float64)) and torch.equal(T.data[:, 3:], R.data)
    assert torch.equal(SO3(T).data, R.data)
    S = Sim3(T)
    assert torch.equal(S.data[:, :7], T.data) and torch.equal(S.data[:, 7], torch.ones(3, dtype=torch.float64))
    assert torch.equal(Sim3(R).data, S.data)
    assert torch.equal(RxSO3(S).data, S.data[:, 3:])
    with pytest.raises(TypeError):
        SO3(RxSO3(S))
    s = torch.tensor([2.0, 3.0, 4.0], dtype=torch.float64)
    Tt, _ = _rand(SE3, (3,), 9)
    Ts = Tt.scale(s)
    assert torch.allclose(Ts.data[:, :3], Tt.data[:, :3] * s[:, None]) and torch.equal(Ts.data[:, 3:], Tt.data[:, 3:])
    C = cat([Tt, Ts], 0)
    assert C.shape == (6,) and stack([Tt, Ts], 1).shape == (3, 2)
    assert [u.shape for u in C.view((2, 3)).unbind(0)] == [(3,), (3,)]
    C[0] = Ts[1]
    assert torch.equal(C.data[0], Ts.data[1])
    assert C.float().dtype == torch.float32 and C.float().double().dtype == torch.float64
    assert SE3.Random(5, sigma=0.1, dtype=torch.float64).shape == (5,)


@pytest.mark.parametrize("name", GROUPS)
def test_gradients_match_central_differences(name):
    cls = _cls(name)
    g = torch.Generator().manual_seed(10)
    # Sim3's exp backward uses the reference's truncated left Jacobian
    # (sim3.h:167-178), exact only to O(|a|^3): small tangents keep it in tolerance.
    s = 0.02 if name == "Sim3" else 0.4
    a = (torch.randn(3, cls.manifold_dim, generator=g, dtype=torch.float64) * s).requires_grad_(True)
    b = (torch.randn(3, cls.manifold_dim, generator=g, dtype=torch.float64) * s).requires_grad_(True)
    p = torch.randn(3, 3, generator=g, dtype=torch.float64).requires_grad_(True)

    def f(a, b, p):
        X, Y = cls.exp(a), cls.exp(b)
        Z = X * Y.inv()
        return torch.cat([Z.log(), Z.act(p), X.adj(b), Y.adjT(a)], -1)

    assert torch.autograd.gradcheck(f, (a, b, p), eps=1e-6, atol=1e-6)


@pytest.mark.parametrize("name", GROUPS)
def test_vec_and_init_from_vec_gradients(name):
    cls = _cls(name)
    s = 0.02 if name == "Sim3" else 0.4  # see test_gradients_match_central_differences
    a = (torch.randn(2, cls.manifold_dim, dtype=torch.float64) * s).requires_grad_(True)
    # d vec(Exp(a)) / da through the projector equals the plain derivative of the embedding
    assert torch.autograd.gradcheck(lambda a: cls.exp(a).vec(), (a,), eps=1e-6, atol=1e-6)
    # InitFromVec(vec(X)) is X; the tangent gradient survives the round trip
    X = cls.exp(a)
    Y = cls.InitFromVec(X.vec())
    assert torch.equal(Y.data, X.data)
    ga = torch.autograd.grad(Y.log().sum(), a)[0]
    gb = torch.autograd.grad(cls.exp(a).log().sum(), a)[0]
    assert torch.allclose(ga, gb, atol=1e-9)


def test_lie_group_parameter():
    from vipe_amd.lietorch import SE3, LieGroupParameter

    X, _ = _rand(SE3, (3,), 11)
    P = LieGroupParameter(X)
    assert P.shape == (3, 6) and P.requires_grad and torch.equal(P.detach(), torch.zeros(3, 6, dtype=torch.float64))
    assert torch.allclose(P.retr().data, X.data, atol=1e-12)
    p = torch.randn(3, 3, dtype=torch.float64)
    loss = (P * p).square().sum()
    loss.backward()
    assert P.grad is not None and P.grad.shape == (3, 6)
    d = torch.randn(3, 6, dtype=torch.float64) * 0.1
    P.add_(d, 0.5)
    assert torch.allclose(P.group.data, (SE3.exp(0.5 * d) * X).data, atol=1e-12)
