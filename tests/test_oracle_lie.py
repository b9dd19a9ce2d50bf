"""Pins oracle/lie.py with checks that do not restate its formulas:
matrix exponentials, finite differences, Lie brackets and closed forms."""
import pytest
import torch

from oracle import lie as L

F64 = torch.float64
GIDS = [1, 2, 3, 4]


def hat_mat(gid, a):
    """Matrix Lie-algebra element (4x4 for all groups) of a tangent vector."""
    n = a.shape[0]
    M = torch.zeros(n, 4, 4, dtype=F64)
    if gid == 1:
        M[:, :3, :3] = L.hat(a)
    elif gid == 2:
        M[:, :3, :3] = L.hat(a[:, :3]) + a[:, 3, None, None] * torch.eye(3, dtype=F64)
    elif gid == 3:
        M[:, :3, :3] = L.hat(a[:, 3:6])
        M[:, :3, 3] = a[:, :3]
    else:
        M[:, :3, :3] = L.hat(a[:, 3:6]) + a[:, 6, None, None] * torch.eye(3, dtype=F64)
        M[:, :3, 3] = a[:, :3]
    return M


def vee_mat(gid, M):
    def vee3(S):
        return torch.stack([S[:, 2, 1], S[:, 0, 2], S[:, 1, 0]], -1)
    A = M[:, :3, :3]
    sym = 0.5 * (A + A.transpose(1, 2))
    sk = 0.5 * (A - A.transpose(1, 2))
    sigma = sym.diagonal(dim1=1, dim2=2).mean(-1)
    if gid == 1:
        return vee3(sk)
    if gid == 2:
        return torch.cat([vee3(sk), sigma[:, None]], -1)
    if gid == 3:
        return torch.cat([M[:, :3, 3], vee3(sk)], -1)
    return torch.cat([M[:, :3, 3], vee3(sk), sigma[:, None]], -1)


def _rand(gid, n, seed, scale=0.8):
    g = torch.Generator().manual_seed(seed)
    G = L.group(gid)
    return torch.randn(n, G.K, generator=g, dtype=F64) * scale, g


@pytest.mark.parametrize("gid", GIDS)
def test_exp_equals_matrix_exponential(gid):
    a, _ = _rand(gid, 64, 1)
    ours = L.as_matrix(gid, L.expm(gid, a))
    ref = torch.linalg.matrix_exp(hat_mat(gid, a))
    assert (ours - ref).abs().max() < 1e-12


@pytest.mark.parametrize("gid", GIDS)
def test_log_inverts_exp(gid):
    a, _ = _rand(gid, 64, 2)
    assert (L.logm(gid, L.expm(gid, a)) - a).abs().max() < 1e-10


@pytest.mark.parametrize("gid", GIDS)
def test_group_axioms_and_action(gid):
    a, g = _rand(gid, 32, 3)
    b, _ = _rand(gid, 32, 4)
    X, Y = L.expm(gid, a), L.expm(gid, b)
    MX, MY = L.as_matrix(gid, X), L.as_matrix(gid, Y)
    assert (L.as_matrix(gid, L.mul(gid, X, Y)) - MX @ MY).abs().max() < 1e-12
    assert (L.as_matrix(gid, L.inv(gid, X)) - torch.linalg.inv(MX)).abs().max() < 1e-12
    p = torch.randn(32, 3, generator=g, dtype=F64)
    hom = torch.cat([p, torch.ones(32, 1, dtype=F64)], -1)
    assert (L.act(gid, X, p) - (MX @ hom[..., None])[:, :3, 0]).abs().max() < 1e-12
    p4 = torch.randn(32, 4, generator=g, dtype=F64)
    assert (L.act4(gid, X, p4) - (MX @ p4[..., None])[..., 0]).abs().max() < 1e-12


@pytest.mark.parametrize("gid", GIDS)
def test_adjoint_conjugation_and_bracket(gid):
    a, g = _rand(gid, 16, 5)
    G = L.group(gid)
    X = L.expm(gid, a)
    v = torch.randn(16, G.K, generator=g, dtype=F64)
    MX = L.as_matrix(gid, X)
    conj = MX @ hat_mat(gid, v) @ torch.linalg.inv(MX)
    assert (L.adj(gid, X, v) - vee_mat(gid, conj)).abs().max() < 1e-10
    w = torch.randn(16, G.K, generator=g, dtype=F64)
    br = hat_mat(gid, v) @ hat_mat(gid, w) - hat_mat(gid, w) @ hat_mat(gid, v)
    assert ((G.adj(v) @ w[..., None])[..., 0] - vee_mat(gid, br)).abs().max() < 1e-12


def _left_jac_fd(gid, a, eps=1e-6):
    """J with exp(a + d) = exp(J d) exp(a), by central differences."""
    G = L.group(gid)
    X0inv = L.inv(gid, L.expm(gid, a))
    cols = []
    for k in range(G.K):
        d = torch.zeros_like(a)
        d[:, k] = eps
        fp = L.logm(gid, L.mul(gid, L.expm(gid, a + d), X0inv))
        fm = L.logm(gid, L.mul(gid, L.expm(gid, a - d), X0inv))
        cols.append((fp - fm) / (2 * eps))
    return torch.stack(cols, -1)


@pytest.mark.parametrize("gid", [1, 2, 3])
def test_left_jacobian_matches_finite_differences(gid):
    a, _ = _rand(gid, 16, 6)
    G = L.group(gid)
    assert (G.left_jac(a) - _left_jac_fd(gid, a)).abs().max() < 1e-7
    assert (G.left_jac_inv(a) - torch.linalg.inv(G.left_jac(a))).abs().max() < 1e-10


def test_sim3_left_jacobian_series_quirk():
    """Sim3 left Jacobian = truncated series without the Xi^5/720 term (sim3.h:160-169):
    exact to O(|a|^5) for small a, off by about Xi^5/720 for larger a."""
    a_small, _ = _rand(4, 16, 7, scale=0.01)
    assert (L.Sim3.left_jac(a_small) - _left_jac_fd(4, a_small)).abs().max() < 1e-8
    a, _ = _rand(4, 16, 8, scale=0.8)
    Xi = L.Sim3.adj(a)
    X5 = torch.linalg.matrix_power(Xi, 5) / 720.0
    X6 = torch.linalg.matrix_power(Xi, 6) / 5040.0
    true = _left_jac_fd(4, a)
    diff = true - L.Sim3.left_jac(a)
    assert (diff - X5).abs().max() < 3 * X6.abs().max() + 1e-6
    assert X5.abs().max() > 1e-4  # the dropped term is visible at this size


def test_sim3_left_jacobian_inverse_series():
    """Sim3 inverse left Jacobian = Bernoulli series X/(e^X - I) truncated after
    the Xi^4 term (sim3.h:171-178): at |a| ~ 0.5 it differs from the inverse of
    the exact Jacobian by the first dropped term, Xi^6/30240, to within the next
    one, while a missing or mis-scaled Xi^2 / Xi^4 term would be far outside."""
    a, _ = _rand(4, 16, 10, scale=0.5)
    Xi = L.Sim3.adj(a)
    P = [torch.linalg.matrix_power(Xi, k) for k in range(9)]
    true_inv = torch.linalg.inv(_left_jac_fd(4, a))
    diff = true_inv - L.Sim3.left_jac_inv(a)
    X6 = P[6] / 30240.0
    X8 = P[8] / 1209600.0
    bound = 3 * X8.abs().max() + 1e-6
    assert (diff - X6).abs().max() < bound
    assert X6.abs().max() > 4 * bound  # the truncation point itself is resolved
    assert (P[4] / 720.0).abs().max() > 50 * bound
    assert (P[2] / 12.0).abs().max() > 50 * bound


@pytest.mark.parametrize("gid", GIDS)
def test_projector_is_data_jacobian(gid):
    a, _ = _rand(gid, 8, 9)
    G = L.group(gid)
    X = L.expm(gid, a)
    eps = 1e-6
    for k in range(G.K):
        d = torch.zeros(8, G.K, dtype=F64)
        d[:, k] = eps
        fd = (L.mul(gid, L.expm(gid, d), X) - L.mul(gid, L.expm(gid, -d), X)) / (2 * eps)
        assert (L.projector(gid, X)[:, :, k] - fd).abs().max() < 1e-8


def _fd_grad(f, x, eps=1e-6):
    out = torch.zeros_like(x)
    for k in range(x.shape[-1]):
        d = torch.zeros_like(x)
        d[:, k] = eps
        out[:, k] = (f(x + d) - f(x - d)) / (2 * eps)
    return out


@pytest.mark.parametrize("gid", GIDS)
def test_backward_ops_match_finite_differences(gid):
    """Every backward formula against central differences of a scalar loss whose
    tangent-space gradient is the upstream gradient."""
    G = L.group(gid)
    K, N = G.K, G.N
    # Sim3 uses truncated Jacobian series (sim3.h:160-178): small angles keep the
    # truncation below the finite-difference tolerance; the quirk has its own test.
    sc = 0.6 if gid != 4 else 0.02
    a, g = _rand(gid, 8, 10, scale=sc)
    b, _ = _rand(gid, 8, 11, scale=sc)
    X, Y = L.expm(gid, a), L.expm(gid, b)
    gk = torch.randn(8, K, generator=g, dtype=F64)
    gN = torch.cat([gk, torch.zeros(8, N - K, dtype=F64)], -1)

    def tangent_loss(Z, Z0):  # <g, log(Z Z0^-1)>, left-tangent gradient g at Z = Z0
        return (gk * L.logm(gid, L.mul(gid, Z, L.inv(gid, Z0)))).sum(-1)

    def perturb(Xb, d):
        return L.mul(gid, L.expm(gid, d), Xb)

    tol = 1e-6
    # exp
    fd = _fd_grad(lambda x: tangent_loss(L.expm(gid, x), X), a)
    assert (L.expm_backward(gid, gN, a)[0] - fd).abs().max() < tol
    # log: f = <g, log(X)>
    fd = _fd_grad(lambda d: (gk * L.logm(gid, perturb(X, d))).sum(-1), torch.zeros(8, K, dtype=F64))
    assert (L.logm_backward(gid, gk, X)[0][:, :K] - fd).abs().max() < tol
    # inv
    Yi = L.inv(gid, X)
    fd = _fd_grad(lambda d: tangent_loss(L.inv(gid, perturb(X, d)), Yi), torch.zeros(8, K, dtype=F64))
    assert (L.inv_backward(gid, gN, X)[0][:, :K] - fd).abs().max() < 1e-6
    # mul
    Z = L.mul(gid, X, Y)
    dX, dY = L.mul_backward(gid, gN, X, Y)
    fd = _fd_grad(lambda d: tangent_loss(L.mul(gid, perturb(X, d), Y), Z), torch.zeros(8, K, dtype=F64))
    assert (dX[:, :K] - fd).abs().max() < 1e-6
    fd = _fd_grad(lambda d: tangent_loss(L.mul(gid, X, perturb(Y, d)), Z), torch.zeros(8, K, dtype=F64))
    assert (dY[:, :K] - fd).abs().max() < 1e-6
    # adj / adjT
    v = torch.randn(8, K, generator=g, dtype=F64)
    dX, dv = L.adj_backward(gid, gk, X, v)
    fd = _fd_grad(lambda d: (gk * L.adj(gid, perturb(X, d), v)).sum(-1), torch.zeros(8, K, dtype=F64))
    assert (dX[:, :K] - fd).abs().max() < 1e-6
    fd = _fd_grad(lambda x: (gk * L.adj(gid, X, x)).sum(-1), v)
    assert (dv - fd).abs().max() < 1e-6
    dX, dv = L.adjT_backward(gid, gk, X, v)
    fd = _fd_grad(lambda d: (gk * L.adjT(gid, perturb(X, d), v)).sum(-1), torch.zeros(8, K, dtype=F64))
    assert (dX[:, :K] - fd).abs().max() < 1e-6
    fd = _fd_grad(lambda x: (gk * L.adjT(gid, X, x)).sum(-1), v)
    assert (dv - fd).abs().max() < 1e-6
    # act / act4
    p = torch.randn(8, 3, generator=g, dtype=F64)
    gq = torch.randn(8, 3, generator=g, dtype=F64)
    dX, dp = L.act_backward(gid, gq, X, p)
    fd = _fd_grad(lambda d: (gq * L.act(gid, perturb(X, d), p)).sum(-1), torch.zeros(8, K, dtype=F64))
    assert (dX[:, :K] - fd).abs().max() < 1e-6
    fd = _fd_grad(lambda x: (gq * L.act(gid, X, x)).sum(-1), p)
    assert (dp - fd).abs().max() < 1e-6
    p4 = torch.randn(8, 4, generator=g, dtype=F64)
    gq4 = torch.randn(8, 4, generator=g, dtype=F64)
    dX, dp = L.act4_backward(gid, gq4, X, p4)
    fd = _fd_grad(lambda d: (gq4 * L.act4(gid, perturb(X, d), p4)).sum(-1), torch.zeros(8, K, dtype=F64))
    assert (dX[:, :K] - fd).abs().max() < 1e-6
    fd = _fd_grad(lambda x: (gq4 * L.act4(gid, X, x)).sum(-1), p4)
    assert (dp - fd).abs().max() < 1e-6


def test_small_angle_branches_and_pi_branch():
    # below EPS the series is used; it must agree with the closed form evaluated just above
    phi = torch.tensor([[3e-7, -2e-7, 1e-7]], dtype=F64)
    q = L.expm(1, phi)
    assert (q[0, :3] - 0.5 * phi[0]).abs().max() < 1e-15
    assert (L.logm(1, q) - phi).abs().max() < 1e-15
    # rotation by pi about x: |w| < EPS takes the +-pi/n branch, -pi when w <= 0 (so3.h:119-125)
    Xpi = torch.tensor([[1.0, 0.0, 0.0, 0.0]], dtype=F64)
    assert (L.logm(1, Xpi) - torch.tensor([[-torch.pi, 0.0, 0.0]], dtype=F64)).abs().max() < 1e-12
    Xpi2 = torch.tensor([[1.0, 0.0, 0.0, 5e-7]], dtype=F64)
    assert (L.logm(1, Xpi2) - torch.tensor([[torch.pi, 0.0, 0.0]], dtype=F64)).abs().max() < 1e-12
    # quaternions are normalised on construction (so3.h:36-38)
    assert (L.act(1, 3.0 * L.expm(1, phi * 1e5), torch.ones(1, 3, dtype=F64)).norm() - 3 ** 0.5).abs() < 1e-12


def test_jinv_is_inverse_left_jacobian_of_log():
    a, g = _rand(3, 8, 12)
    X = L.expm(3, a)
    v = torch.randn(8, 6, generator=g, dtype=F64)
    Jl = L.SE3.left_jac(L.logm(3, X))
    assert ((Jl @ L.Jinv(3, X, v)[..., None])[..., 0] - v).abs().max() < 1e-10


def test_droid_retraction_equals_group_exp_times_pose():
    g = torch.Generator().manual_seed(13)
    xi = torch.randn(16, 6, generator=g, dtype=F64) * 0.1
    T = L.expm(3, torch.randn(16, 6, generator=g, dtype=F64) * 0.5)
    ours = L.droid_retr_se3(xi, T)
    ref = L.mul(3, L.expm(3, xi), T)
    assert (ours - ref).abs().max() < 1e-12
    # below theta = 1e-4 the translation correction is dropped (geom_kernels.cu:163-176)
    xi_s = xi * 1e-4
    assert (L.droid_exp_se3(xi_s)[0] - xi_s[:, :3]).abs().max() == 0.0
