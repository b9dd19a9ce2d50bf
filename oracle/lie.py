"""FP64 restatement of the lietorch group maths and ops (oracle; test-only).

Follows, function by function:
  SO3   csrc/lietorch_ext/so3.h:33-205
  RxSO3 csrc/lietorch_ext/rxso3.h:33-300
  SE3   csrc/lietorch_ext/se3.h:33-208
  Sim3  csrc/lietorch_ext/sim3.h:35-213 (left_jacobian drops the Xi^5/720 term
        because of the stray ';' at sim3.h:167-168; reproduced here)
  ops   csrc/lietorch_ext/lietorch_gpu.cu:24-299 (forward and backward formulas)
  ids   vipe/ext/lietorch/groups.py:244,260,276,298 (SO3=1, RxSO3=2, SE3=3, Sim3=4)

Data layouts (embedding N, tangent K):
  SO3   [qx,qy,qz,qw]            N=4 K=3  tangent [phi]
  RxSO3 [qx,qy,qz,qw,s]          N=5 K=4  tangent [phi, sigma]
  SE3   [tx,ty,tz,qx,qy,qz,qw]   N=7 K=6  tangent [tau, phi]
  Sim3  [tx,ty,tz,qx,qy,qz,qw,s] N=8 K=7  tangent [tau, phi, sigma]

Gradients with respect to a group element are K-vectors stored in the first
K entries of an [n, N] buffer (the rest zero); upstream gradients of a group
output are read from the first K entries of an [n, N] buffer
(lietorch_gpu.cu:38-46, :95-98).

Pinned by tests/test_oracle_lie.py through: Rodrigues and quaternion closed
forms, exp(log X) = X, group axioms, Adj(X) a = log(X exp(a) X^-1) for small a,
act = matrix product, FP64 central differences of every backward formula,
and the small-angle series against the closed forms.
"""
from __future__ import annotations

import math

import torch

EPS = 1e-6  # csrc/lietorch_ext/common.h:12
PI = 3.14159265358979323846

F64 = torch.float64


# ----------------------------------------------------------------------------
# quaternion / matrix helpers (Eigen semantics: coeffs = [x, y, z, w])
# ----------------------------------------------------------------------------

def hat(v: torch.Tensor) -> torch.Tensor:
    """[..., 3] -> [..., 3, 3] skew matrix (so3.h:90-95)."""
    z = torch.zeros_like(v[..., 0])
    x, y, w = v[..., 0], v[..., 1], v[..., 2]
    return torch.stack(
        [torch.stack([z, -w, y], -1), torch.stack([w, z, -x], -1), torch.stack([-y, x, z], -1)], -2
    )


def qnormalize(q: torch.Tensor) -> torch.Tensor:
    """Eigen Quaternion::normalize (so3.h:36-38)."""
    return q / q.norm(dim=-1, keepdim=True)


def qmul(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Hamilton product a*b with [x,y,z,w] storage."""
    av, aw = a[..., :3], a[..., 3:]
    bv, bw = b[..., :3], b[..., 3:]
    v = aw * bv + bw * av + torch.cross(av, bv, dim=-1)
    w = aw * bw - (av * bv).sum(-1, keepdim=True)
    return torch.cat([v, w], -1)


def qconj(q: torch.Tensor) -> torch.Tensor:
    return torch.cat([-q[..., :3], q[..., 3:]], -1)


def qrot(q: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
    """Point rotation with the uv form used by SO3::operator* (so3.h:50-55)."""
    v, w = q[..., :3], q[..., 3:]
    uv = 2.0 * torch.cross(v, p, dim=-1)
    return p + w * uv + torch.cross(v, uv, dim=-1)


def qmat(q: torch.Tensor) -> torch.Tensor:
    """Eigen toRotationMatrix of a unit quaternion (so3.h:63)."""
    x, y, z, w = q.unbind(-1)
    r00 = 1 - 2 * (y * y + z * z)
    r01 = 2 * (x * y - w * z)
    r02 = 2 * (x * z + w * y)
    r10 = 2 * (x * y + w * z)
    r11 = 1 - 2 * (x * x + z * z)
    r12 = 2 * (y * z - w * x)
    r20 = 2 * (x * z - w * y)
    r21 = 2 * (y * z + w * x)
    r22 = 1 - 2 * (x * x + y * y)
    return torch.stack(
        [torch.stack([r00, r01, r02], -1), torch.stack([r10, r11, r12], -1), torch.stack([r20, r21, r22], -1)], -2
    )


def _eye(n, like):
    return torch.eye(n, dtype=like.dtype, device=like.device).expand(like.shape[:-1] + (n, n)).clone()


def _where(c, a, b):
    return torch.where(c, a, b)


# ----------------------------------------------------------------------------
# SO3 core (so3.h)
# ----------------------------------------------------------------------------

def so3_exp_quat(phi: torch.Tensor) -> torch.Tensor:
    """SO3::Exp (so3.h:135-152) -> normalized quaternion."""
    theta2 = (phi * phi).sum(-1, keepdim=True)
    theta = theta2.sqrt()
    theta4 = theta2 * theta2
    small = theta < EPS
    ts = torch.where(small, torch.ones_like(theta), theta)
    imag = _where(small, 0.5 - theta2 / 48.0 + theta4 / 3840.0, torch.sin(0.5 * ts) / ts)
    real = _where(small, 1.0 - theta2 / 8.0 + theta4 / 384.0, torch.cos(0.5 * ts))
    return qnormalize(torch.cat([imag * phi, real], -1))


def so3_log_quat(q: torch.Tensor) -> torch.Tensor:
    """SO3::Log (so3.h:97-133), atan form with the EPS branches."""
    v, w = q[..., :3], q[..., 3:]
    n2 = (v * v).sum(-1, keepdim=True)
    n = n2.sqrt()
    nsafe = torch.where(n > 0, n, torch.ones_like(n))
    wsafe = torch.where(w == 0, torch.ones_like(w), w)
    f_small = 2.0 / wsafe - (2.0 / 3.0) * n2 / (wsafe * wsafe * wsafe)
    f_pi = torch.where(w > 0, PI / nsafe, -PI / nsafe)
    f_gen = 2.0 * torch.atan(n / wsafe) / nsafe
    f = torch.where(n2 < EPS * EPS, f_small, torch.where(w.abs() < EPS, f_pi, f_gen))
    return f * v


def so3_left_jac(phi: torch.Tensor) -> torch.Tensor:
    """SO3::left_jacobian (so3.h:154-170)."""
    Phi = hat(phi)
    Phi2 = Phi @ Phi
    theta2 = (phi * phi).sum(-1)[..., None, None]
    theta = theta2.sqrt()
    small = theta < EPS
    ts = torch.where(small, torch.ones_like(theta), theta)
    c1 = _where(small, 0.5 - theta2 / 24.0, (1.0 - torch.cos(ts)) / (ts * ts))
    c2 = _where(small, 1.0 / 6.0 - theta2 / 120.0, (ts - torch.sin(ts)) / (ts * ts * ts))
    return _eye(3, phi) + c1 * Phi + c2 * Phi2


def so3_left_jac_inv(phi: torch.Tensor) -> torch.Tensor:
    """SO3::left_jacobian_inverse (so3.h:172-187)."""
    Phi = hat(phi)
    Phi2 = Phi @ Phi
    theta = (phi * phi).sum(-1).sqrt()[..., None, None]
    small = theta < EPS
    ts = torch.where(small, torch.ones_like(theta), theta)
    half = 0.5 * ts
    c2 = _where(small, torch.full_like(theta, 1.0 / 12.0), (1.0 - ts * torch.cos(half) / (2.0 * torch.sin(half))) / (ts * ts))
    return _eye(3, phi) - 0.5 * Phi + c2 * Phi2


def so3_projector(q: torch.Tensor) -> torch.Tensor:
    """SO3::orthogonal_projector (so3.h:72-80) -> [..., 4, 4]."""
    v, w = q[..., :3], q[..., 3]
    J = torch.zeros(q.shape[:-1] + (4, 4), dtype=q.dtype)
    J[..., :3, :3] = 0.5 * (w[..., None, None] * _eye(3, v) + hat(-v))
    J[..., 3, :3] = 0.5 * (-v)
    return J


def rxso3_calcW(phi: torch.Tensor, sigma: torch.Tensor) -> torch.Tensor:
    """RxSO3::calcW (rxso3.h:175-215).  phi [...,3], sigma [...]."""
    theta = phi.norm(dim=-1)
    Phi = hat(phi)
    Phi2 = Phi @ Phi
    scale = torch.exp(sigma)
    one = torch.ones_like(sigma)
    sig_small = sigma.abs() < EPS
    th_small = theta.abs() < EPS
    ts = torch.where(th_small, one, theta)
    ss = torch.where(sig_small, one, sigma)
    # sigma small
    A1 = torch.where(th_small, 0.5 * one, (1 - torch.cos(ts)) / (ts * ts))
    B1 = torch.where(th_small, one / 6.0, (ts - torch.sin(ts)) / (ts * ts * ts))
    C1 = one
    # sigma large
    C2 = (scale - 1) / ss
    A2a = ((ss - 1) * scale + 1) / (ss * ss)
    B2a = (scale * 0.5 * ss * ss + scale - 1 - ss * scale) / (ss * ss * ss)
    a = scale * torch.sin(ts)
    b = scale * torch.cos(ts)
    c = ts * ts + ss * ss
    A2b = (a * ss + (1 - b) * ts) / (ts * c)
    B2b = (C2 - ((b - 1) * ss + a * ts) / c) / (ts * ts)
    A2 = torch.where(th_small, A2a, A2b)
    B2 = torch.where(th_small, B2a, B2b)
    A = torch.where(sig_small, A1, A2)[..., None, None]
    B = torch.where(sig_small, B1, B2)[..., None, None]
    C = torch.where(sig_small, C1, C2)[..., None, None]
    return A * Phi + B * Phi2 + C * _eye(3, phi)


# ----------------------------------------------------------------------------
# Groups
# ----------------------------------------------------------------------------

class SO3:
    gid, N, K = 1, 4, 3

    @staticmethod
    def parse(X):
        return qnormalize(X)

    @staticmethod
    def exp(a):
        return so3_exp_quat(a)

    @staticmethod
    def log(X):
        return so3_log_quat(SO3.parse(X))

    @staticmethod
    def inv(X):
        return qconj(SO3.parse(X))

    @staticmethod
    def mul(X, Y):
        return qnormalize(qmul(SO3.parse(X), SO3.parse(Y)))

    @staticmethod
    def Adj(X):
        return qmat(SO3.parse(X))

    @staticmethod
    def adj(a):
        return hat(a)

    @staticmethod
    def act(X, p):
        return qrot(SO3.parse(X), p)

    @staticmethod
    def act4(X, p):
        return torch.cat([SO3.act(X, p[..., :3]), p[..., 3:]], -1)

    @staticmethod
    def matrix4(X):
        T = _eye(4, X[..., :1])
        T[..., :3, :3] = qmat(SO3.parse(X))
        return T

    @staticmethod
    def projector(X):
        return so3_projector(SO3.parse(X))

    @staticmethod
    def left_jac(a):
        return so3_left_jac(a)

    @staticmethod
    def left_jac_inv(a):
        return so3_left_jac_inv(a)

    @staticmethod
    def act_jac(q):
        return hat(-q)

    @staticmethod
    def act4_jac(q4):
        J = torch.zeros(q4.shape[:-1] + (4, 3), dtype=q4.dtype)
        J[..., :3, :] = hat(-q4[..., :3])
        return J


class RxSO3:
    gid, N, K = 2, 5, 4

    @staticmethod
    def parse(X):
        return qnormalize(X[..., :4]), X[..., 4]

    @staticmethod
    def pack(q, s):
        return torch.cat([q, s[..., None]], -1)

    @staticmethod
    def exp(a):
        return RxSO3.pack(so3_exp_quat(a[..., :3]), torch.exp(a[..., 3]))

    @staticmethod
    def log(X):
        q, s = RxSO3.parse(X)
        return torch.cat([so3_log_quat(q), torch.log(s)[..., None]], -1)

    @staticmethod
    def inv(X):
        q, s = RxSO3.parse(X)
        return RxSO3.pack(qconj(q), 1.0 / s)

    @staticmethod
    def mul(X, Y):
        qa, sa = RxSO3.parse(X)
        qb, sb = RxSO3.parse(Y)
        return RxSO3.pack(qnormalize(qmul(qa, qb)), sa * sb)

    @staticmethod
    def Adj(X):
        q, _ = RxSO3.parse(X)
        A = _eye(4, q)
        A[..., :3, :3] = qmat(q)
        return A

    @staticmethod
    def adj(a):
        A = torch.zeros(a.shape[:-1] + (4, 4), dtype=a.dtype)
        A[..., :3, :3] = hat(a[..., :3])
        return A

    @staticmethod
    def act(X, p):
        q, s = RxSO3.parse(X)
        return s[..., None] * qrot(q, p)

    @staticmethod
    def act4(X, p):
        return torch.cat([RxSO3.act(X, p[..., :3]), p[..., 3:]], -1)

    @staticmethod
    def matrix4(X):
        q, s = RxSO3.parse(X)
        T = _eye(4, q)
        T[..., :3, :3] = s[..., None, None] * qmat(q)
        return T

    @staticmethod
    def projector(X):
        q, s = RxSO3.parse(X)
        J = torch.zeros(q.shape[:-1] + (5, 5), dtype=q.dtype)
        J[..., :4, :4] = so3_projector(q)
        J[..., 4, 3] = s
        return J

    @staticmethod
    def left_jac(a):
        J = _eye(4, a)
        J[..., :3, :3] = so3_left_jac(a[..., :3])
        return J

    @staticmethod
    def left_jac_inv(a):
        J = _eye(4, a)
        J[..., :3, :3] = so3_left_jac_inv(a[..., :3])
        return J

    @staticmethod
    def act_jac(q):
        return torch.cat([hat(-q), q[..., :, None]], -1)

    @staticmethod
    def act4_jac(q4):
        J = torch.zeros(q4.shape[:-1] + (4, 4), dtype=q4.dtype)
        J[..., :3, :3] = hat(-q4[..., :3])
        J[..., :3, 3] = q4[..., :3]
        return J


class SE3:
    gid, N, K = 3, 7, 6

    @staticmethod
    def parse(X):
        return X[..., :3], qnormalize(X[..., 3:7])

    @staticmethod
    def pack(t, q):
        return torch.cat([t, q], -1)

    @staticmethod
    def exp(a):
        tau, phi = a[..., :3], a[..., 3:6]
        q = so3_exp_quat(phi)
        t = (so3_left_jac(phi) @ tau[..., None])[..., 0]
        return SE3.pack(t, q)

    @staticmethod
    def log(X):
        t, q = SE3.parse(X)
        phi = so3_log_quat(q)
        tau = (so3_left_jac_inv(phi) @ t[..., None])[..., 0]
        return torch.cat([tau, phi], -1)

    @staticmethod
    def inv(X):
        t, q = SE3.parse(X)
        qi = qconj(q)
        return SE3.pack(-qrot(qi, t), qi)

    @staticmethod
    def mul(X, Y):
        ta, qa = SE3.parse(X)
        tb, qb = SE3.parse(Y)
        return SE3.pack(ta + qrot(qa, tb), qnormalize(qmul(qa, qb)))

    @staticmethod
    def Adj(X):
        t, q = SE3.parse(X)
        R = qmat(q)
        A = torch.zeros(t.shape[:-1] + (6, 6), dtype=t.dtype)
        A[..., :3, :3] = R
        A[..., :3, 3:] = hat(t) @ R
        A[..., 3:, 3:] = R
        return A

    @staticmethod
    def adj(a):
        A = torch.zeros(a.shape[:-1] + (6, 6), dtype=a.dtype)
        A[..., :3, :3] = hat(a[..., 3:6])
        A[..., :3, 3:] = hat(a[..., :3])
        A[..., 3:, 3:] = hat(a[..., 3:6])
        return A

    @staticmethod
    def act(X, p):
        t, q = SE3.parse(X)
        return qrot(q, p) + t

    @staticmethod
    def act4(X, p):
        t, q = SE3.parse(X)
        return torch.cat([qrot(q, p[..., :3]) + t * p[..., 3:], p[..., 3:]], -1)

    @staticmethod
    def matrix4(X):
        t, q = SE3.parse(X)
        T = _eye(4, t)
        T[..., :3, :3] = qmat(q)
        T[..., :3, 3] = t
        return T

    @staticmethod
    def projector(X):
        t, q = SE3.parse(X)
        J = torch.zeros(t.shape[:-1] + (7, 7), dtype=t.dtype)
        J[..., :3, :3] = _eye(3, t)
        J[..., :3, 3:6] = hat(-t)
        J[..., 3:7, 3:7] = so3_projector(q)
        return J

    @staticmethod
    def calcQ(a):
        """SE3::calcQ (se3.h:135-160)."""
        tau, phi = a[..., :3], a[..., 3:6]
        Tau, Phi = hat(tau), hat(phi)
        theta = phi.norm(dim=-1)[..., None, None]
        small = theta < EPS
        ts = torch.where(small, torch.ones_like(theta), theta)
        t2 = ts * ts
        t4 = t2 * t2
        th2 = theta * theta
        c1 = _where(small, 1.0 / 6.0 - th2 / 120.0, (ts - torch.sin(ts)) / (t2 * ts))
        c2 = _where(small, 1.0 / 24.0 - th2 / 720.0, (t2 + 2 * torch.cos(ts) - 2) / (2 * t4))
        c3 = _where(small, 1.0 / 120.0 - th2 / 2520.0, (2 * ts - 3 * torch.sin(ts) + ts * torch.cos(ts)) / (2 * t4 * ts))
        PT = Phi @ Tau
        TP = Tau @ Phi
        PTP = PT @ Phi
        PP = Phi @ Phi
        return (0.5 * Tau + c1 * (PT + TP + PTP) + c2 * (PP @ Tau + TP @ Phi - 3 * PTP)
                + c3 * (PTP @ Phi + PP @ Tau @ Phi))

    @staticmethod
    def left_jac(a):
        phi = a[..., 3:6]
        J = so3_left_jac(phi)
        Q = SE3.calcQ(a)
        out = torch.zeros(a.shape[:-1] + (6, 6), dtype=a.dtype)
        out[..., :3, :3] = J
        out[..., :3, 3:] = Q
        out[..., 3:, 3:] = J
        return out

    @staticmethod
    def left_jac_inv(a):
        phi = a[..., 3:6]
        Ji = so3_left_jac_inv(phi)
        Q = SE3.calcQ(a)
        out = torch.zeros(a.shape[:-1] + (6, 6), dtype=a.dtype)
        out[..., :3, :3] = Ji
        out[..., :3, 3:] = -Ji @ Q @ Ji
        out[..., 3:, 3:] = Ji
        return out

    @staticmethod
    def act_jac(q):
        return torch.cat([_eye(3, q), hat(-q)], -1)

    @staticmethod
    def act4_jac(q4):
        J = torch.zeros(q4.shape[:-1] + (4, 6), dtype=q4.dtype)
        J[..., :3, :3] = q4[..., 3, None, None] * _eye(3, q4[..., :3])
        J[..., :3, 3:] = hat(-q4[..., :3])
        return J


class Sim3:
    gid, N, K = 4, 8, 7

    @staticmethod
    def parse(X):
        return X[..., :3], qnormalize(X[..., 3:7]), X[..., 7]

    @staticmethod
    def pack(t, q, s):
        return torch.cat([t, q, s[..., None]], -1)

    @staticmethod
    def exp(a):
        tau, phi, sigma = a[..., :3], a[..., 3:6], a[..., 6]
        q = so3_exp_quat(phi)
        W = rxso3_calcW(phi, sigma)
        return Sim3.pack((W @ tau[..., None])[..., 0], q, torch.exp(sigma))

    @staticmethod
    def log(X):
        t, q, s = Sim3.parse(X)
        phi = so3_log_quat(q)
        sigma = torch.log(s)
        W = rxso3_calcW(phi, sigma)
        tau = torch.linalg.solve(W, t[..., None])[..., 0]
        return torch.cat([tau, phi, sigma[..., None]], -1)

    @staticmethod
    def inv(X):
        t, q, s = Sim3.parse(X)
        qi, si = qconj(q), 1.0 / s
        return Sim3.pack(-(si[..., None] * qrot(qi, t)), qi, si)

    @staticmethod
    def mul(X, Y):
        ta, qa, sa = Sim3.parse(X)
        tb, qb, sb = Sim3.parse(Y)
        return Sim3.pack(ta + sa[..., None] * qrot(qa, tb), qnormalize(qmul(qa, qb)), sa * sb)

    @staticmethod
    def Adj(X):
        t, q, s = Sim3.parse(X)
        R = qmat(q)
        A = _eye(7, t[..., :1].expand(t.shape[:-1] + (7,)))
        A[..., :3, :3] = s[..., None, None] * R
        A[..., :3, 3:6] = hat(t) @ R
        A[..., :3, 6] = -t
        A[..., 3:6, 3:6] = R
        return A

    @staticmethod
    def adj(a):
        tau, phi, sigma = a[..., :3], a[..., 3:6], a[..., 6]
        A = torch.zeros(a.shape[:-1] + (7, 7), dtype=a.dtype)
        A[..., :3, :3] = hat(phi) + sigma[..., None, None] * _eye(3, phi)
        A[..., :3, 3:6] = hat(tau)
        A[..., :3, 6] = -tau
        A[..., 3:6, 3:6] = hat(phi)
        return A

    @staticmethod
    def act(X, p):
        t, q, s = Sim3.parse(X)
        return s[..., None] * qrot(q, p) + t

    @staticmethod
    def act4(X, p):
        t, q, s = Sim3.parse(X)
        return torch.cat([s[..., None] * qrot(q, p[..., :3]) + p[..., 3:] * t, p[..., 3:]], -1)

    @staticmethod
    def matrix4(X):
        t, q, s = Sim3.parse(X)
        T = _eye(4, t)
        T[..., :3, :3] = s[..., None, None] * qmat(q)
        T[..., :3, 3] = t
        return T

    @staticmethod
    def projector(X):
        t, q, s = Sim3.parse(X)
        J = torch.zeros(t.shape[:-1] + (8, 8), dtype=t.dtype)
        J[..., :3, :3] = _eye(3, t)
        J[..., :3, 3:6] = hat(-t)
        J[..., :3, 6] = t
        J[..., 3:7, 3:7] = so3_projector(q)
        J[..., 7, 6] = s
        return J

    @staticmethod
    def left_jac(a):
        """Truncated series; the Xi^5/720 term is dropped (sim3.h:160-169)."""
        Xi = Sim3.adj(a)
        Xi2 = Xi @ Xi
        Xi4 = Xi2 @ Xi2
        return _eye(7, a) + Xi / 2.0 + Xi2 / 6.0 + (Xi @ Xi2) / 24.0 + Xi4 / 120.0

    @staticmethod
    def left_jac_inv(a):
        """Truncated series (sim3.h:171-178)."""
        Xi = Sim3.adj(a)
        Xi2 = Xi @ Xi
        Xi4 = Xi2 @ Xi2
        return _eye(7, a) - Xi / 2.0 + Xi2 / 12.0 - Xi4 / 720.0

    @staticmethod
    def act_jac(q):
        return torch.cat([_eye(3, q), hat(-q), q[..., :, None]], -1)

    @staticmethod
    def act4_jac(q4):
        J = torch.zeros(q4.shape[:-1] + (4, 7), dtype=q4.dtype)
        J[..., :3, :3] = q4[..., 3, None, None] * _eye(3, q4[..., :3])
        J[..., :3, 3:6] = hat(-q4[..., :3])
        J[..., :3, 6] = q4[..., :3]
        return J


GROUPS = {1: SO3, 2: RxSO3, 3: SE3, 4: Sim3}


def group(gid: int):
    return GROUPS[gid]


# ----------------------------------------------------------------------------
# The 13 ops (lietorch.cpp:23-303 -> lietorch_gpu.cu:24-299), batched [n, .]
# ----------------------------------------------------------------------------

def _pad(g, K_vec):
    out = torch.zeros(K_vec.shape[:-1] + (g.N,), dtype=K_vec.dtype)
    out[..., : g.K] = K_vec
    return out


def _rowvec_mat(v, M):
    return (v[..., None, :] @ M)[..., 0, :]


def expm(gid, a):
    return group(gid).exp(a)


def expm_backward(gid, grad, a):
    g = group(gid)
    return [_rowvec_mat(grad[..., : g.K], g.left_jac(a))]


def logm(gid, X):
    return group(gid).log(X)


def logm_backward(gid, grad, X):
    g = group(gid)
    return [_pad(g, _rowvec_mat(grad, g.left_jac_inv(g.log(X))))]


def inv(gid, X):
    return group(gid).inv(X)


def inv_backward(gid, grad, X):
    g = group(gid)
    return [_pad(g, -_rowvec_mat(grad[..., : g.K], g.Adj(g.inv(X))))]


def mul(gid, X, Y):
    return group(gid).mul(X, Y)


def mul_backward(gid, grad, X, Y):
    g = group(gid)
    dZ = grad[..., : g.K]
    return [_pad(g, dZ), _pad(g, _rowvec_mat(dZ, g.Adj(X)))]


def adj(gid, X, a):
    g = group(gid)
    return (g.Adj(X) @ a[..., None])[..., 0]


def adj_backward(gid, grad, X, a):
    g = group(gid)
    A = g.Adj(X)
    b = (A @ a[..., None])[..., 0]
    return [_pad(g, -_rowvec_mat(grad, g.adj(b))), _rowvec_mat(grad, A)]


def adjT(gid, X, a):
    g = group(gid)
    return (g.Adj(X).transpose(-1, -2) @ a[..., None])[..., 0]


def adjT_backward(gid, grad, X, a):
    g = group(gid)
    A = g.Adj(X)
    Adb = (A @ grad[..., None])[..., 0]
    return [_pad(g, -_rowvec_mat(a, g.adj(Adb))), Adb]


def act(gid, X, p):
    return group(gid).act(X, p)


def act_backward(gid, grad, X, p):
    g = group(gid)
    dp = _rowvec_mat(grad, g.matrix4(X)[..., :3, :3])
    dX = _rowvec_mat(grad, g.act_jac(g.act(X, p)))
    return [_pad(g, dX), dp]


def act4(gid, X, p):
    return group(gid).act4(X, p)


def act4_backward(gid, grad, X, p):
    g = group(gid)
    dp = _rowvec_mat(grad, g.matrix4(X))
    dX = _rowvec_mat(grad, g.act4_jac(g.act4(X, p)))
    return [_pad(g, dX), dp]


def as_matrix(gid, X):
    return group(gid).matrix4(X)


def projector(gid, X):
    return group(gid).projector(X)


def Jinv(gid, X, a):
    g = group(gid)
    return (g.left_jac_inv(g.log(X)) @ a[..., None])[..., 0]


OPS = {
    "expm": expm, "expm_backward": expm_backward,
    "logm": logm, "logm_backward": logm_backward,
    "inv": inv, "inv_backward": inv_backward,
    "mul": mul, "mul_backward": mul_backward,
    "adj": adj, "adj_backward": adj_backward,
    "adjT": adjT, "adjT_backward": adjT_backward,
    "act": act, "act_backward": act_backward,
    "act4": act4, "act4_backward": act4_backward,
    "as_matrix": as_matrix, "projector": projector, "Jinv": Jinv,
}


# ----------------------------------------------------------------------------
# DROID BA retraction (geom_kernels.cu:116-177, :974-991)
# ----------------------------------------------------------------------------

def droid_exp_se3(xi: torch.Tensor):
    """expSE3 of the BA kernels: series for theta^2 < 1e-8, translation
    correction only when theta > 1e-4, quaternion left unnormalised
    (geom_kernels.cu:116-177).  Returns (t, q)."""
    tau, phi = xi[..., :3], xi[..., 3:6]
    theta2 = (phi * phi).sum(-1, keepdim=True)
    theta4 = theta2 * theta2
    theta = theta2.sqrt()
    small = theta2 < 1e-8
    ts = torch.where(small, torch.ones_like(theta), theta)
    imag = torch.where(small, 0.5 - theta2 / 48.0 + theta4 / 3840.0, torch.sin(0.5 * ts) / ts)
    real = torch.where(small, 1.0 - theta2 / 8.0 + theta4 / 384.0, torch.cos(0.5 * ts))
    q = torch.cat([imag * phi, real], -1)
    big = theta > 1e-4
    tb = torch.where(big, theta, torch.ones_like(theta))
    a = (1 - torch.cos(tb)) / (tb * tb)
    b = (tb - torch.sin(tb)) / (tb * tb * tb)
    c1 = torch.cross(phi, tau, dim=-1)
    c2 = torch.cross(phi, c1, dim=-1)
    t = tau + torch.where(big, a * c1 + b * c2, torch.zeros_like(tau))
    return t, q


def droid_retr_se3(xi: torch.Tensor, pose: torch.Tensor) -> torch.Tensor:
    """retrSE3: T <- Exp(xi) * T, quaternion not renormalised (geom_kernels.cu:974-991)."""
    dt, dq = droid_exp_se3(xi)
    t, q = pose[..., :3], pose[..., 3:7]
    q1 = qmul(dq, q)
    t1 = qrot(dq, t) + dt
    return torch.cat([t1, q1], -1)


def rand_group(gid, n, gen, sigma=1.0, dtype=F64):
    """Random element as exp(sigma * xi) (groups.py:114-127)."""
    g = group(gid)
    xi = torch.randn(n, g.K, generator=gen, dtype=dtype) * sigma
    return g.exp(xi)


__all__ = [n for n in dir() if not n.startswith("_")] + ["math"]
