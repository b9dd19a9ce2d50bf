"""SO3 / RxSO3 / SE3 / Sim3 elements over the gfx950 lietorch kernels.

The interface is that of ``vipe.ext.lietorch`` (vipe/ext/lietorch/groups.py:14-338):
- class names and group ids: SO3 = 1, RxSO3 = 2, SE3 = 3, Sim3 = 4;
- embeddings: SO3 [qx,qy,qz,qw], RxSO3 [q, s], SE3 [tx,ty,tz, q], Sim3 [t, q, s];
- tangent sizes 3, 4, 6, 7 in the order (translation, rotation, log-scale);
- the element methods (exp, log, inv, mul, retr, adj, adjT, Jinv, act,
  matrix, translation, vec, ...), ``LieGroupParameter``, ``cat`` and ``stack``.

An element holds ``data`` of shape ``batch + (N,)``.  Each method is one
kernel call from ``ops.run`` over the (broadcast) batch, on the tensors'
device: gfx950 kernels for device tensors, the library's host loop for CPU
tensors.  Gradients flow through the hand-written backward kernels.
"""
from __future__ import annotations

import torch

from . import ops


def _batch_shape(args):
    """``f(2, 3)``, ``f((2, 3))`` and ``f([2, 3])`` all name the batch shape (2, 3)."""
    if len(args) == 1 and isinstance(args[0], (tuple, list, torch.Size)):
        args = tuple(args[0])
    return tuple(int(n) for n in args)


class LieGroup:
    """Base class: a batch of group elements stored in ``data``."""

    group_name = "LieGroup"
    group_id = 0
    manifold_dim = 0  # K, tangent size
    embedded_dim = 0  # N, embedding size
    _identity: tuple = ()

    def __init__(self, data):
        self.data = self._from(data)

    @classmethod
    def _from(cls, data):
        if not isinstance(data, LieGroup):
            return data
        if type(data) is cls:
            return data.data
        convert = _CONVERSIONS.get((type(data), cls))
        if convert is None:
            raise TypeError(f"cannot make a {cls.group_name} from a {data.group_name}")
        return convert(data.data)

    # ---- description --------------------------------------------------------
    def __repr__(self):
        return f"{self.group_name}: size={self.shape}, device={self.device}, dtype={self.dtype}"

    @property
    def shape(self):
        return self.data.shape[:-1]

    @property
    def device(self):
        return self.data.device

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def tangent_shape(self):
        return self.shape + (self.manifold_dim,)

    # ---- constructors ---------------------------------------------------------
    @classmethod
    def _identity_tensor(cls, dtype=None, device=None):
        return torch.tensor(cls._identity, dtype=dtype or torch.float32, device=device)

    @classmethod
    def Identity(cls, *batch_shape, **kwargs):
        batch = _batch_shape(batch_shape)
        ident = cls._identity_tensor(kwargs.get("dtype"), kwargs.get("device"))
        return cls(ident.expand(batch + (cls.embedded_dim,)).contiguous())

    @classmethod
    def IdentityLike(cls, G):
        return cls.Identity(G.shape, device=G.device, dtype=G.dtype)

    @classmethod
    def Random(cls, *batch_shape, sigma=1.0, **kwargs):
        """Exp of a tangent vector drawn from N(0, sigma^2) per component."""
        xi = torch.randn(_batch_shape(batch_shape) + (cls.manifold_dim,), **kwargs)
        return cls.exp(sigma * xi)

    @classmethod
    def InitFromVec(cls, data):
        return cls(ops.embedding(cls.group_id, data, to_embedding=False))

    @classmethod
    def apply_op(cls, op, x, y=None):
        """Element-wise kernel ``op`` ('exp', 'log', 'mul', 'act4', ...) over broadcast operands."""
        return ops.run(op, cls.group_id, *((x,) if y is None else (x, y)))

    @classmethod
    def exp(cls, x):
        return cls(ops.run("exp", cls.group_id, x))

    # ---- group operations -----------------------------------------------------
    def vec(self):
        return ops.embedding(self.group_id, self.data, to_embedding=True)

    def quaternion(self):
        start = 0 if self.group_id in (1, 2) else 3
        return self.data[..., start:start + 4]

    def log(self):
        return ops.run("log", self.group_id, self.data)

    def inv(self):
        return type(self)(ops.run("inv", self.group_id, self.data))

    def mul(self, other):
        return type(self)(ops.run("mul", self.group_id, self.data, other.data))

    def retr(self, a):
        """Exp(a) * X."""
        step = ops.run("exp", self.group_id, a)
        return type(self)(ops.run("mul", self.group_id, step, self.data))

    def adj(self, a):
        """Adj(X) a."""
        return ops.run("adj", self.group_id, self.data, a)

    def adjT(self, a):
        """a Adj(X), i.e. Adj(X)^T a."""
        return ops.run("adjT", self.group_id, self.data, a)

    def Jinv(self, a):
        return ops.run("Jinv", self.group_id, self.data, a)

    def act(self, p):
        """Action on points: [..., 3] Euclidean or [..., 4] homogeneous."""
        width = p.shape[-1]
        if width not in (3, 4):
            raise ValueError(f"points must have 3 or 4 coordinates, got {width}")
        return ops.run("act" if width == 3 else "act4", self.group_id, self.data, p)

    def matrix(self):
        """[..., 4, 4] homogeneous matrices.  Differentiable: with autograd
        recording, column k is the action on the basis point e_k."""
        if not (torch.is_grad_enabled() and self.data.requires_grad):
            return ops.as_matrix(self.group_id, self.data)
        basis = torch.eye(4, dtype=self.dtype, device=self.device)
        return ops.run("act4", self.group_id, self.data.unsqueeze(-2), basis).transpose(-1, -2)

    def translation(self):
        """Image of the homogeneous origin [0, 0, 0, 1]: [t, 1] (SE3, Sim3) or [0, 0, 0, 1]."""
        origin = torch.zeros(4, dtype=self.dtype, device=self.device)
        origin[3] = 1.0
        return ops.run("act4", self.group_id, self.data, origin)

    def __mul__(self, other):
        if isinstance(other, LieGroup):
            return self.mul(other)
        if isinstance(other, torch.Tensor):
            return self.act(other)
        return NotImplemented

    # ---- tensor plumbing ------------------------------------------------------
    def detach(self):
        return type(self)(self.data.detach())

    def view(self, dims):
        return type(self)(self.data.view(_batch_shape((dims,)) + (self.embedded_dim,)))

    def __getitem__(self, index):
        return type(self)(self.data[index])

    def __setitem__(self, index, item):
        self.data[index] = item.data

    def to(self, *args, **kwargs):
        return type(self)(self.data.to(*args, **kwargs))

    def cpu(self):
        return self.to("cpu")

    def cuda(self):
        return type(self)(self.data.cuda())

    def float(self, device=None):
        return type(self)(self.data.float())

    def double(self, device=None):
        return type(self)(self.data.double())

    def unbind(self, dim=0):
        return [type(self)(d) for d in self.data.unbind(dim=dim)]


class SO3(LieGroup):
    group_name, group_id, manifold_dim, embedded_dim = "SO3", 1, 3, 4
    _identity = (0.0, 0.0, 0.0, 1.0)


class RxSO3(LieGroup):
    group_name, group_id, manifold_dim, embedded_dim = "RxSO3", 2, 4, 5
    _identity = (0.0, 0.0, 0.0, 1.0, 1.0)


class SE3(LieGroup):
    group_name, group_id, manifold_dim, embedded_dim = "SE3", 3, 6, 7
    _identity = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)

    def scale(self, s):
        """Translation multiplied by s (batch-shaped), rotation unchanged."""
        scaled = self.data.clone()
        scaled[..., :3] = self.data[..., :3] * s.unsqueeze(-1)
        return SE3(scaled)


class Sim3(LieGroup):
    group_name, group_id, manifold_dim, embedded_dim = "Sim3", 4, 7, 8
    _identity = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0)


for _cls in (SO3, RxSO3, SE3, Sim3):
    # the reference exposes the identity embedding as a class attribute tensor
    _cls.id_elem = _cls._identity_tensor()


def _pad(data, before, after):
    """[..., d] -> [..., before zeros, d, after ones]."""
    parts = []
    if before:
        parts.append(data.new_zeros(data.shape[:-1] + (before,)))
    parts.append(data)
    if after:
        parts.append(data.new_ones(data.shape[:-1] + (after,)))
    return torch.cat(parts, -1)


# (source class, target class) -> embedding conversion; the pairs the
# reference constructors accept (groups.py:244-318)
_CONVERSIONS = {
    (SE3, SO3): lambda d: d[..., 3:7],
    (Sim3, RxSO3): lambda d: d[..., 3:8],
    (SO3, SE3): lambda d: _pad(d, 3, 0),
    (SO3, Sim3): lambda d: _pad(d, 3, 1),
    (SE3, Sim3): lambda d: _pad(d, 0, 1),
}


class LieGroupParameter(torch.Tensor):
    """A tangent-space tensor (zeros, requires grad) bound to a group element:
    the element it stands for is ``group.retr(self)`` = Exp(self) * group."""

    from torch._C import _disabled_torch_function_impl

    __torch_function__ = _disabled_torch_function_impl

    def __new__(cls, group, requires_grad=True):
        zeros = torch.zeros(group.tangent_shape, device=group.device, dtype=group.dtype)
        return torch.Tensor._make_subclass(cls, zeros, requires_grad)

    def __init__(self, group, requires_grad=True):
        self.group = group

    def retr(self):
        return self.group.retr(self)

    def log(self):
        return self.retr().log()

    def inv(self):
        return self.retr().inv()

    def adj(self, a):
        return self.retr().adj(a)

    def __mul__(self, other):
        lhs = self.retr()
        return lhs * (other.retr() if isinstance(other, LieGroupParameter) else other)

    def add_(self, update, alpha):
        """Moves the bound element: group <- Exp(alpha * update) * group."""
        self.group = self.group.exp(alpha * update) * self.group

    def __getitem__(self, index):
        return self.retr()[index]


def cat(group_list, dim):
    """Concatenate elements of one group along batch axis ``dim``."""
    return type(group_list[0])(torch.cat([g.data for g in group_list], dim=dim))


def stack(group_list, dim):
    """Stack elements of one group along a new batch axis ``dim``."""
    return type(group_list[0])(torch.stack([g.data for g in group_list], dim=dim))


__all__ = ["LieGroup", "SO3", "RxSO3", "SE3", "Sim3", "LieGroupParameter", "cat", "stack"]
