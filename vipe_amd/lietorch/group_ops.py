"""Autograd wrappers of the lietorch_ext ops (vipe/ext/lietorch/group_ops.py:10-123).

Each op's forward calls the gfx950 (or host) kernel and its backward calls the
hand-written backward kernel of the same op; gradients with respect to group
elements are tangent vectors stored in the first K entries of an [n, N]
buffer, as in the reference.
"""
from __future__ import annotations

import torch

from ..ext import lietorch_ext as _ops


class GroupOp(torch.autograd.Function):
    """Group operation base class: ``forward_op`` / ``backward_op`` are ext ops."""

    forward_op = None
    backward_op = None

    @classmethod
    def forward(cls, ctx, group_id, *inputs):
        ctx.group_id = group_id
        ctx.save_for_backward(*inputs)
        return cls.forward_op(ctx.group_id, *inputs)

    @classmethod
    def backward(cls, ctx, grad):
        if cls.backward_op is None:
            raise RuntimeError(f"Backward operation not implemented for {cls}")
        inputs = ctx.saved_tensors
        grads = cls.backward_op(ctx.group_id, grad.contiguous(), *inputs)
        return (None,) + tuple(grads)


def _op(name, fwd, bwd):
    return type(name, (GroupOp,), {"forward_op": staticmethod(fwd), "backward_op": staticmethod(bwd) if bwd else None})


Exp = _op("Exp", _ops.expm, _ops.expm_backward)
Log = _op("Log", _ops.logm, _ops.logm_backward)
Inv = _op("Inv", _ops.inv, _ops.inv_backward)
Mul = _op("Mul", _ops.mul, _ops.mul_backward)
Adj = _op("Adj", _ops.adj, _ops.adj_backward)
AdjT = _op("AdjT", _ops.adjT, _ops.adjT_backward)
Act3 = _op("Act3", _ops.act, _ops.act_backward)
Act4 = _op("Act4", _ops.act4, _ops.act4_backward)
Jinv = _op("Jinv", _ops.Jinv, None)
ToMatrix = _op("ToMatrix", _ops.as_matrix, None)


class FromVec(torch.autograd.Function):
    """Vector -> group element; backward maps through pinv of the projector."""

    @staticmethod
    def forward(ctx, group_id, *inputs):
        ctx.group_id = group_id
        ctx.save_for_backward(*inputs)
        return inputs[0]

    @staticmethod
    def backward(ctx, grad):
        inputs = ctx.saved_tensors
        J = _ops.projector(ctx.group_id, *inputs)
        return None, torch.matmul(grad.unsqueeze(-2), torch.linalg.pinv(J)).squeeze(-2)


class ToVec(torch.autograd.Function):
    """Group element -> vector; backward multiplies by the projector."""

    @staticmethod
    def forward(ctx, group_id, *inputs):
        ctx.group_id = group_id
        ctx.save_for_backward(*inputs)
        return inputs[0]

    @staticmethod
    def backward(ctx, grad):
        inputs = ctx.saved_tensors
        J = _ops.projector(ctx.group_id, *inputs)
        return None, torch.matmul(grad.unsqueeze(-2), J).squeeze(-2)
