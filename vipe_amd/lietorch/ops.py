"""Batching and autograd around the ``lietorch_ext`` kernels.

Every group method reduces to one kernel call on flat ``[n, d]`` operands:
``run(op, group_id, x[, y])`` broadcasts the batch axes of its operands
(``torch.broadcast_shapes``), flattens them, calls the kernel through
``LieKernel`` and restores the batch shape.  ``LieKernel``'s backward is the
kernel's hand-written backward (``<op>_backward``), which returns gradients
with respect to group elements as tangent K-vectors in the first K slots of an
``[n, N]`` buffer (lietorch_gpu.cu:95-98).  ``Embedding`` converts between that
tangent form and the plain N-vector of the embedding, which is what
``vec()`` / ``InitFromVec`` expose (vipe/ext/lietorch/group_ops.py:88-123).
"""
from __future__ import annotations

import torch

from ..ext import lietorch_ext as _ext

# op -> (forward kernel, backward kernel); Jinv has no backward in the reference either.
_KERNELS = {
    "exp": (_ext.expm, _ext.expm_backward),
    "log": (_ext.logm, _ext.logm_backward),
    "inv": (_ext.inv, _ext.inv_backward),
    "mul": (_ext.mul, _ext.mul_backward),
    "adj": (_ext.adj, _ext.adj_backward),
    "adjT": (_ext.adjT, _ext.adjT_backward),
    "act": (_ext.act, _ext.act_backward),
    "act4": (_ext.act4, _ext.act4_backward),
    "Jinv": (_ext.Jinv, None),
}


class LieKernel(torch.autograd.Function):
    """One lietorch_ext op on flat contiguous operands, differentiable through its backward kernel."""

    @staticmethod
    def forward(ctx, op, group_id, *operands):
        ctx.op, ctx.group_id = op, group_id
        ctx.save_for_backward(*operands)
        return _KERNELS[op][0](group_id, *operands)

    @staticmethod
    def backward(ctx, grad_out):
        backward_kernel = _KERNELS[ctx.op][1]
        if backward_kernel is None:
            raise RuntimeError(f"lietorch op '{ctx.op}' has no backward")
        grads = backward_kernel(ctx.group_id, grad_out.contiguous(), *ctx.saved_tensors)
        return (None, None, *grads)


class Embedding(torch.autograd.Function):
    """Identity on the data.  Backward maps an upstream gradient between the
    embedding (N numbers per element) and the tangent representation with the
    group's projector P [N, N] (lietorch.cpp:263): to the embedding, g -> g P;
    from the embedding, g -> g pinv(P)."""

    @staticmethod
    def forward(ctx, to_embedding, group_id, data):
        ctx.to_embedding, ctx.group_id = to_embedding, group_id
        ctx.save_for_backward(data)
        return data.view_as(data)

    @staticmethod
    def backward(ctx, grad_out):
        (data,) = ctx.saved_tensors
        P = _ext.projector(ctx.group_id, data)
        if not ctx.to_embedding:
            P = torch.linalg.pinv(P)
        return None, None, (grad_out.unsqueeze(-2) @ P).squeeze(-2)


def _flatten(operands):
    batch = torch.broadcast_shapes(*(t.shape[:-1] for t in operands))
    flat = [t.expand(batch + t.shape[-1:]).reshape(-1, t.shape[-1]).contiguous() for t in operands]
    return flat, batch


def run(op, group_id, *operands):
    """Apply ``op`` element-wise over the broadcast batch of ``operands``; returns ``batch + (d,)``."""
    flat, batch = _flatten(operands)
    return LieKernel.apply(op, group_id, *flat).view(batch + (-1,))


def embedding(group_id, data, to_embedding):
    flat, batch = _flatten([data])
    return Embedding.apply(to_embedding, group_id, flat[0]).view(batch + (-1,))


def as_matrix(group_id, data):
    """4x4 homogeneous matrices of ``data`` (no autograd: the kernel has no backward)."""
    flat, batch = _flatten([data])
    return _ext.as_matrix(group_id, flat[0]).view(batch + (4, 4))
