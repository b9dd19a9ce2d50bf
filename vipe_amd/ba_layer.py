"""One dense BA Gauss-Newton iteration as a differentiable layer.

``BAIteration.apply(targets, weights, poses, disps, intrinsics, disps_sens,
eta, ii, jj, depth_active, t0, t1, lm, ep, motion_only, alpha)`` runs one
iteration of ``slam_ext.ba_extended`` (geom_kernels.cu:1626-1847 semantics)
on copies of ``poses`` / ``disps`` and returns ``(dx, dz)`` - the pose
update [P,6] and the inverse-depth update [|kx|, H*W] (``dz`` is None when
motion only).  Its backward is the implicit adjoint of that iteration: one
extra solve with the forward's Cholesky factor of the reduced camera system
and one pass over the residual rows (``vh_ba_adjoint``), giving
dL/d(targets) and dL/d(weights).  The updated poses and disps follow as
``Exp(dx) * poses`` and ``disps + dz`` (``retract``).

The reference has no backward for its BA layer; this module is an extension
of the drop-in, used by training-time callers that differentiate through BA.
"""
from __future__ import annotations

import torch

from .ext import slam_ext


class BAIteration(torch.autograd.Function):
    @staticmethod
    def forward(ctx, targets, weights, poses, disps, intrinsics, disps_sens, eta, ii, jj, depth_active, t0, t1,
                lm=1e-4, ep=0.1, motion_only=False, alpha=0.05):
        targets = targets.detach().contiguous()
        weights = weights.detach().contiguous()
        p0 = poses.detach().contiguous().clone()
        d0 = disps.detach().contiguous().clone()
        p1, d1 = p0.clone(), d0.clone()
        intr = intrinsics.detach().contiguous().clone()
        dx, dz, _, plan, ws = slam_ext._ba_extended(p1, d1, intr.clone(), disps_sens, targets, weights, eta, ii, jj,
                                                    depth_active, t0, t1, 1, lm, ep, motion_only, alpha, False,
                                                    1e-6, 1e-6, 1.0, False)
        ctx.plan, ctx.ws = plan, ws
        ctx.lm, ctx.motion_only, ctx.P = float(lm), bool(motion_only), int(t1) - int(t0)
        ctx.save_for_backward(targets, weights, p0, d0, intr, depth_active)
        return dx, dz

    @staticmethod
    def backward(ctx, grad_dx, grad_dz):
        targets, weights, p0, d0, intr, depth_active = ctx.saved_tensors
        if grad_dx is None:
            grad_dx = torch.zeros((ctx.P, 6), dtype=torch.float32, device=p0.device)
        if grad_dz is None and not ctx.motion_only:
            grad_dz = torch.zeros((ctx.plan.K, d0.shape[1] * d0.shape[2]), dtype=torch.float32, device=p0.device)
        gt, gw = slam_ext.ba_adjoint(ctx.plan, ctx.ws, p0, d0, intr, targets, weights, depth_active, ctx.lm,
                                     ctx.motion_only, grad_dx, grad_dz)
        return (gt, gw) + (None,) * 14


def ba_iteration(targets, weights, poses, disps, intrinsics, disps_sens, eta, ii, jj, depth_active, t0, t1,
                 lm=1e-4, ep=0.1, motion_only=False, alpha=0.05):
    """(dx, dz) of one BA iteration, differentiable w.r.t. targets and weights."""
    return BAIteration.apply(targets, weights, poses, disps, intrinsics, disps_sens, eta, ii, jj, depth_active,
                             t0, t1, lm, ep, motion_only, alpha)
