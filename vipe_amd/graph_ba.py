"""Caller side of the fused BA: GraphBuffer._fused_ba / bundle_adjustment.

Restates the argument preparation of vipe/slam/components/buffer.py:364-489
(and the fused branch of :501-566) on top of ``slam_ext.ba_extended``:

* rejects layouts the fused kernels do not cover with the reference's
  messages (multi-view or non-pinhole buffers, rig rotation, empty window,
  mismatched edge counts, non-default flow weights, robust kernels);
* with ``ba_config.fused_robust`` set, runs the configured robust kernel
  (``robust_kernel``, ``robust_kernel_threshold``, ``gnc_*``) on the fused
  path instead of rejecting it: Huber / Tukey weights inside the frame kernel
  for ``n_iters`` iterations, GNC-TLS as ``gnc_n_mu_steps`` x
  ``gnc_gn_iters_per_mu`` iterations over the mu schedule of buffer.py:686-695;
* drops self edges ii == jj; merges sparse-track targets and weights into the
  dense ones (weighted mean where the combined weight is positive);
* ``[E, H*W, 2] -> [E, 2, H, W]``; eta = 0.2 * damping[kx] + 1e-7 with
  kx = unique(cat(arange(t0,t1), ii)); pinhole intrinsics at 1/8 resolution
  with intrinsics_scale = 1/8; depth_active = 1, or the [t0, t1) mask for
  limited_disp;
* writes optimised intrinsics back at full resolution and clamps disparities
  at 0.001 afterwards;
* with ``ba_config.fused_rig`` set, runs multi-view rig buffers, non-identity
  rigs and the Mei camera on the fused path (``slam_ext.ba_rig``) instead of
  rejecting them: the edges expand per view as GraphBuffer.expand_edge_multiview
  does (buffer.py:304-362; ii == jj becomes the frame's cross-view edges), the
  disparities are the flattened [N V, h, w] buffer, intrinsics per view at
  1/8 resolution (the Mei k1 unscaled, cameras.py:345-348), fixed.

``buffer`` is any object with the GraphBuffer attributes used here:
``poses [B,7]``, ``disps [B,V,h,w]``, ``disps_sens [B,V,h,w]``,
``intrinsics [V,>=4]``, ``tstamp [B]``, ``rig [V,7]``, ``n_views``,
``camera_type`` (``"pinhole"``), ``height``, ``width``, ``device``,
``ba_config`` (mapping with ``dense_disp_alpha``, optional
``intrinsics_damping_scale``, ``robust_kernel``) and ``sparse_tracks``
(``None`` or an object with ``enabled`` and
``compute_dense_disp_target_weight``).
"""
from __future__ import annotations

import logging

import torch
from einops import rearrange

from .ba_kernel import build_robust_kernel
from .ext import slam_ext

logger = logging.getLogger(__name__)


def _cfg(cfg, key, default=None):
    if cfg is None:
        return default
    if isinstance(cfg, dict):
        return cfg.get(key, default)
    getter = getattr(cfg, "get", None)
    if getter is not None:
        return getter(key, default)
    return getattr(cfg, key, default)


def _fail(reason: str):
    raise RuntimeError(f"Fused BA is enabled, but {reason}. Set ba.fused=false to use the generic BA solver.")


def _camera_name(buffer):
    return str(getattr(buffer.camera_type, "value", getattr(buffer.camera_type, "name", buffer.camera_type))).lower()


def expand_edge_multiview(buffer, ii, jj):
    """GraphBuffer.expand_edge_multiview (buffer.py:304-362, cross=True): every
    edge once per view q (q -> q); an edge ii == jj becomes the frame's
    cross-view edges q -> cross_view_idx[q] (by default q + 1 mod V).  Returns
    (pi, qi, di, pj, qj, dj) with di = pi V + qi, dj = pj V + qj."""
    V = int(buffer.n_views)
    dev = ii.device
    M = ii.shape[0]
    qi = torch.arange(V, device=dev).repeat(M)
    pi = ii.reshape(-1, 1).repeat(1, V).reshape(-1)
    qj = qi.clone()
    pj = jj.reshape(-1, 1).repeat(1, V).reshape(-1)
    cross = pi == pj
    if bool(torch.any(cross)):
        cvi = getattr(buffer, "cross_view_idx", None)
        if cvi is None:
            pj[cross], qj[cross] = pi[cross], (qi[cross] + 1) % V
        else:
            t, v = cvi[pi[cross], qi[cross]].unbind(-1)
            pj[cross], qj[cross] = t.to(pj.dtype), v.to(qj.dtype)
    return pi, qi, pi * V + qi, pj, qj, pj * V + qj


def fused_ba_rig(buffer, target, weight, disp_damping, ii, jj, t0, t1, n_iters, pose_damping, pose_ep, motion_only,
                 limited_disp, optimize_intrinsics, optimize_rig_rotation, verbose=False, robust=None):
    """The fused step on a rig buffer or a Mei camera (``slam_ext.ba_rig``).
    ``target`` / ``weight`` are [M, h*w, 2] for the expanded edges of
    expand_edge_multiview (the generic branch's layout, buffer.py:569-586);
    ``disp_damping`` is [N*V, h, w]."""
    camera = _camera_name(buffer)
    if camera not in ("pinhole", "mei"):
        _fail(f"the {camera} camera has no projection model")
    if optimize_intrinsics:
        _fail("it does not optimize intrinsics on rig or Mei buffers")
    if optimize_rig_rotation:
        _fail("it does not support optimizing rig rotation")
    if t0 >= t1:
        _fail("the active optimization range is empty")
    pi, qi, di, pj, qj, _ = expand_edge_multiview(buffer, ii, jj)
    if target.shape[0] != pi.shape[0] or weight.shape[0] != pi.shape[0]:
        _fail("target/weight edge counts do not match the expanded graph edges")
    keep = ~((pi == pj) & (qi == qj))  # a view of a frame onto itself constrains nothing but the energy
    if not bool(torch.all(keep)):
        pi, qi, pj, qj = pi[keep], qi[keep], pj[keep], qj[keep]
        target, weight = target[keep], weight[keep]
    if pi.numel() == 0:
        _fail("the graph contains no edges between distinct views")
    V = int(buffer.n_views)
    ht, wd = buffer.height // 8, buffer.width // 8
    tracks = getattr(buffer, "sparse_tracks", None)
    if tracks is not None and getattr(tracks, "enabled", False):
        s_target, s_weight = tracks.compute_dense_disp_target_weight(
            source_view_inds=qi, source_frame_inds=buffer.tstamp[pi], target_view_inds=qj,
            target_frame_inds=buffer.tstamp[pj], image_size=(buffer.height, buffer.width), dense_disp_size=(ht, wd))
        s_target, s_weight = s_target.flatten(1, 2), s_weight.flatten(1, 2)
        combined = weight + s_weight
        target = torch.where(combined > 0, (weight * target + s_weight * s_target) / combined.clamp_min(1e-12), target)
        weight = combined
    target = rearrange(target, "k (h w) c -> k c h w", h=ht, w=wd, c=2).contiguous()
    weight = rearrange(weight, "k (h w) c -> k c h w", h=ht, w=wd, c=2).contiguous()
    N = buffer.disps.shape[0]
    if not buffer.disps.is_contiguous():
        raise RuntimeError("disps must be contiguous")
    disps = buffer.disps.view(N * V, ht, wd)
    disps_sens = buffer.disps_sens.reshape(N * V, ht, wd).contiguous()
    scale = torch.ones(5 if camera == "mei" else 4, device=buffer.intrinsics.device)
    scale[:4] = 1.0 / 8.0
    intrinsics = (buffer.intrinsics[:, :scale.shape[0]] * scale).contiguous()
    rig = buffer.rig[:V].contiguous().float()
    plan = slam_ext.get_rig_plan(pi, pj, qi, qj, N, V, ht, wd, t0, t1)
    kx = plan.kx.to(disp_damping.device)
    eta = (0.2 * disp_damping.reshape(N * V, ht, wd)[kx] + 1e-7).contiguous()
    depth_active = torch.ones(N * V, dtype=disps.dtype, device=disps.device)
    if limited_disp:
        depth_active.zero_()
        depth_active[t0 * V:t1 * V] = 1.0
    _, _, energy = slam_ext.ba_rig(buffer.poses, disps, intrinsics, camera, rig, disps_sens, target, weight, eta,
                                   pi.contiguous(), pj.contiguous(), qi.contiguous(), qj.contiguous(), depth_active,
                                   t0, t1, n_iters, pose_damping, pose_ep, motion_only,
                                   float(_cfg(buffer.ba_config, "dense_disp_alpha")), verbose, robust)
    if verbose:
        logger.info(f"BA iters = {n_iters}, energy: {energy[0].item()} -> {energy[-1].item()}")
    return energy


def fused_ba(buffer, target, weight, disp_damping, ii, jj, t0, t1, n_iters, pose_damping, pose_ep, motion_only,
             limited_disp, optimize_intrinsics, optimize_rig_rotation, weight_dense_disp=0.001,
             weight_tracks=0.001, verbose=False, robust=None):
    """GraphBuffer._fused_ba (buffer.py:364-489).  Returns the per-iteration energy
    (empty unless ``verbose``).  ``robust``: (type, threshold, mu, mu_max) of a
    robust kernel on the flow residuals, or None."""
    camera = _camera_name(buffer)
    identity_rig = torch.as_tensor([0, 0, 0, 0, 0, 0, 1], dtype=buffer.rig.dtype, device=buffer.rig.device)
    rig_layout = buffer.n_views != 1 or camera != "pinhole" or not torch.allclose(buffer.rig[0], identity_rig)
    if rig_layout and bool(_cfg(buffer.ba_config, "fused_rig", False)):
        if weight_dense_disp != 0.001 or weight_tracks != 0.001:
            _fail("the configured dense-flow or sparse-track weights are unsupported")
        return fused_ba_rig(buffer, target, weight, disp_damping, ii, jj, t0, t1, n_iters, pose_damping, pose_ep,
                            motion_only, limited_disp, optimize_intrinsics, optimize_rig_rotation, verbose, robust)
    if buffer.n_views != 1 or camera != "pinhole":
        _fail("it only supports single-view pinhole-camera graph buffers")
    if optimize_rig_rotation:
        _fail("it does not support optimizing rig rotation")
    if t0 >= t1:
        _fail("the active optimization range is empty")
    if target.shape[0] != ii.shape[0] or weight.shape[0] != ii.shape[0]:
        _fail("target/weight edge counts do not match the graph edges")
    if weight_dense_disp != 0.001 or weight_tracks != 0.001:
        _fail("the configured dense-flow or sparse-track weights are unsupported")

    edge_mask = ii != jj
    if not torch.any(edge_mask):
        _fail("the graph contains no non-self edges")
    if not torch.all(edge_mask):
        ii, jj = ii[edge_mask], jj[edge_mask]
        target, weight = target[edge_mask], weight[edge_mask]

    if not torch.allclose(buffer.rig[0], identity_rig):
        _fail("it only supports identity rig extrinsics")

    ht, wd = buffer.height // 8, buffer.width // 8
    tracks = getattr(buffer, "sparse_tracks", None)
    if tracks is not None and getattr(tracks, "enabled", False):
        view_inds = torch.zeros_like(ii)
        s_target, s_weight = tracks.compute_dense_disp_target_weight(
            source_view_inds=view_inds, source_frame_inds=buffer.tstamp[ii], target_view_inds=view_inds,
            target_frame_inds=buffer.tstamp[jj], image_size=(buffer.height, buffer.width), dense_disp_size=(ht, wd))
        s_target = s_target.flatten(1, 2)
        s_weight = s_weight.flatten(1, 2)
        combined = weight + s_weight
        target = torch.where(combined > 0, (weight * target + s_weight * s_target) / combined.clamp_min(1e-12), target)
        weight = combined

    target = rearrange(target, "k (h w) c -> k c h w", h=ht, w=wd, c=2).contiguous()
    weight = rearrange(weight, "k (h w) c -> k c h w", h=ht, w=wd, c=2).contiguous()
    device = buffer.poses.device
    kx = torch.unique(torch.cat([torch.arange(t0, t1, device=device), ii]))
    eta = (0.2 * disp_damping[kx] + 1e-7).contiguous()
    intrinsics_scale = 1.0 / 8.0
    intrinsics = (buffer.intrinsics[0, :4] * intrinsics_scale).contiguous()
    depth_active = torch.ones(buffer.disps.shape[0], dtype=buffer.disps.dtype, device=device)
    if limited_disp:
        depth_active.zero_()
        depth_active[t0:t1] = 1.0
    damp_scale = float(_cfg(buffer.ba_config, "intrinsics_damping_scale", 1.0))
    disps = buffer.disps[:, 0]
    if not disps.is_contiguous():
        raise RuntimeError("disps must be contiguous")
    _, _, energy = slam_ext.ba_extended_robust(
        buffer.poses, disps, intrinsics, buffer.disps_sens[:, 0].contiguous(), target, weight, eta,
        ii.contiguous(), jj.contiguous(), depth_active.contiguous(), t0, t1, n_iters, pose_damping, pose_ep,
        motion_only, float(_cfg(buffer.ba_config, "dense_disp_alpha")), bool(optimize_intrinsics),
        1e-6 * damp_scale, 1e-6 * damp_scale, intrinsics_scale, verbose, robust)
    if verbose:
        logger.info(f"BA iters = {n_iters}, energy: {energy[0].item()} -> {energy[-1].item()}")
    if optimize_intrinsics:
        buffer.intrinsics[0, :4] = intrinsics / intrinsics_scale
    return energy


def bundle_adjustment(buffer, target, weight, disp_damping, ii, jj, t0, t1, n_iters, pose_damping, pose_ep,
                      motion_only, limited_disp, optimize_intrinsics, optimize_rig_rotation, verbose=False):
    """Fused branch of GraphBuffer.bundle_adjustment (buffer.py:501-566)."""
    if t0 > t1:
        raise ValueError("t0 must not exceed t1")
    cfg = buffer.ba_config
    kernel = build_robust_kernel(
        _cfg(cfg, "robust_kernel", None), float(_cfg(cfg, "robust_kernel_threshold", 1.0)),
        gnc_mu_init=float(_cfg(cfg, "gnc_mu_init", 1.0)), gnc_mu_step=float(_cfg(cfg, "gnc_mu_step", 1.4)),
        gnc_mu_max=float(_cfg(cfg, "gnc_mu_max", 1.0e6)))
    args = (buffer, target, weight, disp_damping, ii, jj, t0, t1)
    rest = (pose_damping, pose_ep, motion_only, limited_disp, optimize_intrinsics, optimize_rig_rotation, 0.001,
            0.001, verbose)
    if kernel is None:
        energy = fused_ba(*args, n_iters, *rest)
    elif not bool(_cfg(cfg, "fused_robust", False)):
        raise RuntimeError("Fused BA is enabled, but robust kernels are not supported. "
                           "Set ba.fused=false to use the generic BA solver.")
    elif kernel.is_gnc():
        # outer mu schedule, inner GN iterations at a frozen mu (buffer.py:686-695)
        n_mu = max(1, int(_cfg(cfg, "gnc_n_mu_steps", 4)))
        per_mu = max(1, int(_cfg(cfg, "gnc_gn_iters_per_mu", 6)))
        mu = kernel.mu_init
        parts = []
        for _ in range(n_mu):
            kernel.set_mu(mu)
            parts.append(fused_ba(*args, per_mu, *rest, robust=kernel.fused_args()))
            mu = kernel.update_mu(mu)
        energy = torch.cat(parts)
    else:
        energy = fused_ba(*args, n_iters, *rest, robust=kernel.fused_args())
    buffer.disps.clamp_(min=0.001)
    return energy
