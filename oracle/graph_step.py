"""FP64 restatement of the factor-graph update glue (oracle; test-only).

Follows, for a single-view pinhole graph:
  GraphBuffer.reproject_dense_disp   vipe/slam/components/buffer.py:705-726 with
  iproj_i_proj_j_disp                vipe/slam/maths/geom.py:187-263 and the
  pinhole model                      vipe/utils/cameras.py:131-181:
      X0 = ((u - cx)/fx, (v - cy)/fy, 1, d); X1 = T_ij X0 (SE3 on homogeneous
      points: R X0[:3] + t d); z < MIN_DEPTH (0.1) projects with z = 1
  motion features                    vipe/slam/components/factor_graph.py:247-260
      clamp([coords1 - coords0, target - coords1], -64, 64), channels first
  target / weight epilogue           factor_graph.py:268-277
Pinned by tests/test_oracle_graph_step.py: rotation-matrix forms of the
reprojection, identity and pure-translation closed forms, and the clamp.
"""
from __future__ import annotations

import torch

from . import geom

MIN_DEPTH = 0.1


def reproject(poses, disps, intrinsics, ii, jj):
    """coords1 [E,H,W,2] (float64)."""
    P = poses.double()
    D = disps.double()
    N, H, W = D.shape
    fx, fy, cx, cy = [float(v) for v in intrinsics.double()]
    v, u = torch.meshgrid(torch.arange(H, dtype=torch.float64), torch.arange(W, dtype=torch.float64), indexing="ij")
    X = torch.stack([(u - cx) / fx, (v - cy) / fy, torch.ones_like(u)], -1)
    out = torch.empty(ii.shape[0], H, W, 2, dtype=torch.float64)
    for e in range(ii.shape[0]):
        a, b = int(ii[e]), int(jj[e])
        t, q = geom.relative(P[a], P[b])
        X1 = geom._apply(t, q, X, D[a])
        z = torch.where(X1[..., 2] < MIN_DEPTH, torch.ones_like(X1[..., 2]), X1[..., 2])
        out[e, ..., 0] = fx * (X1[..., 0] / z) + cx
        out[e, ..., 1] = fy * (X1[..., 1] / z) + cy
    return out


def motion_features(poses, disps, intrinsics, ii, jj, target):
    """(coords1 [E,H,W,2], motn [E,4,H,W]) in float64."""
    c1 = reproject(poses, disps, intrinsics, ii, jj)
    E, H, W, _ = c1.shape
    v, u = torch.meshgrid(torch.arange(H, dtype=torch.float64), torch.arange(W, dtype=torch.float64), indexing="ij")
    c0 = torch.stack([u, v], -1)
    motn = torch.cat([c1 - c0, target.double() - c1], -1).permute(0, 3, 1, 2).clamp(-64.0, 64.0)
    return c1, motn


def targets(coords1, delta, weight_net, mask=None, frame=None):
    """(target, weight) [E,H,W,2]: coords1 + delta; weight zeroed where mask[frame[e]]."""
    t = coords1.double() + delta.double()
    w = weight_net.double().clone()
    if mask is not None:
        m = mask[frame].bool()
        w[m] = 0.0
    return t, w
