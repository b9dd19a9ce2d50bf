"""FP64 restatement of the slam_ext geometry siblings (oracle; test-only).

Follows csrc/slam_ext/geom_kernels.cu:
  relSE3 / actSE3              :69-114   T_ij = T_j T_i^-1; (X, d) -> (R X + t d, d)
  projmap_kernel               :526-611  coords stay (u, v) unless z > 0.01; third
                                         channel zero; valid = z > MIN_DEPTH (0.25)
  frame_distance_kernel        :613-768  beta-weighted full-motion and translation-only
                                         flow magnitudes; 1000 when the valid weight
                                         is below 0.75 of the total
  depth_filter_kernel          :770-885  neighbours ix-1..ix-3, ix+1..ix+3 in range;
                                         top-left corner inside [0, W-1) x [0, H-1);
                                         a hit when |1/dj - 1/d_corner| < thresh for
                                         any of the 4 corners (at most one per neighbour)
  iproj_kernel                 :887-953  points = (R X + t d) / d
Poses are [t, q xyzw]; the rotation is the unit-quaternion rotation.

Everything is computed in float64 from the float32 inputs.  Pinned by
tests/test_oracle_geom.py: iproj / projmap against rotation-matrix forms of
the same maps built from Rodrigues' formula, identity and pure-translation
closed forms of frame_distance, depth_filter on constant-depth scenes and a
per-pixel loop on small inputs.
"""
from __future__ import annotations

import torch

MIN_DEPTH = 0.25


def _quat_rotate(q, X):
    """Rotate X [..., 3] by unit quaternions q [..., 4] (x, y, z, w)."""
    qv, w = q[..., :3], q[..., 3:4]
    u = 2.0 * torch.cross(qv, X, dim=-1)
    return X + w * u + torch.cross(qv, u, dim=-1)


def _quat_mul(a, b):
    ax, ay, az, aw = a.unbind(-1)
    bx, by, bz, bw = b.unbind(-1)
    return torch.stack([aw * bx + ax * bw + ay * bz - az * by,
                        aw * by - ax * bz + ay * bw + az * bx,
                        aw * bz + ax * by - ay * bx + az * bw,
                        aw * bw - ax * bx - ay * by - az * bz], -1)


def relative(Pi, Pj):
    """T_j T_i^-1 for poses [..., 7]: (t, q)."""
    qi_inv = torch.cat([-Pi[..., 3:6], Pi[..., 6:7]], -1)
    q = _quat_mul(Pj[..., 3:7], qi_inv)
    t = Pj[..., :3] - _quat_rotate(q, Pi[..., :3])
    return t, q


def _pixel_rays(H, W, intr):
    fx, fy, cx, cy = [float(v) for v in intr]
    v, u = torch.meshgrid(torch.arange(H, dtype=torch.float64), torch.arange(W, dtype=torch.float64), indexing="ij")
    X = torch.stack([(u - cx) / fx, (v - cy) / fy, torch.ones_like(u)], -1)
    return u, v, X


def _apply(t, q, X, d):
    """(R X + t d) with t [3], q [4], X [H,W,3], d [H,W]."""
    return _quat_rotate(q.expand(X.shape[:-1] + (4,)), X) + d[..., None] * t


def iproj(poses, disps, intrinsics):
    P = poses.double()
    D = disps.double()
    N, H, W = D.shape
    _, _, X = _pixel_rays(H, W, intrinsics.double())
    out = torch.empty(N, H, W, 3, dtype=torch.float64)
    for n in range(N):
        Y = _apply(P[n, :3], P[n, 3:7], X, D[n])
        out[n] = Y / D[n][..., None]
    return out


def projmap(poses, disps, intrinsics, ii, jj):
    P = poses.double()
    D = disps.double()
    N, H, W = D.shape
    fx, fy, cx, cy = [float(v) for v in intrinsics.double()]
    u, v, X = _pixel_rays(H, W, intrinsics.double())
    E = ii.shape[0]
    coords = torch.zeros(E, H, W, 3, dtype=torch.float64)
    valid = torch.zeros(E, H, W, 1, dtype=torch.float64)
    for e in range(E):
        a, b = int(ii[e]), int(jj[e])
        t, q = relative(P[a], P[b])
        Y = _apply(t, q, X, D[a])
        front = Y[..., 2] > 0.01
        coords[e, ..., 0] = torch.where(front, fx * (Y[..., 0] / Y[..., 2]) + cx, u)
        coords[e, ..., 1] = torch.where(front, fy * (Y[..., 1] / Y[..., 2]) + cy, v)
        valid[e, ..., 0] = (Y[..., 2] > MIN_DEPTH).double()
    return coords, valid


def frame_distance(poses, disps, intrinsics, pi, pj, qi, qj, di, beta):
    P = poses.double()
    D = disps.double()
    I = intrinsics.double()
    N, H, W = D.shape
    beta = float(beta)
    out = torch.empty(pi.shape[0], dtype=torch.float64)
    for b in range(pi.shape[0]):
        a, c, dd = int(pi[b]), int(pj[b]), int(di[b])
        Ii, Ij = I[int(qi[b])], I[int(qj[b])]
        u, v, X = _pixel_rays(H, W, Ii)
        fxj, fyj, cxj, cyj = [float(x) for x in Ij]
        t, q = relative(P[a], P[c])
        d = D[dd]
        Y = _apply(t, q, X, d)
        m_full = torch.sqrt((fxj * Y[..., 0] / Y[..., 2] + cxj - u) ** 2 + (fyj * Y[..., 1] / Y[..., 2] + cyj - v) ** 2)
        Z = X + d[..., None] * t
        m_tr = torch.sqrt((fxj * Z[..., 0] / Z[..., 2] + cxj - u) ** 2 + (fyj * Z[..., 1] / Z[..., 2] + cyj - v) ** 2)
        ok_f = Y[..., 2] > MIN_DEPTH
        ok_t = Z[..., 2] > MIN_DEPTH
        total = H * W * (beta + (1.0 - beta))
        valid = beta * ok_f.sum().item() + (1.0 - beta) * ok_t.sum().item()
        acc = beta * m_full[ok_f].sum().item() + (1.0 - beta) * m_tr[ok_t].sum().item()
        out[b] = 1000.0 if valid / (total + 1e-8) < 0.75 else acc / valid
    return out


def depth_filter(poses, disps, intrinsics, ix, thresh):
    P = poses.double()
    D = disps.double()
    N, H, W = D.shape
    fx, fy, cx, cy = [float(v) for v in intrinsics.double()]
    u, v, X = _pixel_rays(H, W, intrinsics.double())
    out = torch.zeros(ix.shape[0], H, W, dtype=torch.float64)
    for b in range(ix.shape[0]):
        i = int(ix[b])
        th = float(thresh[b])
        d = D[i]
        for n in range(6):
            j = i - n - 1 if n < 3 else i + n - 2
            if j < 0 or j >= N:
                continue
            t, q = relative(P[i], P[j])
            Y = _apply(t, q, X, d)
            uj = fx * (Y[..., 0] / Y[..., 2]) + cx
            vj = fy * (Y[..., 1] / Y[..., 2]) + cy
            dj = d / Y[..., 2]
            u0 = torch.floor(uj)
            v0 = torch.floor(vj)
            inside = (u0 >= 0) & (v0 >= 0) & (u0 < W - 1) & (v0 < H - 1)
            u0c = u0.clamp(0, W - 2).long()
            v0c = v0.clamp(0, H - 2).long()
            Dj = D[j]
            inv = 1.0 / dj
            hit = torch.zeros_like(inside)
            for dv_, du_ in ((0, 0), (0, 1), (1, 0), (1, 1)):
                hit |= (inv - 1.0 / Dj[v0c + dv_, u0c + du_]).abs() < th
            out[b] += (inside & hit).double()
    return out
