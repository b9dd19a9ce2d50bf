"""Restatement of ViPE's generic bundle-adjustment path (oracle and CPU baseline; test-only).

The reference runs dense BA either through the fused ``slam_ext.ba_extended``
kernel (restated in oracle/ba.py) or through its generic Python solver, which
this module restates for a single-view pinhole graph:

  GraphBuffer.bundle_adjustment, generic branch   vipe/slam/components/buffer.py:500-703
  DenseDepthFlowTerm.forward                       vipe/slam/ba/terms.py:98-262
  DispSensRegularizationTerm                       vipe/slam/ba/terms.py:265-327
  geom.iproj_i_proj_j_disp / actp                  vipe/slam/maths/geom.py:91-298
  PinholeCameraModel (MIN_DEPTH 0.1, Z < 0.1 -> 1) vipe/utils/cameras.py:45-217
  Solver.run_inplace + solve_scipy (spsolve)       vipe/slam/ba/solver.py:31-43, 116-202
  damping diag += ep + lm diag / + ep + vector     vipe/slam/maths/matrix.py:178-191, 345-357
  PoseRetractor / DenseDispRetractor (dx > 10 -> 0) / IntrinsicsRetractor
                                                   vipe/slam/maths/retractor.py:29-62

Where the generic path differs from the fused kernel:
- projection clamps Z < 0.1 to Z = 1 (the point still projects) and masks
  the weight with Z > 0.1; the fused kernel zeroes both below 0.25;
- the residual is coords - target, the rhs -J^T w r;
- the depth back-substitution uses every pose (no skip of local pose 0);
- poses of source frames outside [t0, t1) are fixed; a pose that only
  appears as a target stays free;
- the depth damping adds ep = 1e-7 on top of 0.2 damping + 1e-7;
- the depth-sensor prior covers every pixel of a source frame whose sensor
  map has a positive sum (the fused kernel: pixels with a positive reading);
- disparity steps larger than 10 are dropped; poses retract with lietorch's
  SE3 (normalised quaternions);
- the reduced system is solved with scipy's sparse LU (``spsolve``), in the
  tensors' dtype (float32 in the reference).

``dtype`` selects the arithmetic: float32 restates the reference as it runs
(the CPU baseline); float64 is the checker.  Pinned by tests/test_oracle_generic_ba.py:
finite differences of its residuals, a brute-force dense Gauss-Newton solve
over every unknown, the edge rules above on planted points, and agreement
with the fused oracle within the reference test's 2e-4 on a near-converged
graph (tests/test_graph_buffer_fused_ba.py:203-249).
"""
from __future__ import annotations

import numpy as np
import torch

from . import lie

MIN_DEPTH = 0.1


def _relative(poses: torch.Tensor, pi: torch.Tensor, pj: torch.Tensor):
    """Gij = Gj Gi^-1 for SE3 poses [N,7]: translation [E,3], rotation [E,3,3]."""
    ti, qi = poses[pi, :3], lie.qnormalize(poses[pi, 3:7])
    tj, qj = poses[pj, :3], lie.qnormalize(poses[pj, 3:7])
    q = lie.qmul(qj, lie.qconj(qi))
    t = tj - lie.qrot(q, ti)
    return t, lie.qmat(q)


def edge_terms(poses, disps, intr, target, weight, pi, pj, di, intrinsics_factor=1.0, with_focal=False):
    """Residuals and Jacobians of DenseDepthFlowTerm for every edge and pixel.

    poses [N,7] SE3 (world to camera), disps [N,H,W], intr [4] full-resolution
    pinhole (the term scales it by 1 / intrinsics_factor), target / weight
    [E,HW,2] (weight already scaled by 0.001).  Returns a dict of
    r [E,HW,2], w [E,HW,2] (weight times the validity mask), Ji / Jj
    [E,HW,2,6], Jz [E,HW,2], Jf [E,HW,2] (d/d focal of the full-resolution
    intrinsics, both pixels' cameras being the same one), z [E,HW].
    """
    dt = disps.dtype
    E = pi.shape[0]
    N, H, W = disps.shape
    HW = H * W
    fx, fy, cx, cy = [float(v) / intrinsics_factor for v in intr]
    vv, uu = torch.meshgrid(torch.arange(H, dtype=dt), torch.arange(W, dtype=dt), indexing="ij")
    X0 = ((uu - cx) / fx).reshape(HW)
    Y0 = ((vv - cy) / fy).reshape(HW)
    d = disps[di].reshape(E, HW)
    t, R = _relative(poses, pi, pj)
    # act4: [X, Y, Z, d] = [R [X0, Y0, 1] + t d, d]
    P0 = torch.stack([X0.expand(E, HW), Y0.expand(E, HW), torch.ones(E, HW, dtype=dt)], -1)
    X1 = (R[:, None] @ P0[..., None])[..., 0] + t[:, None, :] * d[..., None]
    X, Y, Z = X1.unbind(-1)
    valid = Z > MIN_DEPTH
    Zc = torch.where(Z < MIN_DEPTH, torch.ones_like(Z), Z)
    inv = 1.0 / Zc
    coords = torch.stack([fx * (X * inv) + cx, fy * (Y * inv) + cy], -1)
    r = coords - target
    w = weight * valid[..., None].to(dt)
    o = torch.zeros_like(inv)
    # d(proj)/d[X, Y, Z] at the clamped depth (the homogeneous 4th column is zero)
    Jp = torch.stack([torch.stack([fx * inv, o, -fx * X * inv * inv], -1),
                      torch.stack([o, fy * inv, -fy * Y * inv * inv], -1)], -2)  # [E,HW,2,3]
    # d[X, Y, Z]/d(xi_j) for Exp(xi) acting on [X, Y, Z, d]: [d I, -hat(X)] (actp)
    Ja = torch.zeros(E, HW, 3, 6, dtype=dt)
    Ja[..., 0, 0] = d
    Ja[..., 1, 1] = d
    Ja[..., 2, 2] = d
    Ja[..., :, 3:] = -lie.hat(X1)
    Jj = Jp @ Ja  # [E,HW,2,6]
    A = torch.zeros(E, 6, 6, dtype=dt)
    A[:, :3, :3] = R
    A[:, :3, 3:] = lie.hat(t) @ R
    A[:, 3:, 3:] = R
    Ji = -(Jj @ A[:, None])
    Jz = (Jp @ t[:, None, :, None])[..., 0]  # T acting on [0, 0, 0, 1] gives [t, 1]
    Jf = None
    if with_focal:
        # dX0/df = [-X0/fx, -Y0/fy, 0, 0] through R, plus the projection's own
        # dependence [X/Z, Y/Z]; scaled by 1 / factor (J_scale)
        dP = torch.stack([-X0 / fx, -Y0 / fy, torch.zeros_like(X0)], -1)
        dX1 = (R[:, None] @ dP[None, :, :, None])[..., 0]
        Jf = ((Jp @ dX1[..., None])[..., 0] + torch.stack([X * inv, Y * inv], -1)) / intrinsics_factor
    return {"r": r, "w": w, "Ji": Ji, "Jj": Jj, "Jz": Jz, "Jf": Jf, "z": Z}


def _spsolve(rows, cols, vals, rhs):
    """solve_scipy (solver.py:31-43): coo -> csr -> scipy spsolve, in rhs's dtype."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.linalg import spsolve

    n = rhs.shape[0]
    m = coo_matrix((vals.numpy(), (rows.numpy(), cols.numpy())), shape=(n, n)).tocsr()
    x = spsolve(m, rhs.numpy())
    return torch.from_numpy(np.asarray(x, dtype=rhs.numpy().dtype))


def gn_step(poses, disps, intr, disps_sens, target, weight, disp_damping, ii, jj, t0, t1, lm, ep, motion_only,
            limited_disp, optimize_intrinsics, alpha=0.001, intrinsics_factor=1.0, intrinsics_damping=1e-6,
            dtype=torch.float32, return_energy=False):
    """One Solver.run_inplace of the generic branch of bundle_adjustment, in place.

    poses [N,7], disps [N,H,W], intr [4] (full resolution; focal updated in
    place when optimising), disps_sens [N,H,W], target [E,HW,2], weight
    [E,HW,2] (raw: 0.001 is applied here), disp_damping [N,H,W], ii / jj [E].
    Returns the energy (sum of w r^2 and the prior) before the update."""
    N, H, W = disps.shape
    HW = H * W
    E = ii.shape[0]
    pi, pj, di = ii.long(), jj.long(), ii.long()
    pz = poses.to(dtype)
    dz = disps.to(dtype)
    T = edge_terms(pz, dz, intr.to(dtype), target.to(dtype), 0.001 * weight.to(dtype), pi, pj, di,
                   intrinsics_factor, with_focal=optimize_intrinsics)
    r, w = T["r"].reshape(E, HW * 2), T["w"].reshape(E, HW * 2)
    energy = (w * r * r).sum()

    # free variables: poses of the window (source poses outside it are fixed,
    # buffer.py:660-663), disps of the source frames (all of them fixed when
    # motion only; those of edges leaving the window with limited_disp)
    pi_unique = torch.unique(pi)
    fixed_pose = set(pi_unique[(pi_unique < t0) | (pi_unique >= t1)].tolist()) if t0 < t1 else set(range(N))
    pose_ids = sorted((set(pi.tolist()) | set(pj.tolist())) - fixed_pose)
    pslot = {p: k for k, p in enumerate(pose_ids)}
    di_unique = torch.unique(di)
    if motion_only:
        disp_ids = []
    else:
        fixed_d = set(di[(pi < t0) | (pi >= t1)].tolist()) if limited_disp else set()
        disp_ids = [k for k in di_unique.tolist() if k not in fixed_d]
    dslot = {k: s for s, k in enumerate(disp_ids)}
    nP, nD = len(pose_ids), len(disp_ids)
    nI = 1 if optimize_intrinsics else 0
    n = 6 * nP + nI

    Jp = torch.cat([T["Ji"], T["Jj"]], 0).reshape(2 * E, HW * 2, 6)
    side_pose = torch.cat([pi, pj])
    wr = torch.cat([w, w])
    rr = torch.cat([r, r])
    # pose-pose blocks (upper) and rhs
    H_pp = torch.zeros(n, n, dtype=dtype)
    b = torch.zeros(n, dtype=dtype)
    for a in range(2 * E):
        pa = int(side_pose[a])
        if pa not in pslot:
            continue
        sa = pslot[pa]
        wa = wr[a][:, None] * Jp[a]
        b[6 * sa:6 * sa + 6] -= (wa * rr[a][:, None]).sum(0)
        e = a % E
        for c in (e, e + E):
            pc = int(side_pose[c])
            if pc not in pslot:
                continue
            H_pp[6 * sa:6 * sa + 6, 6 * pslot[pc]:6 * pslot[pc] + 6] += wa.T @ Jp[c]
    if optimize_intrinsics:
        Jf = T["Jf"].reshape(E, HW * 2)
        H_pp[n - 1, n - 1] += (w * Jf * Jf).sum()
        b[n - 1] -= (w * r * Jf).sum()
        for a in range(2 * E):
            pa = int(side_pose[a])
            if pa in pslot:
                sa = pslot[pa]
                v = (wr[a][:, None] * Jp[a] * Jf[a % E][:, None]).sum(0)
                H_pp[6 * sa:6 * sa + 6, n - 1] += v
                H_pp[n - 1, 6 * sa:6 * sa + 6] += v
    # damping (apply_damping_assume_coalesced): pose diag += ep + lm diag,
    # intrinsics diag += 1e-6 + 1e-6 diag (intrinsics_damping_scale 1)
    diag = torch.diagonal(H_pp)
    lmv = torch.full((n,), float(lm), dtype=dtype)
    epv = torch.full((n,), float(ep), dtype=dtype)
    if optimize_intrinsics:
        lmv[-1] = intrinsics_damping
        epv[-1] = intrinsics_damping
    H_pp = H_pp + torch.diag(epv + lmv * diag)

    # dense disparity: per pixel C, the pose coupling E and the rhs, marginalised
    if nD:
        Jz = T["Jz"].reshape(E, HW, 2)
        wz = w.reshape(E, HW, 2)
        rz = r.reshape(E, HW, 2)
        C = torch.zeros(nD, HW, dtype=dtype)
        bz = torch.zeros(nD, HW, dtype=dtype)
        dsl = torch.tensor([dslot.get(int(k), -1) for k in di])
        sel = dsl >= 0
        C.index_add_(0, dsl[sel], (wz * Jz * Jz).sum(-1)[sel])
        bz.index_add_(0, dsl[sel], -(wz * rz * Jz).sum(-1)[sel])
        ids = torch.tensor(disp_ids)
        # damping vector 0.2 damping + 1e-7 over every source frame, ep 1e-7
        # (buffer.py:667-675).  It is placed by searchsorted among the free
        # frames (matrix.py:187-191), so with limited_disp a fixed frame's
        # damping lands on the next free frame; past the last one the
        # reference raises an index error, and so does this restatement.
        pos = torch.searchsorted(ids, di_unique)
        if int(pos.max()) >= nD:
            raise IndexError("dense_disp damping index past the last free frame (as in the reference)")
        damp = 0.2 * disp_damping.to(dtype).reshape(N, HW)[di_unique] + 1e-7
        C.index_add_(0, pos, damp + 1e-7)
        # depth-sensor prior on source frames whose sensor map has a positive sum
        ds = disps_sens.to(dtype).reshape(N, HW)
        has_sens = ds[ids].sum(1) > 0
        res = dz.reshape(N, HW)[ids] - ds[ids]
        C += torch.where(has_sens[:, None], torch.full_like(C, alpha), torch.zeros_like(C))
        bz -= torch.where(has_sens[:, None], alpha * res, torch.zeros_like(res))
    if return_energy and not motion_only:
        # the prior's energy covers every sensor frame among the sources, fixed or not
        ds_all = disps_sens.to(dtype).reshape(N, HW)[di_unique]
        res_all = dz.reshape(N, HW)[di_unique] - ds_all
        energy = energy + (alpha * res_all * res_all)[ds_all.sum(1) > 0].sum()
    if nD:
        # E rows (pose side a, disp frame of edge a % E): [HW, 6] + intrinsics column
        Ez = []
        for a in range(2 * E):
            e = a % E
            Ja = Jp[a].reshape(HW, 2, 6)
            Ez.append((wz[e][..., None] * Jz[e][..., None] * Ja).sum(1))  # [HW,6]
        Cinv = 1.0 / C
        S = torch.zeros(n, n, dtype=dtype)
        v = torch.zeros(n, dtype=dtype)
        Ef = None
        if optimize_intrinsics:
            Jf = T["Jf"].reshape(E, HW, 2)
            Ef = torch.zeros(nD, HW, dtype=dtype)
            Ef.index_add_(0, dsl[sel], (wz * Jz * Jf).sum(-1)[sel])
        # rows sharing a disparity frame couple their poses (Schur complement)
        by_frame = {}
        for a in range(2 * E):
            pa, ka = int(side_pose[a]), int(dsl[a % E])
            if pa in pslot and ka >= 0:
                by_frame.setdefault(ka, []).append((a, pslot[pa]))
        for ka, members in by_frame.items():
            for a, sa in members:
                EaC = Ez[a] * Cinv[ka][:, None]
                v[6 * sa:6 * sa + 6] += (EaC * bz[ka][:, None]).sum(0)
                if Ef is not None:
                    cf = (EaC * Ef[ka][:, None]).sum(0)
                    S[6 * sa:6 * sa + 6, n - 1] += cf
                    S[n - 1, 6 * sa:6 * sa + 6] += cf
                for c, sc in members:
                    S[6 * sa:6 * sa + 6, 6 * sc:6 * sc + 6] += EaC.T @ Ez[c]
        if Ef is not None:
            S[n - 1, n - 1] += (Ef * Cinv * Ef).sum()
            v[n - 1] += (Ef * Cinv * bz).sum()
        Ared, bred = H_pp - S, b - v
    else:
        Ared, bred = H_pp, b
    # sparse LU of the reduced system (only its non-zero entries are passed)
    nz = torch.nonzero(Ared != 0, as_tuple=True)
    x = _spsolve(nz[0], nz[1], Ared[nz], bred) if n else torch.zeros(0, dtype=dtype)

    # retractions
    for p, s in pslot.items():
        xi = x[6 * s:6 * s + 6].to(torch.float64)
        newp = lie.SE3.mul(lie.SE3.exp(xi[None]), poses[p:p + 1].to(torch.float64))
        poses[p] = newp[0].to(poses.dtype)
    if optimize_intrinsics:
        intr[:2] = (intr[:2].to(dtype) + x[n - 1]).to(intr.dtype)
    if nD:
        # back-substitution over every pose row: dz = C^-1 (bz - E^T x)
        dw = torch.zeros(nD, HW, dtype=dtype)
        for ka, members in by_frame.items():
            for a, sa in members:
                dw[ka] += Ez[a] @ x[6 * sa:6 * sa + 6]
        dxz = Cinv * (bz - dw)
        if Ef is not None:
            dxz = dxz - Cinv * Ef * x[n - 1]
        dxz = torch.where(dxz > 10, torch.zeros_like(dxz), dxz)
        flat = disps.view(N, HW)
        flat[ids] = (flat[ids].to(dtype) + dxz).to(disps.dtype)
    return float(energy)


def bundle_adjustment(poses, disps, intr, disps_sens, target, weight, disp_damping, ii, jj, t0, t1, n_iters, lm, ep,
                      motion_only=False, limited_disp=False, optimize_intrinsics=False, alpha=0.001,
                      intrinsics_factor=1.0, dtype=torch.float32):
    """n_iters generic iterations, then the disparity clamp at 0.001
    (buffer.py:697-702).  Returns the per-iteration energies."""
    en = [gn_step(poses, disps, intr, disps_sens, target, weight, disp_damping, ii, jj, t0, t1, lm, ep, motion_only,
                  limited_disp, optimize_intrinsics, alpha, intrinsics_factor, dtype=dtype, return_energy=True)
          for _ in range(n_iters)]
    disps.clamp_(min=0.001)
    return en
