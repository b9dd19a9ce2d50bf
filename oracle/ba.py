"""FP64 restatement of slam_ext.ba / ba_extended (oracle; test-only).

Follows csrc/slam_ext/geom_kernels.cu:
  projective_transform_kernel :178-524  per-edge residuals, Jacobians, Hessian blocks,
                                         per-pixel depth terms (MIN_DEPTH 0.25, :33;
                                         weight scale 0.001, :321-322; stereo edge
                                         ii == jj -> Tij = (-0.1,0,0; I) and no pose terms,
                                         :231-241, :359, :407)
  SparseBlock / SparseSystem  :1197-1424 block triplets, damping diag += ep + lm*diag,
                                         LLT, zeros on failure (:1269-1288)
  accum_cuda                  :1043-1089 segmented sums
  schur_block (+_with_intr.)  :1426-1624 Schur rows share a depth frame
  EvT6x1_kernel               :1179-1195 back-substitution skips rows with local pose
                                         index <= 0 (pose t0 never reaches the depth
                                         update)
  ba_cuda_impl                :1626-1820 iteration order, energy before the update,
                                         retractions (:993-1041)
and the caller contract of vipe/slam/components/buffer.py:364-489.

Sim3 poses [N,8] (t, q, s) run the same step with 7-dof pose blocks: the
point map s R X + t d and the actp Jacobian of vipe/slam/maths/geom.py:147-182
(whose scale column the projection sends to zero on the j side), the Sim3
adjoint for the i side (sim3.h Adj), and the retraction T <- Exp(dx) T with
lietorch's Sim3 Exp and product (sim3.h, groups.py:157-160).  The reference's
fused kernel is SE3 only; this branch is pinned by the same finite-difference
and dense Gauss-Newton checks as the SE3 one.

Everything is computed in float64 on the CPU from the float32 inputs; poses,
disps and intrinsics are updated in place (rounded back to the caller's
dtype at the end of each iteration, as the reference stores them).

Pinned by tests/test_oracle_ba.py through: finite differences of the
residuals against the analytic Jacobians; per-edge Hessian blocks equal
J^T W J of the explicit Jacobian; the Schur path equals a brute-force dense
Gauss-Newton solve over all pose and depth unknowns (without the
back-substitution quirk, which is then checked on its own); the reference
tests' invariants (ba == ba_extended, inactive and motion-only frames
unchanged bit-exactly, positive finite energy).
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import robust as robust_mod

from . import lie

F64 = torch.float64
MIN_DEPTH = 0.25


@dataclass
class Bookkeeping:
    """Index bookkeeping shared by every iteration (geom_kernels.cu:1648-1655, 1437-1468)."""
    P: int
    ts: torch.Tensor
    ii_exp: torch.Tensor
    jj_exp: torch.Tensor
    kx: torch.Tensor
    kk_exp: torch.Tensor


def bookkeeping(ii: torch.Tensor, jj: torch.Tensor, t0: int, t1: int) -> Bookkeeping:
    ts = torch.arange(t0, t1, dtype=torch.long)
    ii_exp = torch.cat([ts, ii.long()])
    jj_exp = torch.cat([ts, jj.long()])
    kx, kk_exp = torch.unique(ii_exp, sorted=True, return_inverse=True)
    return Bookkeeping(t1 - t0, ts, ii_exp, jj_exp, kx, kk_exp)


def pose_dof(poses: torch.Tensor) -> int:
    """6 for SE3 poses [N,7] (t, q), 7 for Sim3 poses [N,8] (t, q, s)."""
    return 7 if poses.shape[-1] == 8 else 6


def relative_poses(poses: torch.Tensor, ii: torch.Tensor, jj: torch.Tensor):
    """Tij = Tj * Ti^-1 (relSE3, geom_kernels.cu:101-114); stereo edges get
    (t=(-0.1,0,0), q=identity) (:231-241).  For Sim3 poses [N,8] the relative
    scale is s_j / s_i and t = t_j - s R t_i (group product of sim3.h).
    Returns t [E,3], q [E,4], s [E] (ones for SE3)."""
    P = poses.to(F64)
    ti, qi = P[ii, :3], P[ii, 3:7]
    tj, qj = P[jj, :3], P[jj, 3:7]
    qij = lie.qmul(qj, lie.qconj(qi))
    if pose_dof(poses) == 7:
        sij = P[jj, 7] / P[ii, 7]
    else:
        sij = torch.ones(ii.shape[0], dtype=F64)
    tij = tj - sij[:, None] * lie.qrot(qij, ti)
    stereo = (ii == jj)
    if stereo.any():
        tij = tij.clone()
        qij = qij.clone()
        sij = sij.clone()
        tij[stereo] = torch.tensor([-0.1, 0.0, 0.0], dtype=F64)
        qij[stereo] = torch.tensor([0.0, 0.0, 0.0, 1.0], dtype=F64)
        sij[stereo] = 1.0
    return tij, qij, sij


def adj_T(t: torch.Tensor, q: torch.Tensor, s: torch.Tensor | None = None, dof: int = 6) -> torch.Tensor:
    """Adj(Tij) for the [tau, phi(, sigma)] ordering: SE3 [E,6,6] =
    [[R, [t]x R], [0, R]]; Sim3 [E,7,7] = [[sR, [t]x R, -t], [0, R, 0], [0, 0, 1]]
    (se3.h / sim3.h Adj)."""
    R = lie.qmat(q)
    A = torch.zeros(t.shape[0], dof, dof, dtype=F64)
    A[:, :3, :3] = R if s is None or dof == 6 else s[:, None, None] * R
    A[:, :3, 3:6] = lie.hat(t) @ R
    A[:, 3:6, 3:6] = R
    if dof == 7:
        A[:, :3, 6] = -t
        A[:, 6, 6] = 1.0
    return A


@dataclass
class EdgeTerms:
    ru: torch.Tensor      # [E, HW]
    rv: torch.Tensor
    wu: torch.Tensor      # 0.001*weight, zeroed where z < MIN_DEPTH
    wv: torch.Tensor
    Ju: torch.Tensor      # [E, HW, D]  d(proj_u)/d(xi_j), D = 6 (SE3) or 7 (Sim3)
    Jv: torch.Tensor
    Jzu: torch.Tensor     # [E, HW]     d(proj_u)/d(disp_i)
    Jzv: torch.Tensor
    Jfu: torch.Tensor     # [E, HW]     d(proj_u)/d(f) * intrinsics_scale
    Jfv: torch.Tensor
    Adj: torch.Tensor     # [E, D, D]
    proj_u: torch.Tensor
    proj_v: torch.Tensor
    z: torch.Tensor


def edge_terms(poses, disps, intrinsics, targets, weights, ii, jj, intrinsics_scale=1.0, robust=None) -> EdgeTerms:
    """Per-pixel residuals and Jacobians of projective_transform_kernel (:298-435).
    ``robust`` = (kind, threshold, mu, mu_max) multiplies each residual
    component's weight by the robust kernel's weight of that residual
    (vipe/slam/ba/terms.py:73-75; weights in oracle/robust.py)."""
    E = ii.shape[0]
    N, H, W = disps.shape
    HW = H * W
    D = pose_dof(poses)
    fx, fy, cx, cy = [float(v) for v in intrinsics.to(F64)]
    tij, qij, sij = relative_poses(poses, ii, jj)
    # the homogeneous point [X0, Y0, 1, d] maps to [s R X + t d, d] (act4)
    R = sij[:, None, None] * lie.qmat(qij)
    vv, uu = torch.meshgrid(torch.arange(H, dtype=F64), torch.arange(W, dtype=F64), indexing="ij")
    X0 = ((uu - cx) / fx).reshape(HW)
    Y0 = ((vv - cy) / fy).reshape(HW)
    d = disps.to(F64)[ii].reshape(E, HW)
    Xi3 = torch.stack([X0.expand(E, HW), Y0.expand(E, HW), torch.ones(E, HW, dtype=F64)], -1)  # [E,HW,3]
    Xj = (R[:, None] @ Xi3[..., None])[..., 0] + tij[:, None, :] * d[..., None]
    x, y, z = Xj[..., 0], Xj[..., 1], Xj[..., 2]
    h = d
    valid = z >= MIN_DEPTH
    zs = torch.where(valid, z, torch.ones_like(z))
    dinv = torch.where(valid, 1.0 / zs, torch.zeros_like(z))
    d2 = dinv * dinv
    wgt = weights.to(F64).reshape(E, 2, HW)
    tgt = targets.to(F64).reshape(E, 2, HW)
    wu = torch.where(valid, 0.001 * wgt[:, 0], torch.zeros_like(z))
    wv = torch.where(valid, 0.001 * wgt[:, 1], torch.zeros_like(z))
    proj_u = fx * dinv * x + cx
    proj_v = fy * dinv * y + cy
    ru = tgt[:, 0] - proj_u
    rv = tgt[:, 1] - proj_v
    if robust is not None and robust[0] != robust_mod.NONE:
        kind, c, mu, mu_max = robust
        wu = wu * robust_mod.weight(ru, kind, c, mu, mu_max)
        wv = wv * robust_mod.weight(rv, kind, c, mu, mu_max)
    o = torch.zeros_like(z)
    # d(proj)/d(xi_j) = d(proj)/dXj * d(Exp(xi) Xj)/dxi at xi = 0 (geom.py:105-182 actp).
    # For Sim3 the scale column of the point Jacobian is [x, y, z], which the
    # projection maps to fx (x/z - x z/z^2) = 0: the j-side sigma column is zero.
    Ju = fx * torch.stack([h * dinv, o, -x * h * d2, -x * y * d2, 1 + x * x * d2, -y * dinv] + [o] * (D - 6), -1)
    Jv = fy * torch.stack([o, h * dinv, -y * h * d2, -1 - y * y * d2, x * y * d2, x * dinv] + [o] * (D - 6), -1)
    Jzu = fx * (tij[:, 0:1] * dinv - tij[:, 2:3] * (x * d2))
    Jzv = fy * (tij[:, 1:2] * dinv - tij[:, 2:3] * (y * d2))
    # intrinsics (:325-331): dXi/df = (-Xi0/fx, -Xi1/fy, 0, 0) mapped by R only
    dXi = torch.stack([-X0 / fx, -Y0 / fy, torch.zeros_like(X0)], -1)  # [HW,3]
    dXj = (R[:, None] @ dXi[None, :, :, None])[..., 0]  # [E,HW,3]
    s = float(intrinsics_scale)
    Jfu = s * (fx * dinv * dXj[..., 0] - fx * x * d2 * dXj[..., 2] + x * dinv)
    Jfv = s * (fy * dinv * dXj[..., 1] - fy * y * d2 * dXj[..., 2] + y * dinv)
    return EdgeTerms(ru, rv, wu, wv, Ju, Jv, Jzu, Jzv, Jfu, Jfv, adj_T(tij, qij, sij, D), proj_u, proj_v, z)


def _sym(E):
    return E


def ba_extended(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, depth_active,
                t0, t1, iterations, lm, ep, motion_only, alpha, optimize_intrinsics,
                intrinsics_lm, intrinsics_ep, intrinsics_scale, compute_energy=False,
                skip_quirk=True, solver="cholesky", robust=None):
    """Restatement of ba_cuda_impl (geom_kernels.cu:1626-1820).

    Mutates ``poses``, ``disps`` and ``intrinsics`` in place like the reference
    and returns [dx [P,D], dz [|kx|,HW] or None, energy [iterations] or [0]]
    (D = 6 for SE3 poses [N,7], 7 for Sim3 poses [N,8])
    in float64.  ``skip_quirk=False`` disables the EvT skip of local pose 0
    (used only to compare against the brute-force dense solve).
    """
    E = ii.shape[0]
    N, H, W = disps.shape
    HW = H * W
    bk = bookkeeping(ii.cpu(), jj.cpu(), t0, t1)
    P, ts, ii_exp, jj_exp, kx, kk_exp = bk.P, bk.ts, bk.ii_exp, bk.jj_exp, bk.kx, bk.kk_exp
    K = kx.shape[0]
    ii = ii.long().cpu()
    jj = jj.long().cpu()
    dact = depth_active.to(F64).cpu()
    stereo = (ii == jj)
    dx = dz = None
    energy = torch.zeros(iterations if compute_energy else 0, dtype=F64)
    D = pose_dof(poses)
    dim = D * P + (1 if optimize_intrinsics else 0)

    for itr in range(iterations):
        T = edge_terms(poses, disps, intrinsics, targets, weights, ii, jj, intrinsics_scale, robust)
        act_d = dact[ii][:, None]
        wdu, wdv = T.wu * act_d, T.wv * act_d
        Cii = wdu * T.Jzu ** 2 + wdv * T.Jzv ** 2
        bz = wdu * T.ru * T.Jzu + wdv * T.rv * T.Jzv
        Efi = wdu * T.Jzu * T.Jfu + wdv * T.Jzv * T.Jfv
        Eij = (wdu * T.Jzu)[..., None] * T.Ju + (wdv * T.Jzv)[..., None] * T.Jv  # [E,HW,D]
        At = T.Adj.transpose(1, 2)
        Jiu = -(At[:, None] @ T.Ju[..., None])[..., 0]
        Jiv = -(At[:, None] @ T.Jv[..., None])[..., 0]
        Eii = (wdu * T.Jzu)[..., None] * Jiu + (wdv * T.Jzv)[..., None] * Jiv
        pu = torch.where(stereo[:, None], torch.zeros_like(T.wu), T.wu)
        pv = torch.where(stereo[:, None], torch.zeros_like(T.wv), T.wv)
        Jxu = torch.cat([Jiu, T.Ju], -1)  # [E,HW,2D]
        Jxv = torch.cat([Jiv, T.Jv], -1)
        Hx = torch.einsum("ep,epa,epb->eab", pu, Jxu, Jxu) + torch.einsum("ep,epa,epb->eab", pv, Jxv, Jxv)
        vx = torch.einsum("ep,epa->ea", pu * T.ru, Jxu) + torch.einsum("ep,epa->ea", pv * T.rv, Jxv)
        Fx = torch.einsum("ep,epa->ea", pu * T.Jfu, Jxu) + torch.einsum("ep,epa->ea", pv * T.Jfv, Jxv)
        Hff = (pu * T.Jfu ** 2 + pv * T.Jfv ** 2).sum()
        vf = (pu * T.ru * T.Jfu + pv * T.rv * T.Jfv).sum()
        flow_energy = (T.wu * T.ru ** 2 + T.wv * T.rv ** 2).sum()

        # ---- pose system A, b (update_lhs/update_rhs; indices outside [0,P) dropped)
        A = torch.zeros(dim, dim, dtype=F64)
        b = torch.zeros(dim, dtype=F64)
        li, lj = ii - t0, jj - t0
        inw = lambda k: (0 <= k) and (k < P)  # noqa: E731
        for e in range(E):
            a_, b_ = int(li[e]), int(lj[e])
            blocks = [(a_, a_, Hx[e, :D, :D]), (a_, b_, Hx[e, :D, D:]), (b_, a_, Hx[e, D:, :D]), (b_, b_, Hx[e, D:, D:])]
            for r_, c_, blk in blocks:
                if inw(r_) and inw(c_):
                    A[D * r_:D * r_ + D, D * c_:D * c_ + D] += blk
            if inw(a_):
                b[D * a_:D * a_ + D] += vx[e, :D]
                if optimize_intrinsics:
                    A[D * a_:D * a_ + D, D * P] += Fx[e, :D]
                    A[D * P, D * a_:D * a_ + D] += Fx[e, :D]
            if inw(b_):
                b[D * b_:D * b_ + D] += vx[e, D:]
                if optimize_intrinsics:
                    A[D * b_:D * b_ + D, D * P] += Fx[e, D:]
                    A[D * P, D * b_:D * b_ + D] += Fx[e, D:]
        if optimize_intrinsics:
            A[D * P, D * P] += Hff
            b[D * P] += vf

        if compute_energy:
            en = flow_energy
        if not motion_only:
            active = dact[kx][:, None]
            dk = disps.to(F64)[kx].reshape(K, HW)
            dsk = disps_sens.to(F64)[kx].reshape(K, HW)
            m = (dsk > 0).to(F64) * active
            if compute_energy:
                en = en + (m * alpha * (dk - dsk) ** 2).sum()
            kpos = torch.searchsorted(kx, ii)  # edge -> position in kx
            Csum = torch.zeros(K, HW, dtype=F64).index_add_(0, kpos, Cii)
            bsum = torch.zeros(K, HW, dtype=F64).index_add_(0, kpos, bz)
            C = active * (Csum + eta.to(F64).reshape(K, HW) + m * alpha) + (1 - active)
            w = active * (bsum - m * alpha * (dk - dsk))
            Q = 1.0 / C
            # E rows: [ts rows: sum of Eii over edges with ii == t] + [edge rows: Eij]
            Ei = torch.zeros(P, HW, D, dtype=F64)
            sel = (ii >= t0) & (ii < t1)
            Ei.index_add_(0, (ii - t0)[sel], Eii[sel])
            Er = torch.cat([Ei, Eij], 0) * dact[ii_exp][:, None, None]  # [P+E, HW, D]
            rows_pose = jj_exp - t0
            Ef = None
            if optimize_intrinsics:
                Ef = torch.zeros(K, HW, dtype=F64).index_add_(0, kpos, Efi) * active
            # Schur: S(pa,pb) += sum_p E_a Q E_b^T for rows sharing a depth frame
            S = torch.zeros(dim, dim, dtype=F64)
            v = torch.zeros(dim, dtype=F64)
            for k in range(K):
                rows = torch.nonzero(kk_exp == k).flatten()
                Qk = Q[k]
                for ra in rows.tolist():
                    pa = int(rows_pose[ra])
                    if pa < 0 or pa >= P:
                        continue
                    Ea = Er[ra]  # [HW,D]
                    v[D * pa:D * pa + D] += (Ea * (Qk * w[k])[:, None]).sum(0)
                    if optimize_intrinsics:
                        spf = (Ea * (Qk * Ef[k])[:, None]).sum(0)
                        S[D * pa:D * pa + D, D * P] += spf
                        S[D * P, D * pa:D * pa + D] += spf
                    for rb in rows.tolist():
                        pb = int(rows_pose[rb])
                        if pb < 0 or pb >= P:
                            continue
                        S[D * pa:D * pa + D, D * pb:D * pb + D] += (Ea * Qk[:, None]).T @ Er[rb]
            if optimize_intrinsics:
                S[D * P, D * P] += (Ef * Q * Ef).sum()
                v[D * P] += (Ef * Q * w).sum()
            Ared, bred = A - S, b - v
        else:
            Ared, bred = A, b
        if compute_energy:
            energy[itr] = en

        x, ok = damped_solve(Ared, bred, D * P, lm, ep, intrinsics_lm, intrinsics_ep, optimize_intrinsics, solver)
        dxv = x[:D * P].reshape(P, D)
        df = float(x[D * P]) if optimize_intrinsics else 0.0

        if not motion_only:
            # EvT back-substitution (:1179-1195, :1797-1807)
            dw = torch.zeros(P + E, HW, dtype=F64)
            for r in range(P + E):
                pr = int(rows_pose[r])
                lo = 1 if skip_quirk else 0
                if pr < lo or pr >= P:
                    continue
                dw[r] = Er[r] @ dxv[pr]
            dwk = torch.zeros(K, HW, dtype=F64).index_add_(0, kk_exp, dw)
            dzv = Q * (w - dwk)
            if optimize_intrinsics:
                dzv = dzv - Q * Ef * df
        # retractions (:1809-1819); Sim3 poses: T <- Exp(dx) * T with lietorch's
        # Sim3 Exp and product (sim3.h Exp, operator*; groups.py:157-160 retr)
        if D == 7:
            newp = lie.Sim3.mul(lie.Sim3.exp(dxv), poses.to(F64)[t0:t1])
        else:
            newp = lie.droid_retr_se3(dxv, poses.to(F64)[t0:t1])
        poses[t0:t1] = newp.to(poses.dtype)
        if optimize_intrinsics:
            intrinsics[0:2] = (intrinsics.to(F64)[0:2] + df * float(intrinsics_scale)).to(intrinsics.dtype)
        if not motion_only:
            dflat = disps.view(N, HW)
            dflat[kx] = (dflat[kx].to(F64) + dzv).to(disps.dtype)
            dz = dzv
        dx = dxv
    return [dx, dz, energy]


def damped_solve(A, b, n_pose, lm, ep, ilm, iep, with_intr, solver="cholesky"):
    """SparseBlock/SparseSystem::solve: diag += ep + lm*diag, LLT, zeros on failure
    (geom_kernels.cu:1269-1288, 1389-1417).  Eigen's SimplicialLLT fails when a
    pivot is not positive; a dense Cholesky of the same matrix fails on the same
    condition (principal minors do not depend on the elimination order)."""
    L = A.clone()
    n = L.shape[0]
    d = torch.diagonal(L).clone()
    for i in range(n):
        is_s = i >= n_pose  # the intrinsics scalar follows the pose unknowns
        lmv = ilm if is_s else lm
        epv = iep if is_s else ep
        L[i, i] = d[i] + epv + lmv * d[i]
    if solver == "cholesky":
        Lc, info = torch.linalg.cholesky_ex(L)
        if int(info) != 0:
            return torch.zeros(n, dtype=F64), False
        return torch.cholesky_solve(b[:, None], Lc)[:, 0], True
    return torch.linalg.solve(L, b), True


def ba(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, t0, t1, iterations, lm, ep,
       motion_only, alpha=0.05):
    """slam_ext.ba (geom_kernels.cu:1822-1830): depth_active = 1, no intrinsics, no energy."""
    depth_active = torch.ones(disps.shape[0], dtype=F64)
    return ba_extended(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, depth_active,
                       t0, t1, iterations, lm, ep, motion_only, alpha, False, 1e-6, 1e-6, 1.0, False)


# ----------------------------------------------------------------------------
# Brute-force dense Gauss-Newton (pins the Schur path on tiny graphs)
# ----------------------------------------------------------------------------

def dense_gn_step(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, depth_active,
                  t0, t1, lm, ep, alpha, return_system=False, robust=None):
    """One GN step over [poses in window (D P), all depth pixels of kx (K*HW)] with the explicit
    Jacobian J of every residual row, the same damping of the pose block and the same depth
    prior / eta terms, solved with torch.linalg.solve.  Returns (dx [P,D], dz [K,HW]).

    Independent of the Schur assembly: it forms the full normal equations from an explicit
    stacked Jacobian.  Without intrinsics; depth_active must be all ones.  The pose damping
    is applied to the diagonal of H (the Schur path damps H - S), so the two agree for
    lm = 0."""
    E = ii.shape[0]
    N, H, W = disps.shape
    HW = H * W
    bk = bookkeeping(ii, jj, t0, t1)
    P, kx = bk.P, bk.kx
    K = kx.shape[0]
    T = edge_terms(poses, disps, intrinsics, targets, weights, ii, jj, robust=robust)
    At = T.Adj.transpose(1, 2)
    Jiu = -(At[:, None] @ T.Ju[..., None])[..., 0]
    Jiv = -(At[:, None] @ T.Jv[..., None])[..., 0]
    D = pose_dof(poses)
    nvar = D * P + K * HW
    stereo = (ii == jj)
    kpos = torch.searchsorted(kx, ii)
    rows = []
    rhs = []
    wts = []
    for e in range(E):
        for comp in range(2):
            Jj = T.Ju[e] if comp == 0 else T.Jv[e]
            Ji = Jiu[e] if comp == 0 else Jiv[e]
            Jz = T.Jzu[e] if comp == 0 else T.Jzv[e]
            r = T.ru[e] if comp == 0 else T.rv[e]
            wgt = T.wu[e] if comp == 0 else T.wv[e]
            J = torch.zeros(HW, nvar, dtype=F64)
            a_, b_ = int(ii[e]) - t0, int(jj[e]) - t0
            if not stereo[e]:
                if 0 <= a_ < P:
                    J[:, D * a_:D * a_ + D] += Ji
                if 0 <= b_ < P:
                    J[:, D * b_:D * b_ + D] += Jj
            col0 = D * P + int(kpos[e]) * HW
            J[torch.arange(HW), col0 + torch.arange(HW)] = Jz
            rows.append(J)
            rhs.append(r)
            wts.append(wgt)
    J = torch.cat(rows)
    r = torch.cat(rhs)
    wv = torch.cat(wts)
    Hm = J.T @ (wv[:, None] * J)
    g = J.T @ (wv * r)
    # pose damping as in SparseBlock::solve; depth: + eta + m*alpha, prior residual
    for i in range(D * P):
        Hm[i, i] = Hm[i, i] + ep + lm * Hm[i, i]
    dk = disps.to(F64)[kx].reshape(K, HW)
    dsk = disps_sens.to(F64)[kx].reshape(K, HW)
    m = (dsk > 0).to(F64)
    idx = D * P + torch.arange(K * HW)
    Hm[idx, idx] += (eta.to(F64).reshape(K, HW) + m * alpha).reshape(-1)
    g[idx] -= (m * alpha * (dk - dsk)).reshape(-1)
    x = torch.linalg.solve(Hm, g)
    if return_system:
        return x[:D * P].reshape(P, D), x[D * P:].reshape(K, HW), Hm, g
    return x[:D * P].reshape(P, D), x[D * P:].reshape(K, HW)
