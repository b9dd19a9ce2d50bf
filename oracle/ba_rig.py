"""FP64 restatement of the fused BA on multi-view rigs and Mei cameras (oracle; test-only).

The reference's fused kernel (csrc/slam_ext/geom_kernels.cu) covers
single-view pinhole buffers only; GraphBuffer._fused_ba rejects the rest
(vipe/slam/components/buffer.py:384-410) and sends them to the generic
solver.  This module states the fused step for those buffers: the fused
kernel's conventions (oracle/ba.py: MIN_DEPTH 0.25 with weight and Jacobians
zeroed below it, weight scale 0.001, per-depth-frame Schur rows, the
back-substitution skip of rows whose local pose index is <= 0, damping
ep + lm diag, zeros on a failed factorisation, energy before the update, SE3
retraction) applied to the generic path's multi-view projection model:

  edge expansion        buffer.py:304-362   depth frame d = p V + q of
                                             pose p and view q (flattened
                                             disparities [N V, H, W])
  transform             geom.py:187-298      T = rig[qj]^-1 Gj Gi^-1 rig[qi],
                                             rig = camera-to-rig SE3 per view
  Jacobians             geom.py:265-283      Jj = Jc Adj(rig[qj]^-1),
                                             Ji = -Jj Adj(Gj Gi^-1),
                                             Jc = dproj/dX [X d-block, -hat(X)]
  pinhole / Mei models  vipe/utils/cameras.py:123-217 (pinhole), :220-354 (Mei:
                                             iproj through q = sqrt(1 + (1 -
                                             k1^2) r^2); proj x = fx X /
                                             (Z + k1 |X|) + cx)

Edges within one frame (pi == pj, views qi != qj) have Ji = -Jj exactly
(Gj Gi^-1 = I): their pose columns cancel, so they carry no pose terms and
no Schur rows, and constrain the depth only.  Intrinsics are fixed; poses are
SE3 [N,7].

Pinned by tests/test_oracle_ba_rig.py: with one view, an identity rig and the
pinhole model, ba_rig equals oracle/ba.py's ba_extended; residual Jacobians
(both pose sides, the disparity) against central differences for both camera
models under a non-identity rig; the Mei model against its closed forms (k1 = 0
is the pinhole model, proj(iproj(u, v)) = (u, v)); one step against a
brute-force dense Gauss-Newton solve over every pose and depth unknown.
"""
from __future__ import annotations

import torch

from . import ba as OB
from . import lie
from . import robust as robust_mod

F64 = torch.float64
MIN_DEPTH = OB.MIN_DEPTH
PINHOLE, MEI = 0, 1


def iproj(cam: int, K: torch.Tensor, u: torch.Tensor, v: torch.Tensor):
    """(X0, Y0) of the homogeneous point [X0, Y0, 1, d] of pixel (u, v)
    (cameras.py:131-159 pinhole; :228-251 Mei, whose Z0 is 1 as well)."""
    fx, fy, cx, cy = K[..., 0], K[..., 1], K[..., 2], K[..., 3]
    ub = (u - cx) / fx
    vb = (v - cy) / fy
    if cam == PINHOLE:
        return ub, vb
    k1 = K[..., 4]
    r2 = ub * ub + vb * vb
    q = torch.sqrt(1 + (1 - k1 * k1) * r2)
    factor = (k1 + q) / (1 + r2)
    s = factor / (factor - k1)
    return ub * s, vb * s


def project(cam: int, K: torch.Tensor, X, Y, Z):
    """Pixel (u, v) of the camera-frame point and d(u, v)/d(X, Y, Z) as two
    3-vectors (cameras.py:161-209 pinhole, :283-337 Mei), with the fused
    kernel's rule: Z < MIN_DEPTH gives no projection (inverse 0).  K [..., 4
    or 5] broadcasts against X, Y, Z."""
    fx, fy, cx, cy = K[..., 0], K[..., 1], K[..., 2], K[..., 3]
    valid = Z >= MIN_DEPTH
    if cam == PINHOLE:
        inv = torch.where(valid, 1.0 / torch.where(valid, Z, torch.ones_like(Z)), torch.zeros_like(Z))
        u = fx * X * inv + cx
        v = fy * Y * inv + cy
        o = torch.zeros_like(Z)
        gu = torch.stack([fx * inv, o, -fx * X * inv * inv], -1)
        gv = torch.stack([o, fy * inv, -fy * Y * inv * inv], -1)
        return u, v, gu, gv, valid
    k1 = K[..., 4]
    r = torch.sqrt(X * X + Y * Y + Z * Z)
    rb = Z + k1 * r
    inv = torch.where(valid, 1.0 / torch.where(valid, rb, torch.ones_like(rb)), torch.zeros_like(rb))
    u = fx * X * inv + cx
    v = fy * Y * inv + cy
    den = torch.where(valid, inv * inv / torch.where(valid, r, torch.ones_like(r)), torch.zeros_like(r))
    gu = (fx * den)[..., None] * torch.stack([rb * r - k1 * X * X, -k1 * X * Y, -X * (k1 * Z + r)], -1)
    gv = (fy * den)[..., None] * torch.stack([-k1 * X * Y, rb * r - k1 * Y * Y, -Y * (k1 * Z + r)], -1)
    return u, v, gu, gv, valid


def _se3(P):
    """(R [...,3,3], t [...,3]) of SE3 [..., 7] (t, q xyzw); the quaternion is
    used as stored, as the fused kernel does (geom_kernels.cu:101-114)."""
    return lie.qmat(P[..., 3:7]), P[..., :3]


def _compose(Ra, ta, Rb, tb):
    return Ra @ Rb, (Ra @ tb[..., None])[..., 0] + ta


def _inverse(R, t):
    Rt = R.transpose(-1, -2)
    return Rt, -(Rt @ t[..., None])[..., 0]


def adj(R, t):
    """Adj of SE3 (R, t) in the [tau, phi] ordering: [[R, hat(t) R], [0, R]]."""
    A = torch.zeros(R.shape[:-2] + (6, 6), dtype=R.dtype)
    A[..., :3, :3] = R
    A[..., :3, 3:] = lie.hat(t) @ R
    A[..., 3:, 3:] = R
    return A


def edge_geometry(poses, rig, pi, pj, qi, qj):
    """Per edge: G = Gj Gi^-1 (identity when pi == pj), the camera transform
    T = rig[qj]^-1 G rig[qi], A = Adj(rig[qj]^-1) and Adj(G)."""
    P = poses.to(F64)
    # G as the fused kernel forms it: q = qj conj(qi), t = tj - R(q) ti (relSE3)
    qG = lie.qmul(P[pj, 3:7], lie.qconj(P[pi, 3:7]))
    RG = lie.qmat(qG)
    tG = P[pj, :3] - (RG @ P[pi, :3, None])[..., 0]
    same = (pi == pj)
    if same.any():
        RG = torch.where(same[:, None, None], torch.eye(3, dtype=F64).expand_as(RG), RG)
        tG = torch.where(same[:, None], torch.zeros_like(tG), tG)
    Rr, tr = _se3(rig.to(F64))
    Rqj_inv, tqj_inv = _inverse(Rr[qj], tr[qj])
    RM, tM = _compose(Rqj_inv, tqj_inv, RG, tG)
    RT, tT = _compose(RM, tM, Rr[qi], tr[qi])
    return RT, tT, adj(Rqj_inv, tqj_inv), adj(RG, tG), same


def edge_terms(poses, disps, intr, cam, rig, targets, weights, pi, pj, qi, qj, V, robust=None):
    """Per-pixel terms of every edge: residuals r = target - proj, weights
    (0.001 weight, zero below MIN_DEPTH), Jacobians Jj / Ji [E,HW,2,6] of the
    two poses (left perturbations), Jz [E,HW,2] of the source disparity."""
    E = pi.shape[0]
    NV, H, W = disps.shape
    HW = H * W
    K = intr.to(F64)
    RT, tT, A, AG, same = edge_geometry(poses, rig, pi, pj, qi, qj)
    vv, uu = torch.meshgrid(torch.arange(H, dtype=F64), torch.arange(W, dtype=F64), indexing="ij")
    uu, vv = uu.reshape(HW), vv.reshape(HW)
    Ki = K[qi][:, None, :].expand(E, HW, K.shape[1])
    Kj = K[qj][:, None, :].expand(E, HW, K.shape[1])
    X0, Y0 = iproj(cam, Ki, uu.expand(E, HW), vv.expand(E, HW))
    d = disps.to(F64)[pi * V + qi].reshape(E, HW)
    P0 = torch.stack([X0, Y0, torch.ones_like(X0)], -1)
    X1 = (RT[:, None] @ P0[..., None])[..., 0] + tT[:, None, :] * d[..., None]
    X, Y, Z = X1.unbind(-1)
    u, v, gu, gv, valid = project(cam, Kj, X, Y, Z)
    wgt = weights.to(F64).reshape(E, 2, HW)
    tgt = targets.to(F64).reshape(E, 2, HW)
    wu = torch.where(valid, 0.001 * wgt[:, 0], torch.zeros_like(Z))
    wv = torch.where(valid, 0.001 * wgt[:, 1], torch.zeros_like(Z))
    ru = tgt[:, 0] - u
    rv = tgt[:, 1] - v
    if robust is not None and robust[0] != robust_mod.NONE:
        kind, c, mu, mu_max = robust
        wu = wu * robust_mod.weight(ru, kind, c, mu, mu_max)
        wv = wv * robust_mod.weight(rv, kind, c, mu, mu_max)
    # d(T X)/d(xi) for a left perturbation of T on [X, d]: [d I, -hat(X)]
    Ja = torch.zeros(E, HW, 3, 6, dtype=F64)
    Ja[..., 0, 0] = d
    Ja[..., 1, 1] = d
    Ja[..., 2, 2] = d
    Ja[..., :, 3:] = -lie.hat(X1)
    Jcu = (gu[..., None, :] @ Ja)[..., 0, :]  # [E,HW,6]
    Jcv = (gv[..., None, :] @ Ja)[..., 0, :]
    Jju = (Jcu[..., None, :] @ A[:, None])[..., 0, :]
    Jjv = (Jcv[..., None, :] @ A[:, None])[..., 0, :]
    Jiu = -(Jju[..., None, :] @ AG[:, None])[..., 0, :]
    Jiv = -(Jjv[..., None, :] @ AG[:, None])[..., 0, :]
    Jzu = (gu * tT[:, None, :]).sum(-1)
    Jzv = (gv * tT[:, None, :]).sum(-1)
    return dict(ru=ru, rv=rv, wu=wu, wv=wv, Jju=Jju, Jjv=Jjv, Jiu=Jiu, Jiv=Jiv, Jzu=Jzu, Jzv=Jzv, z=Z,
                same=same, u=u, v=v)


def bookkeeping(pi, qi, V, t0, t1):
    """Depth frames kx = unique of every view of the window's frames and the
    edges' source depth frames (the flattened-buffer form of
    geom_kernels.cu:1648-1652); the edge -> depth-frame slot map."""
    win = torch.arange(t0 * V, t1 * V, dtype=torch.long)
    di = pi * V + qi
    kx = torch.unique(torch.cat([win, di]), sorted=True)
    return kx, torch.searchsorted(kx, di)


def ba_rig(poses, disps, intr, cam, rig, disps_sens, targets, weights, eta, pi, pj, qi, qj, depth_active,
           t0, t1, iterations, lm, ep, motion_only, alpha, compute_energy=False, skip_quirk=True, robust=None):
    """``iterations`` fused Gauss-Newton steps on a multi-view / Mei buffer, in
    place on poses [N,7] and disps [N V, H, W] (intr [V, 4 or 5], rig [V, 7],
    disps_sens [N V, H, W], eta [|kx|, H, W], depth_active [N V]).  Returns
    [dx [P,6], dz [|kx|,HW] or None, energy].  ``skip_quirk=False`` turns the
    back-substitution skip off (for the dense comparison)."""
    pi, pj, qi, qj = (x.long().cpu() for x in (pi, pj, qi, qj))
    E = pi.shape[0]
    NV, H, W = disps.shape
    V = rig.shape[0]
    HW = H * W
    P = t1 - t0
    kx, kpos = bookkeeping(pi, qi, V, t0, t1)
    K = kx.shape[0]
    kpose = kx // V
    dact = depth_active.to(F64).cpu()
    energy = torch.zeros(iterations if compute_energy else 0, dtype=F64)
    dx = dz = None
    inw = lambda k: 0 <= k < P  # noqa: E731
    for itr in range(iterations):
        T = edge_terms(poses, disps, intr, cam, rig, targets, weights, pi, pj, qi, qj, V, robust)
        keep = (~T["same"]).to(F64)[:, None]
        pu, pv = T["wu"] * keep, T["wv"] * keep
        act_e = dact[pi * V + qi][:, None]
        wdu, wdv = T["wu"] * act_e, T["wv"] * act_e
        Cii = wdu * T["Jzu"] ** 2 + wdv * T["Jzv"] ** 2
        bz = wdu * T["ru"] * T["Jzu"] + wdv * T["rv"] * T["Jzv"]
        # E rows of the two pose sides (zero for edges within one frame)
        cu, cv = (act_e * wdu * T["Jzu"] * keep)[..., None], (act_e * wdv * T["Jzv"] * keep)[..., None]
        Eij = cu * T["Jju"] + cv * T["Jjv"]
        Eii = cu * T["Jiu"] + cv * T["Jiv"]
        Jxu = torch.cat([T["Jiu"], T["Jju"]], -1)
        Jxv = torch.cat([T["Jiv"], T["Jjv"]], -1)
        Hx = torch.einsum("ep,epa,epb->eab", pu, Jxu, Jxu) + torch.einsum("ep,epa,epb->eab", pv, Jxv, Jxv)
        vx = torch.einsum("ep,epa->ea", pu * T["ru"], Jxu) + torch.einsum("ep,epa->ea", pv * T["rv"], Jxv)
        A = torch.zeros(6 * P, 6 * P, dtype=F64)
        b = torch.zeros(6 * P, dtype=F64)
        for e in range(E):
            a_, b_ = int(pi[e]) - t0, int(pj[e]) - t0
            for r_, c_, blk in ((a_, a_, Hx[e, :6, :6]), (a_, b_, Hx[e, :6, 6:]), (b_, a_, Hx[e, 6:, :6]),
                                (b_, b_, Hx[e, 6:, 6:])):
                if inw(r_) and inw(c_):
                    A[6 * r_:6 * r_ + 6, 6 * c_:6 * c_ + 6] += blk
            if inw(a_):
                b[6 * a_:6 * a_ + 6] += vx[e, :6]
            if inw(b_):
                b[6 * b_:6 * b_ + 6] += vx[e, 6:]
        en = (T["wu"] * T["ru"] ** 2 + T["wv"] * T["rv"] ** 2).sum()
        if not motion_only:
            active = dact[kx][:, None]
            dk = disps.to(F64)[kx].reshape(K, HW)
            dsk = disps_sens.to(F64)[kx].reshape(K, HW)
            m = (dsk > 0).to(F64) * active
            en = en + (m * alpha * (dk - dsk) ** 2).sum()
            Csum = torch.zeros(K, HW, dtype=F64).index_add_(0, kpos, Cii)
            bsum = torch.zeros(K, HW, dtype=F64).index_add_(0, kpos, bz)
            C = active * (Csum + eta.to(F64).reshape(K, HW) + m * alpha) + (1 - active)
            w = active * (bsum - m * alpha * (dk - dsk))
            Q = 1.0 / C
            # rows of depth frame k: its pose's summed i-side row, then each edge's j-side row
            rows = []  # (depth slot, pose, E row [HW,6])
            for k in range(K):
                sel = torch.nonzero(kpos == k).flatten()
                if t0 <= int(kpose[k]) < t1:
                    rows.append((k, int(kpose[k]) - t0, Eii[sel].sum(0) if len(sel) else torch.zeros(HW, 6,
                                                                                                  dtype=F64)))
                for e in sel.tolist():
                    rows.append((k, int(pj[e]) - t0, Eij[e]))
            S = torch.zeros(6 * P, 6 * P, dtype=F64)
            vS = torch.zeros(6 * P, dtype=F64)
            for ka, pa, Ea in rows:
                if not inw(pa):
                    continue
                vS[6 * pa:6 * pa + 6] += (Ea * (Q[ka] * w[ka])[:, None]).sum(0)
                for kb, pb, Eb in rows:
                    if kb != ka or not inw(pb):
                        continue
                    S[6 * pa:6 * pa + 6, 6 * pb:6 * pb + 6] += (Ea * Q[ka][:, None]).T @ Eb
            Ared, bred = A - S, b - vS
        else:
            Ared, bred = A, b
        if compute_energy:
            energy[itr] = en
        x, _ = OB.damped_solve(Ared, bred, 6 * P, lm, ep, 0.0, 0.0, False)
        dxv = x.reshape(P, 6)
        if not motion_only:
            lo = 1 if skip_quirk else 0
            dw = torch.zeros(K, HW, dtype=F64)
            for ka, pa, Ea in rows:
                if lo <= pa < P:
                    dw[ka] += Ea @ dxv[pa]
            dzv = Q * (w - dw)
        poses[t0:t1] = lie.droid_retr_se3(dxv, poses.to(F64)[t0:t1]).to(poses.dtype)
        if not motion_only:
            flat = disps.view(NV, HW)
            flat[kx] = (flat[kx].to(F64) + dzv).to(disps.dtype)
            dz = dzv
        dx = dxv
    return [dx, dz, energy]


def dense_gn_step_rig(poses, disps, intr, cam, rig, disps_sens, targets, weights, eta, pi, pj, qi, qj, t0, t1, lm,
                      ep, alpha):
    """One Gauss-Newton step over [the window's poses (6 P), every pixel of the
    depth frames kx] from the explicitly stacked Jacobian of every residual
    row, with ba_rig's weights, depth prior, eta and pose damping (applied to
    the diagonal of H, so the two agree at lm = 0), solved densely.  Depth
    active everywhere.  Returns (dx [P,6], dz [K,HW])."""
    pi, pj, qi, qj = (x.long().cpu() for x in (pi, pj, qi, qj))
    E = pi.shape[0]
    NV, H, W = disps.shape
    V = rig.shape[0]
    HW = H * W
    P = t1 - t0
    kx, kpos = bookkeeping(pi, qi, V, t0, t1)
    K = kx.shape[0]
    T = edge_terms(poses, disps, intr, cam, rig, targets, weights, pi, pj, qi, qj, V)
    nvar = 6 * P + K * HW
    rows, rhs, wts = [], [], []
    for e in range(E):
        for comp in ("u", "v"):
            J = torch.zeros(HW, nvar, dtype=F64)
            a_, b_ = int(pi[e]) - t0, int(pj[e]) - t0
            # an edge within one frame puts Ji + Jj = 0 on its pose: written out, not skipped
            if 0 <= a_ < P:
                J[:, 6 * a_:6 * a_ + 6] += T["Ji" + comp][e]
            if 0 <= b_ < P:
                J[:, 6 * b_:6 * b_ + 6] += T["Jj" + comp][e]
            col0 = 6 * P + int(kpos[e]) * HW
            J[torch.arange(HW), col0 + torch.arange(HW)] = T["Jz" + comp][e]
            rows.append(J)
            rhs.append(T["r" + comp][e])
            wts.append(T["w" + comp][e])
    J = torch.cat(rows)
    r = torch.cat(rhs)
    wv = torch.cat(wts)
    Hm = J.T @ (wv[:, None] * J)
    g = J.T @ (wv * r)
    for i in range(6 * P):
        Hm[i, i] = Hm[i, i] + ep + lm * Hm[i, i]
    dk = disps.to(F64)[kx].reshape(K, HW)
    dsk = disps_sens.to(F64)[kx].reshape(K, HW)
    m = (dsk > 0).to(F64)
    idx = 6 * P + torch.arange(K * HW)
    Hm[idx, idx] += (eta.to(F64).reshape(K, HW) + m * alpha).reshape(-1)
    g[idx] -= (m * alpha * (dk - dsk)).reshape(-1)
    x = torch.linalg.solve(Hm, g)
    return x[:6 * P].reshape(P, 6), x[6 * P:].reshape(K, HW)
