"""Seeded synthetic frame graphs for the SLAM inner loop.

One generator, CPU ``torch.Generator`` so inputs are bit-identical on every
machine.  Value distributions (per the workload definitions this package is
measured on):

* poses: SE3 ``exp([t ~ N(0, 0.1), w ~ N(0, 0.05 rad)])`` per keyframe
  (Sim3 adds ``log s ~ N(0, 0.02)``), stored ``[tx,ty,tz,qx,qy,qz,qw(,s)]``;
* inverse depth ``U[0.2, 1.0]``; pinhole intrinsics at 1/8 resolution;
* features ``N(0,1)/sqrt(C)`` float32, ``[N, C, H, W]``;
* BA targets = reprojection of every edge + ``N(0, 0.5 px)``, weights ``U[0,1]``.

Edge lists: window edges ``(i, i+d)`` ordered by ``|d|`` then ``i`` then sign,
truncated to ``n_edges - n_random``, followed by ``n_random`` long-range edges
(source round-robin over frames, target uniform with ``|i-j| > window``,
no duplicates).  For 8 frames / 24 edges / window 2 / 4 random this keeps all
14 ``|d|=1`` edges and 6 of the ``|d|=2`` edges; for 64 frames / 384 edges /
window 3 / 2 random per frame it keeps every ``|d|<=2`` edge, 6 ``|d|=3`` edges
and 128 random edges, so every frame has out-degree 4..7.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch


@dataclass
class FrameGraph:
    poses: torch.Tensor        # [N, 7] (SE3) or [N, 8] (Sim3) float32
    disps: torch.Tensor        # [N, H, W] float32
    disps_sens: torch.Tensor   # [N, H, W] float32 (zeros: no depth sensor)
    intrinsics: torch.Tensor   # [4] float32 (fx, fy, cx, cy)
    ii: torch.Tensor           # [E] int64
    jj: torch.Tensor           # [E] int64
    targets: torch.Tensor      # [E, 2, H, W] float32
    weights: torch.Tensor      # [E, 2, H, W] float32
    eta: torch.Tensor          # [|kx|, H, W] float32 for the window [t0, t1)
    fmaps: torch.Tensor        # [N, C, H, W] float32
    coords: torch.Tensor       # [E, H, W, 2] float32 reprojected pixel coordinates
    t0: int
    t1: int


# name -> (frames, edges, H, W, fx, fy, cx, cy, window, n_random, group)
CONFIGS = {
    # 8 keyframes, 24 edges, 48x64 (1/8 of 384x512)
    "kf8_e24_48x64": dict(n_frames=8, n_edges=24, H=48, W=64, intr=(60.0, 60.0, 32.0, 24.0),
                          window=2, n_random=4, sim3=False),
    # 64 keyframes, 384 edges, 64x80 (1/8 of 512x640)
    "kf64_e384_64x80": dict(n_frames=64, n_edges=384, H=64, W=80, intr=(80.0, 80.0, 40.0, 32.0),
                            window=3, n_random=128, sim3=False),
    # Sim3 poses at the 8-keyframe size
    "kf8_e24_48x64_sim3": dict(n_frames=8, n_edges=24, H=48, W=64, intr=(60.0, 60.0, 32.0, 24.0),
                               window=2, n_random=4, sim3=True),
    # Sim3 poses at the 64-keyframe size
    "kf64_e384_64x80_sim3": dict(n_frames=64, n_edges=384, H=64, W=80, intr=(80.0, 80.0, 40.0, 32.0),
                                 window=3, n_random=128, sim3=True),
    # 32 keyframes, 192 edges, 48x64
    "kf32_e192_48x64": dict(n_frames=32, n_edges=192, H=48, W=64, intr=(60.0, 60.0, 32.0, 24.0),
                            window=3, n_random=64, sim3=False),
}


def make_edges(n_frames: int, n_edges: int, window: int, n_random: int, gen: torch.Generator):
    cands = []
    for d in range(1, window + 1):
        for i in range(n_frames):
            for sgn in (1, -1):
                j = i + sgn * d
                if 0 <= j < n_frames:
                    cands.append((i, j))
    n_win = n_edges - n_random
    if n_win > len(cands):
        raise ValueError("not enough window edges for the requested edge count")
    edges = cands[:n_win]
    seen = set(edges)
    k = 0
    guard = 0
    while len(edges) < n_edges:
        i = k % n_frames
        j = int(torch.randint(0, n_frames, (1,), generator=gen))
        guard += 1
        if guard > 100000:
            raise ValueError("cannot place random edges")
        if abs(i - j) <= window or (i, j) in seen:
            continue
        edges.append((i, j))
        seen.add((i, j))
        k += 1
    ii = torch.tensor([e[0] for e in edges], dtype=torch.long)
    jj = torch.tensor([e[1] for e in edges], dtype=torch.long)
    return ii, jj


def _quat_exp(phi: torch.Tensor) -> torch.Tensor:
    th = phi.norm(dim=-1, keepdim=True)
    ths = torch.where(th > 1e-12, th, torch.ones_like(th))
    imag = torch.where(th > 1e-12, torch.sin(0.5 * ths) / ths, 0.5 * torch.ones_like(th))
    return torch.cat([imag * phi, torch.cos(0.5 * th)], -1)


def _qrot(q, p):
    v, w = q[..., :3], q[..., 3:]
    uv = 2.0 * torch.cross(v, p, dim=-1)
    return p + w * uv + torch.cross(v, uv, dim=-1)


def _qmul(a, b):
    av, aw = a[..., :3], a[..., 3:]
    bv, bw = b[..., :3], b[..., 3:]
    return torch.cat([aw * bv + bw * av + torch.cross(av, bv, dim=-1), aw * bw - (av * bv).sum(-1, keepdim=True)], -1)


def reproject(poses, disps, intrinsics, ii, jj):
    """Pixel coordinates of frame ii's pixels in frame jj (float64 maths):
    X = [(u-cx)/fx, (v-cy)/fy, 1, d]; Xj = Tj Ti^-1 X; returns [E, H, W, 2] and z."""
    P = poses.double()
    N, H, W = disps.shape
    fx, fy, cx, cy = [float(v) for v in intrinsics]
    qi, qj = P[ii, 3:7], P[jj, 3:7]
    ti, tj = P[ii, :3], P[jj, :3]
    si = P[ii, 7] if P.shape[1] == 8 else torch.ones(ii.shape[0], dtype=torch.float64)
    sj = P[jj, 7] if P.shape[1] == 8 else torch.ones(ii.shape[0], dtype=torch.float64)
    qinv = torch.cat([-qi[:, :3], qi[:, 3:]], -1)
    qij = _qmul(qj, qinv)
    sij = sj / si
    tij = tj - sij[:, None] * _qrot(qij, ti)
    vv, uu = torch.meshgrid(torch.arange(H, dtype=torch.float64), torch.arange(W, dtype=torch.float64), indexing="ij")
    X = torch.stack([(uu - cx) / fx, (vv - cy) / fy, torch.ones_like(uu)], -1)  # [H,W,3]
    d = disps.double()[ii]  # [E,H,W]
    Xr = sij[:, None, None, None] * _qrot(qij[:, None, None, :].expand(-1, H, W, -1), X.expand(ii.shape[0], H, W, 3))
    Xj = Xr + tij[:, None, None, :] * d[..., None]
    z = Xj[..., 2]
    zs = torch.where(z.abs() > 1e-9, z, torch.full_like(z, 1e-9))
    u = fx * Xj[..., 0] / zs + cx
    v = fy * Xj[..., 1] / zs + cy
    return torch.stack([u, v], -1), z


def make_graph(name: str = "kf8_e24_48x64", seed: int = 0, C: int = 128, with_fmaps: bool = True,
               t0: int = 1, t1: int | None = None) -> FrameGraph:
    return make_graph_from(CONFIGS[name], seed, C, with_fmaps, t0, t1)


# Batched uneven workload: window sizes, out-degrees and resolutions of the clips
BATCH_WINDOWS = (8, 16, 24, 48, 64, 96)
BATCH_DEGREES = (2, 12)  # inclusive range of edges per frame
# (H, W, fx, fy, cx, cy): 1/8 of 384x512, 512x640, and a 2:1 grid of 320x640
BATCH_RESOLUTIONS = ((48, 64, 60.0, 60.0, 32.0, 24.0), (64, 80, 80.0, 80.0, 40.0, 32.0),
                     (40, 80, 40.0, 40.0, 40.0, 20.0))


def batch_configs(n_clips: int = 16, seed: int = 0, windows=BATCH_WINDOWS, resolutions=BATCH_RESOLUTIONS):
    """Per-clip graph configurations of the batched uneven workload: the window
    size is drawn from ``windows``, the out-degree uniformly from 2..12 (at most
    frames-1), the resolution from ``resolutions``; a quarter of each frame's
    edges (rounded down) are long-range, the rest a +-window band."""
    gen = torch.Generator().manual_seed(seed)
    out = []
    for _ in range(n_clips):
        n = windows[int(torch.randint(len(windows), (1,), generator=gen))]
        deg = int(torch.randint(BATCH_DEGREES[0], BATCH_DEGREES[1] + 1, (1,), generator=gen))
        deg = min(deg, n - 1)
        H, W, fx, fy, cx, cy = resolutions[int(torch.randint(len(resolutions), (1,), generator=gen))]
        n_rand_pf = deg // 4
        window = max(1, (deg - n_rand_pf + 1) // 2)
        # every source frame must have n_rand_pf targets farther than the window
        worst = min(max(0, i - window) + max(0, n - 1 - i - window) for i in range(n))
        if worst < n_rand_pf:
            n_rand_pf = 0
            window = min(max(1, (deg + 1) // 2), n - 1)
        band = deg - n_rand_pf
        n_cands = sum(2 * (n - d) for d in range(1, window + 1))
        n_win = min(n * band, n_cands)
        n_random = n * n_rand_pf
        out.append(dict(n_frames=n, n_edges=n_win + n_random, H=H, W=W, intr=(fx, fy, cx, cy), window=window,
                        n_random=n_random, sim3=False))
    return out


def make_batch(n_clips: int = 16, seed: int = 0, windows=BATCH_WINDOWS, resolutions=BATCH_RESOLUTIONS):
    """Independent clips of the batched uneven workload (no feature maps),
    clip c seeded with ``seed * 1000 + c``; each optimises poses [1, N)."""
    return [make_graph_from(cfg, seed * 1000 + c, with_fmaps=False)
            for c, cfg in enumerate(batch_configs(n_clips, seed, windows, resolutions))]


def make_graph_from(cfg: dict, seed: int = 0, C: int = 128, with_fmaps: bool = True, t0: int = 1,
                    t1: int | None = None) -> FrameGraph:
    gen = torch.Generator().manual_seed(seed)
    N, E, H, W = cfg["n_frames"], cfg["n_edges"], cfg["H"], cfg["W"]
    t1 = N if t1 is None else t1
    ii, jj = make_edges(N, E, cfg["window"], cfg["n_random"], gen)
    tau = torch.randn(N, 3, generator=gen, dtype=torch.float64) * 0.1
    phi = torch.randn(N, 3, generator=gen, dtype=torch.float64) * 0.05
    q = _quat_exp(phi)
    # SE3 exp translation: V(phi) tau
    th = phi.norm(dim=-1, keepdim=True)
    a = (1 - torch.cos(th)) / (th * th)
    b = (th - torch.sin(th)) / (th ** 3)
    c1 = torch.cross(phi, tau, dim=-1)
    t = tau + a * c1 + b * torch.cross(phi, c1, dim=-1)
    if cfg["sim3"]:
        s = torch.exp(torch.randn(N, 1, generator=gen, dtype=torch.float64) * 0.02)
        poses = torch.cat([t, q, s], -1)
    else:
        poses = torch.cat([t, q], -1)
    disps = 0.2 + 0.8 * torch.rand(N, H, W, generator=gen, dtype=torch.float64)
    intr = torch.tensor(cfg["intr"], dtype=torch.float64)
    coords, _ = reproject(poses, disps, intr, ii, jj)
    noise = torch.randn(E, H, W, 2, generator=gen, dtype=torch.float64) * 0.5
    targets = (coords + noise).permute(0, 3, 1, 2).contiguous()
    weights = torch.rand(E, 2, H, W, generator=gen, dtype=torch.float64)
    kx = torch.unique(torch.cat([torch.arange(t0, t1), ii]))
    eta = 0.2 * (0.05 + 0.1 * torch.rand(kx.shape[0], H, W, generator=gen, dtype=torch.float64)) + 1e-7
    if with_fmaps:
        fmaps = torch.randn(N, C, H, W, generator=gen, dtype=torch.float32) / math.sqrt(C)
    else:
        fmaps = torch.zeros(0)
    return FrameGraph(
        poses=poses.float().contiguous(), disps=disps.float().contiguous(),
        disps_sens=torch.zeros(N, H, W), intrinsics=intr.float(), ii=ii, jj=jj,
        targets=targets.float().contiguous(), weights=weights.float().contiguous(),
        eta=eta.float().contiguous(), fmaps=fmaps, coords=coords.float().contiguous(), t0=t0, t1=t1,
    )


@dataclass
class RigGraph:
    poses: torch.Tensor        # [N, 7] SE3
    disps: torch.Tensor        # [N V, H, W] (the flattened [N, V, H, W] buffer)
    disps_sens: torch.Tensor   # [N V, H, W]
    intrinsics: torch.Tensor   # [V, 4] pinhole or [V, 5] Mei (fx fy cx cy k1)
    rig: torch.Tensor          # [V, 7] camera-to-rig SE3 of each view
    camera: str                # "pinhole" or "mei"
    pi: torch.Tensor           # [M] edge source pose
    pj: torch.Tensor           # [M] edge target pose
    qi: torch.Tensor           # [M] edge source view
    qj: torch.Tensor           # [M] edge target view
    targets: torch.Tensor      # [M, 2, H, W]
    weights: torch.Tensor      # [M, 2, H, W]
    eta: torch.Tensor          # [|kx|, H, W]
    t0: int
    t1: int


def _rig_project(cam, K, Xc):
    """Pixel coordinates of camera-frame points Xc [..., 3] (float64): pinhole
    fx X / Z + cx, or Mei fx X / (Z + k1 |X|) + cx."""
    X, Y, Z = Xc.unbind(-1)
    den = Z + (K[..., 4] * Xc.norm(dim=-1) if cam == "mei" else 0.0)
    den = torch.where(den.abs() > 1e-9, den, torch.full_like(den, 1e-9))
    return torch.stack([K[..., 0] * X / den + K[..., 2], K[..., 1] * Y / den + K[..., 3]], -1)


def _rig_ray(cam, K, u, v):
    """The unit-depth ray [X0, Y0, 1] of pixel (u, v) (float64)."""
    ub, vb = (u - K[..., 2]) / K[..., 0], (v - K[..., 3]) / K[..., 1]
    if cam == "mei":
        k1 = K[..., 4]
        r2 = ub * ub + vb * vb
        q = torch.sqrt(1 + (1 - k1 * k1) * r2)
        f = (k1 + q) / (1 + r2)
        ub, vb = ub * f / (f - k1), vb * f / (f - k1)
    return torch.stack([ub, vb, torch.ones_like(ub)], -1)


def make_rig_graph(name: str = "kf64_e384_64x80", camera: str = "mei", views: int = 2, seed: int = 0,
                   t0: int = 1, noise: float = 0.5) -> RigGraph:
    """A rig version of the frame graph ``name``: the same poses and edges, seen
    by ``views`` cameras (view q a 0.1 baseline step to the right of view q-1,
    rotated 0.05 rad about y), every edge once per view (q -> q) and, for each
    frame of the window, its cross-view edges q -> q + 1 mod V.  Mei
    intrinsics keep the pinhole focal length at the centre (fx (1 + k1), k1
    0.6 and 0.5).  Targets are the projections + N(0, noise px), weights U[0, 1]."""
    cfg = CONFIGS[name]
    g = make_graph_from(cfg, seed, with_fmaps=False, t0=t0)
    gen = torch.Generator().manual_seed(seed + 1)
    N, H, W = g.disps.shape
    V = views
    fx, fy, cx, cy = cfg["intr"]
    Ks = []
    for q in range(V):
        k1 = 0.6 - 0.1 * (q % 2)
        if camera == "mei":
            Ks.append([fx * (1 + k1), fy * (1 + k1), cx, cy, k1])
        else:
            Ks.append([fx, fy, cx, cy])
    K = torch.tensor(Ks, dtype=torch.float64)
    rig = torch.zeros(V, 7, dtype=torch.float64)
    for q in range(V):
        rig[q, 0] = 0.1 * q
        rig[q, 3:7] = _quat_exp(torch.tensor([0.0, 0.05 * q, 0.0], dtype=torch.float64))
    keep = g.ii != g.jj
    ii, jj = g.ii[keep], g.jj[keep]
    pi = ii.repeat_interleave(V)
    pj = jj.repeat_interleave(V)
    qi = torch.arange(V).repeat(ii.shape[0])
    qj = qi.clone()
    if V > 1:
        fr = torch.arange(t0, N)
        pi = torch.cat([pi, fr.repeat_interleave(V)])
        pj = torch.cat([pj, fr.repeat_interleave(V)])
        qi = torch.cat([qi, torch.arange(V).repeat(fr.shape[0])])
        qj = torch.cat([qj, (torch.arange(V) + 1).remainder(V).repeat(fr.shape[0])])
    M = pi.shape[0]
    disps = (0.2 + 0.8 * torch.rand(N * V, H, W, generator=gen, dtype=torch.float64))
    P = g.poses.double()
    vv, uu = torch.meshgrid(torch.arange(H, dtype=torch.float64), torch.arange(W, dtype=torch.float64), indexing="ij")
    targets = torch.empty(M, 2, H, W, dtype=torch.float64)
    for e0 in range(0, M, 256):
        sl = slice(e0, min(M, e0 + 256))
        a, b, c, d = pi[sl], pj[sl], qi[sl], qj[sl]
        ray = _rig_ray(camera, K[c][:, None, None, :], uu, vv)                    # [m,H,W,3]
        X = _qrot(rig[c, 3:7][:, None, None, :], ray) + rig[c, :3][:, None, None, :] * disps[a * V + c][..., None]
        # rig -> world -> rig of frame b -> camera d (poses are world-to-rig)
        qa_inv = torch.cat([-P[a, 3:6], P[a, 6:7]], -1)[:, None, None, :]
        X = _qrot(qa_inv, X - P[a, :3][:, None, None, :] * disps[a * V + c][..., None])
        X = _qrot(P[b, 3:7][:, None, None, :], X) + P[b, :3][:, None, None, :] * disps[a * V + c][..., None]
        qd_inv = torch.cat([-rig[d, 3:6], rig[d, 6:7]], -1)[:, None, None, :]
        X = _qrot(qd_inv, X - rig[d, :3][:, None, None, :] * disps[a * V + c][..., None])
        uv = _rig_project(camera, K[d][:, None, None, :], X)
        targets[sl] = uv.permute(0, 3, 1, 2)
    targets += torch.randn(M, 2, H, W, generator=gen, dtype=torch.float64) * noise
    weights = torch.rand(M, 2, H, W, generator=gen)
    kx = torch.unique(torch.cat([torch.arange(t0 * V, N * V), pi * V + qi]))
    eta = 0.2 * (0.05 + 0.1 * torch.rand(kx.shape[0], H, W, generator=gen, dtype=torch.float64)) + 1e-7
    return RigGraph(poses=g.poses.contiguous(), disps=disps.float().contiguous(),
                    disps_sens=torch.zeros(N * V, H, W), intrinsics=K.float().contiguous(),
                    rig=rig.float().contiguous(), camera=camera, pi=pi, pj=pj, qi=qi, qj=qj,
                    targets=targets.float().contiguous(), weights=weights.contiguous(), eta=eta.float().contiguous(),
                    t0=t0, t1=N)
