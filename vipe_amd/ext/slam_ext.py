"""Drop-in for ``vipe.ext.slam_ext`` bundle adjustment (csrc/slam_ext/slam.cpp:41-51).

    ba(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj,
       t0, t1, iterations, lm, ep, motion_only, alpha=0.05) -> [dx, dz, energy]
    ba_extended(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj,
                depth_active, t0, t1, iterations, lm, ep, motion_only, alpha,
                optimize_intrinsics, intrinsics_lm, intrinsics_ep, intrinsics_scale,
                compute_energy=False) -> [dx, dz, energy]

``poses`` [N,7], ``disps`` [N,H,W] and (when optimising intrinsics)
``intrinsics`` [4] are updated in place; ``dx`` is [t1-t0, 6] and ``dz``
[|kx|, H*W] from the last iteration (``None`` when motion only or when no
iteration ran, as the reference returns undefined tensors), ``energy`` has
``iterations`` entries when requested and is empty otherwise
(geom_kernels.cu:1626-1847).  Beyond the reference, ``poses`` may also be
Sim3 [N,8] (t, q, s): the same step with 7-dof pose blocks, ``dx`` [t1-t0, 7]
and the retraction of lietorch's Sim3.  Non-contiguous inputs raise
``RuntimeError("<name> must be contiguous")`` (geom_kernels.cu:30-31,
1634-1642).

The whole solve runs on the device: per-edge Jacobians and Hessian blocks,
the Schur complement, the FP64 Cholesky of the reduced camera system and the
retractions.  The edge list is copied to the host to build (or fetch from a
cache) the index bookkeeping, except when the call passes the same unmodified
``ii`` / ``jj`` tensors as the previous call.
"""
from __future__ import annotations

import ctypes as C
import weakref
from collections import OrderedDict

import torch

from .. import _lib

_PLANS: "OrderedDict[tuple, _Plan]" = OrderedDict()
_PLAN_CACHE_SIZE = 32


class _Plan:
    def __init__(self, ii_cpu, jj_cpu, N, H, W, t0, t1, pixel_tile=0):
        lib = _lib.load()
        ii_c = ii_cpu.to(torch.int64).contiguous()
        jj_c = jj_cpu.to(torch.int64).contiguous()
        h = lib.vh_ba_plan_create_tiled(C.c_void_p(ii_c.data_ptr()), C.c_void_p(jj_c.data_ptr()),
                                        int(ii_c.numel()), int(N), int(H), int(W), int(t0), int(t1),
                                        int(pixel_tile))
        if not h:
            raise RuntimeError(f"slam_ext.ba: {_lib.last_error()}")
        self.h = C.c_void_p(h)
        self.K = lib.vh_ba_plan_num_depth_frames(self.h)
        self.ws_bytes = int(lib.vh_ba_workspace_size(self.h))
        self.pixel_tile = int(lib.vh_ba_plan_pixel_tile(self.h))
        kx = torch.empty(self.K, dtype=torch.int64)
        lib.vh_ba_plan_depth_frames(self.h, C.c_void_p(kx.data_ptr()))
        self.kx = kx

    def __del__(self):
        try:
            _lib.load().vh_ba_plan_destroy(self.h)
        except Exception:
            pass


_LAST = {"ii": None, "jj": None, "iv": -1, "jv": -1, "key": None}


def get_plan(ii, jj, N, H, W, t0, t1, pixel_tile=0) -> _Plan:
    """The (cached) plan of an edge list; pixel_tile 0 lets the library choose
    the frame kernel's pixels per pass, 64 / 128 / 256 force it."""
    # Fast path: the same (unmodified) edge tensors as the previous call reuse the
    # plan without copying them to the host.
    dims = (int(N), int(H), int(W), int(t0), int(t1), int(pixel_tile))
    if (_LAST["ii"] is not None and _LAST["ii"]() is ii and _LAST["jj"]() is jj and ii._version == _LAST["iv"]
            and jj._version == _LAST["jv"] and _LAST["key"] is not None and _LAST["key"][2:] == dims):
        plan = _PLANS.get(_LAST["key"])
        if plan is not None:
            return plan
    ii_cpu = ii.detach().to("cpu", torch.int64)
    jj_cpu = jj.detach().to("cpu", torch.int64)
    key = (ii_cpu.numpy().tobytes(), jj_cpu.numpy().tobytes()) + dims
    plan = _PLANS.get(key)
    if plan is None:
        plan = _Plan(ii_cpu, jj_cpu, N, H, W, t0, t1, pixel_tile)
        _PLANS[key] = plan
        while len(_PLANS) > _PLAN_CACHE_SIZE:
            _PLANS.popitem(last=False)
    else:
        _PLANS.move_to_end(key)
    try:
        _LAST.update(ii=weakref.ref(ii), jj=weakref.ref(jj), iv=ii._version, jv=jj._version, key=key)
    except TypeError:
        _LAST.update(ii=None, jj=None, key=None)
    return plan


def _check(**named):
    for k, v in named.items():
        if not v.is_contiguous():
            raise RuntimeError(f"{k} must be contiguous")


def _f32(name, t):
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be float32 (got {t.dtype})")


def ba_extended(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, depth_active, t0, t1,
                iterations, lm, ep, motion_only, alpha, optimize_intrinsics, intrinsics_lm, intrinsics_ep,
                intrinsics_scale, compute_energy=False):
    dx, dz, energy, _, _ = _ba_extended(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj,
                                        depth_active, t0, t1, iterations, lm, ep, motion_only, alpha,
                                        optimize_intrinsics, intrinsics_lm, intrinsics_ep, intrinsics_scale,
                                        compute_energy)
    return [dx, dz, energy]


def ba_extended_robust(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, depth_active, t0, t1,
                       iterations, lm, ep, motion_only, alpha, optimize_intrinsics, intrinsics_lm, intrinsics_ep,
                       intrinsics_scale, compute_energy=False, robust=None):
    """ba_extended with a robust kernel on the flow residuals: ``robust`` is
    (type, threshold, mu, mu_max) as given by ``RobustKernel.fused_args()``
    (vipe_amd.ba_kernel), or None for plain least squares.  This package's
    extension: the reference's fused BA rejects robust kernels."""
    dx, dz, energy, _, _ = _ba_extended(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj,
                                        depth_active, t0, t1, iterations, lm, ep, motion_only, alpha,
                                        optimize_intrinsics, intrinsics_lm, intrinsics_ep, intrinsics_scale,
                                        compute_energy, robust)
    return [dx, dz, energy]


def _ba_extended(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, depth_active, t0, t1,
                 iterations, lm, ep, motion_only, alpha, optimize_intrinsics, intrinsics_lm, intrinsics_ep,
                 intrinsics_scale, compute_energy=False, robust=None):
    """ba_extended that also returns the plan and the workspace of the run."""
    _check(targets=targets, weights=weights, poses=poses, disps=disps, intrinsics=intrinsics,
           disps_sens=disps_sens, ii=ii, jj=jj, depth_active=depth_active)
    for name, t in (("poses", poses), ("disps", disps), ("intrinsics", intrinsics), ("disps_sens", disps_sens),
                    ("targets", targets), ("weights", weights), ("depth_active", depth_active)):
        _f32(name, t)
        if t.device.type != "cuda":
            raise RuntimeError(f"slam_ext.ba runs on the GPU; {name} is on {t.device}")
    N, H, W = disps.shape
    E = ii.shape[0]
    t0, t1, iterations = int(t0), int(t1), int(iterations)
    if poses.shape not in ((N, 7), (N, 8)):
        raise RuntimeError(f"poses must be [N,7] SE3 or [N,8] Sim3 (got {tuple(poses.shape)})")
    sim3 = poses.shape[1] == 8
    if targets.shape != (E, 2, H, W) or weights.shape != (E, 2, H, W):
        raise RuntimeError("targets/weights must be [E,2,H,W]")
    plan = get_plan(ii, jj, N, H, W, t0, t1)
    P = t1 - t0
    dev = poses.device
    eta_c = eta.contiguous().to(torch.float32) if not motion_only else eta
    if not motion_only and eta_c.numel() != plan.K * H * W:
        raise RuntimeError(f"eta must hold |kx| = {plan.K} maps of {H}x{W}")
    energy = torch.empty((iterations if compute_energy else 0,), dtype=torch.float32, device=dev)
    if iterations <= 0:
        return None, None, energy, plan, None
    dx = torch.empty((P, 7 if sim3 else 6), dtype=torch.float32, device=dev)
    dz = None if motion_only else torch.empty((plan.K, H * W), dtype=torch.float32, device=dev)
    ws = torch.empty((max(plan.ws_bytes, 1),), dtype=torch.uint8, device=dev)
    flags = (1 if motion_only else 0) | (2 if optimize_intrinsics else 0) | (4 if compute_energy else 0) | \
        (8 if sim3 else 0)
    lib = _lib.load()
    rk = (0, 0.0, 0.0, 0.0) if robust is None else tuple(robust)
    with torch.cuda.device(dev):
        if rk[0] == 0:
            rc = lib.vh_ba_run(plan.h, _lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics), _lib.ptr(disps_sens),
                               _lib.ptr(targets), _lib.ptr(weights), None if motion_only else _lib.ptr(eta_c),
                               _lib.ptr(depth_active), iterations, float(lm), float(ep), float(alpha),
                               float(intrinsics_lm), float(intrinsics_ep), float(intrinsics_scale), flags,
                               _lib.ptr(dx), _lib.ptr(dz), _lib.ptr(energy) if compute_energy else None,
                               _lib.ptr(ws), _lib.stream_of(poses))
        else:
            rc = lib.vh_ba_run_robust(plan.h, _lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics),
                                      _lib.ptr(disps_sens), _lib.ptr(targets), _lib.ptr(weights),
                                      None if motion_only else _lib.ptr(eta_c), _lib.ptr(depth_active), iterations,
                                      float(lm), float(ep), float(alpha), float(intrinsics_lm), float(intrinsics_ep),
                                      float(intrinsics_scale), flags, int(rk[0]), float(rk[1]), float(rk[2]),
                                      float(rk[3]), _lib.ptr(dx), _lib.ptr(dz),
                                      _lib.ptr(energy) if compute_energy else None, _lib.ptr(ws),
                                      _lib.stream_of(poses))
    _lib.check(rc, "slam_ext.ba")
    return dx, dz, energy, plan, ws


def run_planned(plan, ws, poses, disps, intrinsics, disps_sens, targets, weights, eta, depth_active, iterations,
                lm, ep, alpha, dx, dz, motion_only=False):
    """``iterations`` GN iterations of ``ba_extended`` on a prepared plan with
    caller-allocated workspace and outputs (dx [P,6] or [P,7] for Sim3 poses
    [N,8], dz [|kx|,H*W] or None).
    It allocates nothing, copies nothing to the host and launches on torch's
    current stream only, so it may be captured into a graph.  The inputs must be
    contiguous float32 device tensors of the plan's shapes (not re-checked)."""
    flags = (1 if motion_only else 0) | (8 if poses.shape[-1] == 8 else 0)
    rc = _lib.load().vh_ba_run(plan.h, _lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics), _lib.ptr(disps_sens),
                               _lib.ptr(targets), _lib.ptr(weights), None if motion_only else _lib.ptr(eta),
                               _lib.ptr(depth_active), int(iterations), float(lm), float(ep), float(alpha), 1e-6,
                               1e-6, 1.0, flags, _lib.ptr(dx), None if motion_only else _lib.ptr(dz), None,
                               _lib.ptr(ws), _lib.stream_of(poses))
    _lib.check(rc, "slam_ext.ba")


def ba_adjoint(plan, ws, poses, disps, intrinsics, targets, weights, depth_active, lm, motion_only, grad_dx,
               grad_dz):
    """Implicit adjoint of the single iteration that just ran on (plan, ws):
    returns [dL/dtargets, dL/dweights] ([E,2,H,W] float32) from dL/d(dx) [P,6]
    and dL/d(dz) [|kx|,H*W].  poses / disps / intrinsics are the values before
    that iteration.  This package's extension: the reference BA has no backward."""
    if poses.shape[-1] != 7:
        raise RuntimeError("slam_ext.ba_adjoint: SE3 poses [N,7] only")
    E = targets.shape[0]
    gt = torch.empty_like(targets)
    gw = torch.empty_like(weights)
    gdx = grad_dx.contiguous().float()
    gdz = None if motion_only else grad_dz.contiguous().float()
    lib = _lib.load()
    with torch.cuda.device(poses.device):
        rc = lib.vh_ba_adjoint(plan.h, _lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics), _lib.ptr(targets),
                               _lib.ptr(weights), _lib.ptr(depth_active), float(lm), 1 if motion_only else 0,
                               _lib.ptr(gdx), _lib.ptr(gdz), _lib.ptr(gt) if E else None,
                               _lib.ptr(gw) if E else None, _lib.ptr(ws), _lib.stream_of(poses))
    _lib.check(rc, "slam_ext.ba_adjoint")
    return [gt, gw]


class _RigPlan(_Plan):
    def __init__(self, pi, pj, qi, qj, N, V, H, W, t0, t1, pixel_tile=0):
        lib = _lib.load()
        idx = [x.to(torch.int64).contiguous() for x in (pi, pj, qi, qj)]
        h = lib.vh_ba_plan_create_rig(*[C.c_void_p(x.data_ptr()) for x in idx], int(idx[0].numel()), int(N), int(V),
                                      int(H), int(W), int(t0), int(t1), int(pixel_tile))
        if not h:
            raise RuntimeError(f"slam_ext.ba_rig: {_lib.last_error()}")
        self.h = C.c_void_p(h)
        self.K = lib.vh_ba_plan_num_depth_frames(self.h)
        self.ws_bytes = int(lib.vh_ba_workspace_size(self.h))
        self.pixel_tile = int(lib.vh_ba_plan_pixel_tile(self.h))
        kx = torch.empty(self.K, dtype=torch.int64)
        lib.vh_ba_plan_depth_frames(self.h, C.c_void_p(kx.data_ptr()))
        self.kx = kx


def get_rig_plan(pi, pj, qi, qj, N, V, H, W, t0, t1, pixel_tile=0) -> _RigPlan:
    """The (cached) plan of a rig edge list: edge e from view qi[e] of pose
    pi[e] to view qj[e] of pose pj[e]; ``plan.kx`` lists the depth frames
    (p V + q) whose maps ``eta`` must hold, in that order."""
    idx = [x.detach().to("cpu", torch.int64) for x in (pi, pj, qi, qj)]
    key = ("rig",) + tuple(x.numpy().tobytes() for x in idx) + (int(N), int(V), int(H), int(W), int(t0), int(t1),
                                                              int(pixel_tile))
    plan = _PLANS.get(key)
    if plan is None:
        plan = _RigPlan(*idx, N, V, H, W, t0, t1, pixel_tile)
        _PLANS[key] = plan
        while len(_PLANS) > _PLAN_CACHE_SIZE:
            _PLANS.popitem(last=False)
    else:
        _PLANS.move_to_end(key)
    return plan


CAMERAS = {"pinhole": 0, "mei": 1}


def ba_rig(poses, disps, intrinsics, camera, rig, disps_sens, targets, weights, eta, pi, pj, qi, qj, depth_active,
           t0, t1, iterations, lm, ep, motion_only, alpha, compute_energy=False, robust=None):
    """Fused BA on a multi-view rig buffer or a Mei camera: the layouts the
    reference's fused path rejects (buffer.py:384-410).  This package's
    extension, with the fused kernel's conventions.

    poses [N,7] SE3 and disps [N*V,H,W] (the flattened [N,V,H,W] buffer) are
    updated in place; intrinsics [V,4] (pinhole) or [V,5] (Mei, fx fy cx cy
    k1) at the maps' resolution, fixed; ``camera`` "pinhole" or "mei"; rig
    [V,7] camera-to-rig SE3 per view; disps_sens [N*V,H,W]; targets / weights
    [E,2,H,W]; eta [|kx|,H,W] with kx = ``get_rig_plan(...).kx``;
    depth_active [N*V]; edge e runs from view qi[e] of pose pi[e] to view
    qj[e] of pose pj[e].  Returns [dx [t1-t0,6], dz [|kx|,H*W] or None,
    energy]."""
    if camera not in CAMERAS:
        raise ValueError(f"slam_ext.ba_rig: camera {camera!r} (pinhole or mei)")
    _check(targets=targets, weights=weights, poses=poses, disps=disps, intrinsics=intrinsics, rig=rig,
           disps_sens=disps_sens, pi=pi, pj=pj, qi=qi, qj=qj, depth_active=depth_active)
    for name, t in (("poses", poses), ("disps", disps), ("intrinsics", intrinsics), ("rig", rig),
                    ("disps_sens", disps_sens), ("targets", targets), ("weights", weights),
                    ("depth_active", depth_active)):
        _f32(name, t)
        if t.device.type != "cuda":
            raise RuntimeError(f"slam_ext.ba_rig runs on the GPU; {name} is on {t.device}")
    V = rig.shape[0]
    NV, H, W = disps.shape
    if rig.shape != (V, 7) or NV % V != 0:
        raise RuntimeError(f"rig must be [V,7] and disps [N*V,H,W] (got {tuple(rig.shape)}, {tuple(disps.shape)})")
    N = NV // V
    kd = 5 if camera == "mei" else 4
    if poses.shape != (N, 7):
        raise RuntimeError(f"poses must be [N,7] SE3 (got {tuple(poses.shape)})")
    if intrinsics.shape != (V, kd):
        raise RuntimeError(f"intrinsics must be [V,{kd}] for the {camera} model (got {tuple(intrinsics.shape)})")
    E = pi.shape[0]
    if targets.shape != (E, 2, H, W) or weights.shape != (E, 2, H, W):
        raise RuntimeError("targets/weights must be [E,2,H,W]")
    if depth_active.shape != (NV,) or disps_sens.shape != disps.shape:
        raise RuntimeError("depth_active must be [N*V] and disps_sens [N*V,H,W]")
    t0, t1, iterations = int(t0), int(t1), int(iterations)
    plan = get_rig_plan(pi, pj, qi, qj, N, V, H, W, t0, t1)
    dev = poses.device
    if not motion_only and eta.numel() != plan.K * H * W:
        raise RuntimeError(f"eta must hold |kx| = {plan.K} maps of {H}x{W}")
    eta_c = eta.contiguous().to(torch.float32)
    energy = torch.empty((iterations if compute_energy else 0,), dtype=torch.float32, device=dev)
    if iterations <= 0:
        return [None, None, energy]
    dx = torch.empty((t1 - t0, 6), dtype=torch.float32, device=dev)
    dz = None if motion_only else torch.empty((plan.K, H * W), dtype=torch.float32, device=dev)
    ws = torch.empty((max(plan.ws_bytes, 1),), dtype=torch.uint8, device=dev)
    flags = (1 if motion_only else 0) | (4 if compute_energy else 0)
    rk = (0, 0.0, 0.0, 0.0) if robust is None else tuple(robust)
    with torch.cuda.device(dev):
        rc = _lib.load().vh_ba_run_rig(plan.h, _lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics),
                                       CAMERAS[camera], _lib.ptr(rig), _lib.ptr(disps_sens), _lib.ptr(targets),
                                       _lib.ptr(weights), None if motion_only else _lib.ptr(eta_c),
                                       _lib.ptr(depth_active), iterations, float(lm), float(ep), float(alpha), flags,
                                       int(rk[0]), float(rk[1]), float(rk[2]), float(rk[3]), _lib.ptr(dx),
                                       _lib.ptr(dz), _lib.ptr(energy) if compute_energy else None, _lib.ptr(ws),
                                       _lib.stream_of(poses))
    _lib.check(rc, "slam_ext.ba_rig")
    return [dx, dz, energy]


def ba(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, t0, t1, iterations, lm, ep,
       motion_only, alpha=0.05):
    """Legacy entry: depth_active = 1, no intrinsics, no energy (geom_kernels.cu:1822-1830)."""
    depth_active = torch.ones(disps.shape[0], dtype=disps.dtype, device=disps.device)
    return ba_extended(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, depth_active, t0, t1,
                       iterations, lm, ep, motion_only, alpha, False, 1e-6, 1e-6, 1.0, False)


# ---------------------------------------------------------------------------
# geometry siblings (slam.cpp:52-55; geom_kernels.cu:526-953, 1849-1950)
# ---------------------------------------------------------------------------
def _geom_inputs(fn, **named):
    _check(**named)
    for k, v in named.items():
        if v.device.type != "cuda":
            raise RuntimeError(f"slam_ext.{fn} runs on the GPU; {k} is on {v.device}")


def _i64(name, t):
    if t.dtype != torch.int64:
        raise RuntimeError(f"{name} must be int64 (got {t.dtype})")


def iproj(poses, disps, intrinsics):
    """points [N,H,W,3]: each pixel back-projected with its disparity and mapped
    by its frame's pose, divided by the disparity (iproj_kernel, :887-953)."""
    _geom_inputs("iproj", poses=poses, disps=disps, intrinsics=intrinsics)
    for name, t in (("poses", poses), ("disps", disps), ("intrinsics", intrinsics)):
        _f32(name, t)
    N, H, W = disps.shape
    points = torch.empty((N, H, W, 3), dtype=torch.float32, device=disps.device)
    with torch.cuda.device(disps.device):
        rc = _lib.load().vh_geom_iproj(_lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics), N, H, W,
                                       _lib.ptr(points), _lib.stream_of(disps))
    _lib.check(rc, "slam_ext.iproj")
    return points


def projmap(poses, disps, intrinsics, ii, jj):
    """[coords [E,H,W,3], valid [E,H,W,1]]: pixels of frame ii[e] reprojected
    into frame jj[e]; coordinates stay (u, v) for points behind z = 0.01, the
    third channel is zero, valid = z > MIN_DEPTH (projmap_kernel, :526-611)."""
    _geom_inputs("projmap", poses=poses, disps=disps, intrinsics=intrinsics, ii=ii, jj=jj)
    for name, t in (("poses", poses), ("disps", disps), ("intrinsics", intrinsics)):
        _f32(name, t)
    _i64("ii", ii)
    _i64("jj", jj)
    N, H, W = disps.shape
    E = ii.shape[0]
    coords = torch.empty((E, H, W, 3), dtype=torch.float32, device=disps.device)
    valid = torch.empty((E, H, W, 1), dtype=torch.float32, device=disps.device)
    with torch.cuda.device(disps.device):
        rc = _lib.load().vh_geom_projmap(_lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics), _lib.ptr(ii),
                                         _lib.ptr(jj), E, N, H, W, _lib.ptr(coords), _lib.ptr(valid),
                                         _lib.stream_of(disps))
    _lib.check(rc, "slam_ext.projmap")
    return [coords, valid]


def frame_distance(poses, disps, intrinsics, pi, pj, qi, qj, di, beta):
    """dist [num] between frames pi[b] -> pj[b] using the disparity map di[b]
    and intrinsics rows qi[b], qj[b] of intrinsics [R,4]; 1000 when less than
    75 % of the beta-weighted pixel terms are valid (frame_distance_kernel,
    :613-768)."""
    _geom_inputs("frame_distance", poses=poses, disps=disps, intrinsics=intrinsics, pi=pi, pj=pj, qi=qi, qj=qj,
                 di=di)
    for name, t in (("poses", poses), ("disps", disps), ("intrinsics", intrinsics)):
        _f32(name, t)
    for name, t in (("pi", pi), ("pj", pj), ("qi", qi), ("qj", qj), ("di", di)):
        _i64(name, t)
    N, H, W = disps.shape
    num = pi.shape[0]
    dist = torch.empty((num,), dtype=torch.float32, device=disps.device)
    with torch.cuda.device(disps.device):
        rc = _lib.load().vh_geom_frame_distance(_lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics),
                                                int(intrinsics.shape[0]), _lib.ptr(pi), _lib.ptr(pj), _lib.ptr(qi),
                                                _lib.ptr(qj), _lib.ptr(di), num, N, H, W, float(beta),
                                                _lib.ptr(dist), _lib.stream_of(disps))
    _lib.check(rc, "slam_ext.frame_distance")
    return dist


def depth_filter(poses, disps, intrinsics, ix, thresh):
    """counter [num,H,W]: how many of the frames ix[b]-3 .. ix[b]+3 (excluding
    ix[b], within [0, N)) see each pixel of frame ix[b] at a consistent depth
    (depth_filter_kernel, :770-885)."""
    _geom_inputs("depth_filter", poses=poses, disps=disps, intrinsics=intrinsics, ix=ix, thresh=thresh)
    for name, t in (("poses", poses), ("disps", disps), ("intrinsics", intrinsics), ("thresh", thresh)):
        _f32(name, t)
    _i64("ix", ix)
    N, H, W = disps.shape
    num = ix.shape[0]
    counter = torch.empty((num, H, W), dtype=torch.float32, device=disps.device)
    with torch.cuda.device(disps.device):
        rc = _lib.load().vh_geom_depth_filter(_lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intrinsics), _lib.ptr(ix),
                                              _lib.ptr(thresh), num, N, H, W, _lib.ptr(counter),
                                              _lib.stream_of(disps))
    _lib.check(rc, "slam_ext.depth_filter")
    return counter
