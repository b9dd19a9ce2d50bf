"""Drop-in for ``vipe.ext.droid_net_ext`` (csrc/droid_net_ext/droid.cpp:24-82).

Same names, positional arguments and list returns as the reference binding:

    corr_index_forward(volume, coords, radius)               -> [corr]
    corr_index_backward(volume, coords, corr_grad, radius)    -> [volume_grad]
    altcorr_forward(fmap1, fmap2, coords, radius)             -> [corr]
    altcorr_backward(fmap1, fmap2, coords, corr_grad, radius) -> [fmap1_grad, fmap2_grad, coords_grad]
    altcorr_index_forward(fmap1, fmap2_pyramid, coords, ii, jj, radius) -> [corr]

Non-contiguous inputs raise ``RuntimeError("<name> must be contiguous")``
(droid.cpp:11-12).  The ops run on gfx950 on the current stream and need
device tensors, as the reference's CUDA-only kernels do.

``corr_build`` is this package's all-pairs volume + pyramid op behind
``vipe_amd.corr.CorrBlock`` (the reference does the same with torch.matmul and
avg_pool2d, vipe/slam/networks/droid_net.py:56-69, 94-102).
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _lib


def _contig(**named):
    for k, v in named.items():
        if not v.is_contiguous():
            raise RuntimeError(f"{k} must be contiguous")


def _gpu(*ts):
    for t in ts:
        if t.device.type != "cuda":
            raise RuntimeError("droid_net_ext ops run on the GPU; got a tensor on " + str(t.device))


def _code(t, allowed):
    if t.dtype not in allowed:
        raise RuntimeError(f"unsupported dtype {t.dtype}")
    return _lib.DTYPE_CODE[t.dtype]


_VOL_DT = (torch.float16, torch.float32, torch.float64)


def corr_index_forward(volume, coords, radius):
    _contig(volume=volume, coords=coords)
    _gpu(volume, coords)
    if coords.dtype != torch.float32:
        raise RuntimeError("coords must be float32")
    B, H1, W1, H2, W2 = volume.shape
    rd = 2 * int(radius) + 1
    corr = torch.empty((B, rd, rd, H1, W1), dtype=volume.dtype, device=volume.device)
    with torch.cuda.device(volume.device):
        rc = _lib.load().vh_corr_index_forward(_code(volume, _VOL_DT), B, H1, W1, H2, W2, _lib.ptr(volume),
                                               _lib.ptr(coords), int(radius), _lib.ptr(corr), _lib.stream_of(volume))
    _lib.check(rc, "corr_index_forward")
    return [corr]


def corr_index_backward(volume, coords, corr_grad, radius):
    _contig(volume=volume, coords=coords, corr_grad=corr_grad)
    _gpu(volume, coords, corr_grad)
    B, H1, W1, H2, W2 = volume.shape
    volume_grad = torch.zeros_like(volume)
    g = corr_grad if corr_grad.dtype == volume.dtype else corr_grad.to(volume.dtype).contiguous()
    with torch.cuda.device(volume.device):
        rc = _lib.load().vh_corr_index_backward(_code(volume, _VOL_DT), B, H1, W1, H2, W2, _lib.ptr(coords),
                                                _lib.ptr(g), int(radius), _lib.ptr(volume_grad),
                                                _lib.stream_of(volume))
    _lib.check(rc, "corr_index_backward")
    return [volume_grad]


def altcorr_forward(fmap1, fmap2, coords, radius):
    _contig(fmap1=fmap1, fmap2=fmap2, coords=coords)
    _gpu(fmap1, fmap2, coords)
    B, N, H1, W1, _ = coords.shape
    _, H2, W2, Cc = fmap2.shape
    rd = 2 * int(radius) + 1
    corr = torch.empty((B, N, rd * rd, H1, W1), dtype=fmap1.dtype, device=fmap1.device)
    with torch.cuda.device(fmap1.device):
        rc = _lib.load().vh_altcorr_forward(_code(fmap1, _VOL_DT), B, N, H1, W1, H2, W2,
                                            fmap1.shape[3], _lib.ptr(fmap1), _lib.ptr(fmap2), _lib.ptr(coords),
                                            int(radius), _lib.ptr(corr), _lib.stream_of(fmap1))
    _lib.check(rc, "altcorr_forward")
    return [corr]


def altcorr_backward(fmap1, fmap2, coords, corr_grad, radius):
    _contig(fmap1=fmap1, fmap2=fmap2, coords=coords, corr_grad=corr_grad)
    _gpu(fmap1, fmap2, coords, corr_grad)
    B, N, H1, W1, _ = coords.shape
    _, H2, W2, Cc = fmap2.shape
    # both are fully written by the call; the workspace holds the per-window tap
    # gradients from which fmap2_grad is gathered without atomics
    fmap1_grad = torch.empty((B, H1, W1, fmap1.shape[3]), dtype=fmap1.dtype, device=fmap1.device)
    fmap2_grad = torch.empty((B, H2, W2, Cc), dtype=fmap1.dtype, device=fmap1.device)
    coords_grad = torch.zeros((B, N, H1, W1, 2), dtype=fmap1.dtype, device=fmap1.device)
    lib = _lib.load()
    ws = torch.empty((int(lib.vh_altcorr_backward_workspace_size(B, N, H1, W1)),), dtype=torch.uint8,
                     device=fmap1.device)
    with torch.cuda.device(fmap1.device):
        rc = lib.vh_altcorr_backward(_code(fmap1, (torch.float32,)), B, N, H1, W1, H2, W2, fmap1.shape[3],
                                     _lib.ptr(fmap1), _lib.ptr(fmap2), _lib.ptr(coords), _lib.ptr(corr_grad),
                                     int(radius), _lib.ptr(fmap1_grad), _lib.ptr(fmap2_grad), _lib.ptr(ws),
                                     _lib.stream_of(fmap1))
    _lib.check(rc, "altcorr_backward")
    return [fmap1_grad, fmap2_grad, coords_grad]


_FEAT_DT = (torch.float32, torch.float16, torch.bfloat16, torch.float64)


def feat_to_nhwc(fmap, scale=1.0, out_dtype=torch.float32):
    """[E, C, H, W] (f32 / f64 / f16 / bf16) -> [E, H, W, C] times scale, in
    out_dtype (float32 or the map's dtype)."""
    _contig(fmap=fmap)
    _gpu(fmap)
    E, Cc, H, W = fmap.shape
    out = torch.empty((E, H, W, Cc), dtype=out_dtype, device=fmap.device)
    with torch.cuda.device(fmap.device):
        rc = _lib.load().vh_feat_to_nhwc(_code(fmap, _FEAT_DT), _lib.ptr(fmap), E, Cc, H * W, float(scale),
                                         _code(out, _FEAT_DT), _lib.ptr(out), _lib.stream_of(fmap))
    _lib.check(rc, "feat_to_nhwc")
    return out


def pool2x2_nhwc(fmap):
    """[E, H, W, C] -> [E, H/2, W/2, C]: means of 2x2 blocks (floor of odd sizes),
    taken in FP32 (FP64 for float64 maps) and rounded to the map's dtype."""
    _contig(fmap=fmap)
    _gpu(fmap)
    E, H, W, Cc = fmap.shape
    out = torch.empty((E, H // 2, W // 2, Cc), dtype=fmap.dtype, device=fmap.device)
    with torch.cuda.device(fmap.device):
        rc = _lib.load().vh_pool2x2_nhwc(_code(fmap, _FEAT_DT), _lib.ptr(fmap), E, H, W, Cc, _lib.ptr(out),
                                         _lib.stream_of(fmap))
    _lib.check(rc, "pool2x2_nhwc")
    return out


def corr_feature_grad(g1_levels, g2_levels, scale, dtype):
    """Gradients of CorrBlock's two feature maps from the per-level altcorr
    backward outputs: g1_levels[l] f32 [E, H, W, C], g2_levels[l] f32
    [E, H >> l, W >> l, C] -> (df1, df2) [E, C, H, W] in dtype, with
    df1 = scale * sum_l g1_l and df2 = scale * sum_l 4^-l g2_l spread over the
    2^l x 2^l blocks."""
    L = len(g1_levels)
    for t in (*g1_levels, *g2_levels):
        _contig(grad=t)
    _gpu(*g1_levels, *g2_levels)
    E, H, W, Cc = g1_levels[0].shape
    df1 = torch.empty((E, Cc, H, W), dtype=dtype, device=g1_levels[0].device)
    df2 = torch.empty((E, Cc, H, W), dtype=dtype, device=g1_levels[0].device)
    p1 = (C.c_void_p * L)(*[t.data_ptr() for t in g1_levels])
    p2 = (C.c_void_p * L)(*[t.data_ptr() for t in g2_levels])
    with torch.cuda.device(df1.device):
        rc = _lib.load().vh_corr_feature_grad(_code(df1, _FEAT_DT), E, Cc, H, W, L, p1, p2, float(scale),
                                              _lib.ptr(df1), _lib.ptr(df2), _lib.stream_of(df1))
    _lib.check(rc, "corr_feature_grad")
    return df1, df2


def altcorr_index_forward(fmap1, fmap2_pyramid, coords, ii, jj, radius):
    _contig(fmap1=fmap1, coords=coords, ii=ii, jj=jj)
    for f in fmap2_pyramid:
        _contig(fmap2=f)
    _gpu(fmap1, coords, ii, jj, *fmap2_pyramid)
    B, Nf, H, W, Cc = fmap1.shape
    _, M, _, _, _ = coords.shape
    L = len(fmap2_pyramid)
    rd = 2 * int(radius) + 1
    corr = torch.empty((B, M, L * rd * rd, H, W), dtype=torch.float32, device=fmap1.device)
    lv = (C.c_void_p * max(L, 1))(*[f.data_ptr() for f in fmap2_pyramid])
    Hl = (C.c_int * max(L, 1))(*[f.shape[2] for f in fmap2_pyramid])
    Wl = (C.c_int * max(L, 1))(*[f.shape[3] for f in fmap2_pyramid])
    iil = ii if ii.dtype == torch.int64 else ii.long().contiguous()
    jjl = jj if jj.dtype == torch.int64 else jj.long().contiguous()
    with torch.cuda.device(fmap1.device):
        rc = _lib.load().vh_altcorr_index_forward(_code(fmap1, _VOL_DT), B, Nf, M, H, W, Cc,
                                                  _lib.ptr(fmap1), L, lv, Hl, Wl, _lib.ptr(coords), _lib.ptr(iil),
                                                  _lib.ptr(jjl), int(radius), _lib.ptr(corr), _lib.stream_of(fmap1))
    _lib.check(rc, "altcorr_index_forward")
    return [corr]


PIECE_BYTES = 16  # one tile row of one plane (kPieceBytes in corr.hip)
TILE_PLANES = 4  # source planes sharing a tile (kTilePlanes)
TILE_ROWS = 2  # rows of a plane in a tile (kTileRows)
TILE_BYTES = PIECE_BYTES * TILE_PLANES * TILE_ROWS


def tiled_plane(Hl, Wl, dtype):
    """(Hp, Wp, tw) of a tiled volume plane: Hl, Wl padded to TILE_ROWS x tw
    blocks (tw columns = 16 bytes); the block of TILE_PLANES consecutive planes
    is one TILE_BYTES tile."""
    tw = PIECE_BYTES // torch.empty((), dtype=dtype).element_size()
    return (Hl + TILE_ROWS - 1) // TILE_ROWS * TILE_ROWS, (Wl + tw - 1) // tw * tw, tw


def _check_groups(n, what):
    if n % TILE_PLANES:
        raise RuntimeError(f"{what}: tiled levels group the planes of an edge by {TILE_PLANES}; "
                           f"H*W must be a multiple of it (got {n})")


def untile(level, Hl, Wl):
    """Tiled level [B,H1,W1,Hp*Wp] -> planar [B,H1,W1,Hl,Wl] (copy).  Planes p of a
    group of TILE_PLANES share tiles of TILE_ROWS rows x tw columns, 16-byte piece
    (row % TILE_ROWS) * TILE_PLANES + p % TILE_PLANES."""
    B, H1, W1, _ = level.shape
    _check_groups(H1 * W1, "untile")
    Hp, Wp, tw = tiled_plane(Hl, Wl, level.dtype)
    v = level.view(B * H1 * W1 // TILE_PLANES, Hp // TILE_ROWS, Wp // tw, TILE_ROWS, TILE_PLANES, tw)
    return v.permute(0, 4, 1, 3, 2, 5).reshape(B, H1, W1, Hp, Wp)[..., :Hl, :Wl].contiguous()


def tile(volume):
    """Planar [B,H1,W1,Hl,Wl] -> tiled [B,H1,W1,Hp*Wp] with zero padding (copy)."""
    B, H1, W1, Hl, Wl = volume.shape
    _check_groups(H1 * W1, "tile")
    Hp, Wp, tw = tiled_plane(Hl, Wl, volume.dtype)
    pad = torch.zeros((B, H1, W1, Hp, Wp), dtype=volume.dtype, device=volume.device)
    pad[..., :Hl, :Wl] = volume
    v = pad.view(B * H1 * W1 // TILE_PLANES, TILE_PLANES, Hp // TILE_ROWS, TILE_ROWS, Wp // tw, tw)
    return v.permute(0, 2, 4, 3, 1, 5).reshape(B, H1, W1, Hp * Wp).contiguous()


def corr_lookup_pyramid(levels, coords, radius, level_dims=None, layout="planar"):
    """Fused CorrBlock lookup: levels[l] [B,H1,W1,Hl,Wl] (planar) or [B,H1,W1,Hp*Wp]
    (tiled, with level_dims[l] = (Hl, Wl)), coords [B,H1,W1,2] (level-0 pixels)
    -> [B, L*(2r+1)^2, H1, W1] in the volume dtype."""
    _contig(coords=coords)
    for lvl in levels:
        _contig(volume=lvl)
    _gpu(coords, *levels)
    if coords.dtype != torch.float32:
        raise RuntimeError("coords must be float32")
    tiled = layout == "tiled"
    if tiled and level_dims is None:
        raise RuntimeError("tiled levels need level_dims")
    B, H1, W1 = levels[0].shape[:3]
    L = len(levels)
    rd = 2 * int(radius) + 1
    out = torch.empty((B, L * rd * rd, H1, W1), dtype=levels[0].dtype, device=coords.device)
    lv = (C.c_void_p * L)(*[t.data_ptr() for t in levels])
    dims = level_dims if tiled else [(t.shape[3], t.shape[4]) for t in levels]
    Hl = (C.c_int * L)(*[d[0] for d in dims])
    Wl = (C.c_int * L)(*[d[1] for d in dims])
    with torch.cuda.device(coords.device):
        rc = _lib.load().vh_corr_lookup_pyramid(_code(levels[0], _VOL_DT), B, H1, W1, L, lv, Hl, Wl,
                                                _lib.VH_LAYOUT_TILED if tiled else _lib.VH_LAYOUT_PLANAR,
                                                _lib.ptr(coords), int(radius), _lib.ptr(out),
                                                _lib.stream_of(coords))
    _lib.check(rc, "corr_lookup_pyramid")
    return out


_MMA_CODE = {torch.float32: _lib.VH_F32, torch.float16: _lib.VH_F16, torch.bfloat16: _lib.VH_BF16,
             torch.float8_e4m3fn: _lib.VH_FP8E4M3}


def corr_build(fmap1, fmap2, num_levels=4, out_dtype=None, contraction_dtype=None, layout="planar"):
    """All-pairs correlation volume pyramid of CorrBlock.  fmap1, fmap2 [E,C,H,W]
    (f32 / f16 / bf16) -> list of level-l volumes [E,H,W,H>>l,W>>l] in out_dtype
    (default: f16 for 16-bit inputs, f32 for f32 inputs).  contraction_dtype
    selects the MFMA operand type (default: the input dtype); float8_e4m3fn
    balances each channel between the two maps and scales each pixel row into
    e4m3, accumulating in FP32.  layout "tiled" stores each level's planes as
    TILE_BYTES tiles shared by groups of TILE_PLANES planes, [E,H,W,Hp*Wp] (H*W a
    multiple of TILE_PLANES; see tiled_plane / untile)."""
    _contig(fmap1=fmap1, fmap2=fmap2)
    _gpu(fmap1, fmap2)
    if fmap2.shape != fmap1.shape or fmap2.dtype != fmap1.dtype:
        raise RuntimeError("fmap1 and fmap2 must have the same shape and dtype")
    E, Cc, H, W = fmap1.shape
    if out_dtype is None:
        out_dtype = torch.float32 if fmap1.dtype == torch.float32 else torch.float16
    mma = _MMA_CODE.get(contraction_dtype if contraction_dtype is not None else fmap1.dtype, -1)
    tiled = layout == "tiled"
    levels = []
    h, w = H, W
    for _ in range(num_levels):
        if tiled:
            Hp, Wp, _ = tiled_plane(h, w, out_dtype)
            levels.append(torch.empty((E, H, W, Hp * Wp), dtype=out_dtype, device=fmap1.device))
        else:
            levels.append(torch.empty((E, H, W, h, w), dtype=out_dtype, device=fmap1.device))
        h, w = h // 2, w // 2
    lib = _lib.load()
    ws = torch.empty((max(int(lib.vh_corr_build_workspace(mma, E, Cc, H, W)), 1),), dtype=torch.uint8,
                     device=fmap1.device)
    lv = (C.c_void_p * num_levels)(*[t.data_ptr() for t in levels])
    with torch.cuda.device(fmap1.device):
        rc = lib.vh_corr_build(_code(fmap1, (torch.float32, torch.float16, torch.bfloat16)), mma,
                               _lib.DTYPE_CODE.get(out_dtype, -1), E, Cc, H, W, _lib.ptr(fmap1), _lib.ptr(fmap2),
                               num_levels, lv, _lib.VH_LAYOUT_TILED if tiled else _lib.VH_LAYOUT_PLANAR,
                               _lib.ptr(ws), _lib.stream_of(fmap1))
    _lib.check(rc, "corr_build")
    return levels
