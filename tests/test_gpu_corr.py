"""gfx950 correlation volume, pyramid and lookups vs the FP64 oracle."""
import math

import pytest
import torch

from oracle import corr as OC

pytestmark = pytest.mark.gpu


def _coords(B, H, W, H2, W2, g, spread=3.0):
    yy, xx = torch.meshgrid(torch.arange(H, dtype=torch.float32), torch.arange(W, dtype=torch.float32), indexing="ij")
    base = torch.stack([xx * (W2 / W), yy * (H2 / H)], 0)[None].repeat(B, 1, 1, 1)
    return (base + spread * torch.randn(B, 2, H, W, generator=g)).contiguous()


# float64 volumes still use float bilinear weights, as the reference forms them
# (scalar_t(dx * dy) with float dx, dy): agreement with the FP64 oracle is ~1e-7.
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.float64, 5e-8), (torch.float16, 2e-3)])
def test_corr_index_forward_backward(dtype, tol):
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(0)
    B, H1, W1, H2, W2, r = 3, 6, 7, 9, 11, 3
    vol = torch.randn(B, H1, W1, H2, W2, generator=g).to(dtype)
    coords = _coords(B, H1, W1, H2, W2, g)
    (out,) = D.corr_index_forward(vol.cuda(), coords.cuda(), r)
    ref = OC.corr_index_forward(vol.double(), coords.double(), r)
    assert out.shape == ref.shape and out.dtype == dtype
    assert (out.double().cpu() - ref).abs().max().item() <= tol * 8
    cg = torch.randn(ref.shape, generator=g).to(dtype)
    (vg,) = D.corr_index_backward(vol.cuda(), coords.cuda(), cg.cuda(), r)
    vref = OC.corr_index_backward(vol.shape, coords.double(), cg.double(), r)
    assert (vg.double().cpu() - vref).abs().max().item() <= tol * 8


# (input, contraction, output, abs tol against FP64 of the same rounded inputs).
# FP8 is checked against the unquantised FP64 volume: its bound is relative to
# the volume's largest magnitude (e4m3 keeps 3 mantissa bits).
BUILD_CASES = [
    (torch.float32, torch.float32, torch.float32, 2e-5),
    (torch.float16, torch.float16, torch.float16, 3e-3),
    (torch.bfloat16, torch.bfloat16, torch.float32, 2e-5),
    (torch.float32, torch.float16, torch.float32, 2e-3),
    (torch.float32, torch.bfloat16, torch.float16, 2e-2),
    (torch.float32, torch.float8_e4m3fn, torch.float32, None),
]


@pytest.mark.parametrize("in_dt,mma_dt,out_dt,tol", BUILD_CASES)
@pytest.mark.parametrize("E,C,H,W", [(3, 128, 12, 16), (2, 40, 13, 18), (1, 128, 24, 40), (1, 32, 11, 21)])
def test_corr_build_and_pyramid(in_dt, mma_dt, out_dt, tol, E, C, H, W):
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(1)
    f1 = (torch.randn(E, C, H, W, generator=g) / math.sqrt(C))
    f2 = (torch.randn(E, C, H, W, generator=g) / math.sqrt(C))
    levels = D.corr_build(f1.to(in_dt).cuda(), f2.to(in_dt).cuda(), 4, out_dtype=out_dt, contraction_dtype=mma_dt)
    # reference inputs: the input rounding, then the contraction rounding (f16 / bf16)
    r1, r2 = f1.to(in_dt).double(), f2.to(in_dt).double()
    if mma_dt in (torch.float16, torch.bfloat16):
        r1, r2 = r1.to(mma_dt).double(), r2.to(mma_dt).double()
    ref = OC.corr_pyramid(OC.corr_volume(r1[None], r2[None]), 4)
    for lv, rv in zip(levels, ref):
        assert lv.shape == rv.shape and lv.dtype == out_dt
        err = (lv.double().cpu() - rv).abs().max().item()
        if tol is None:
            bound = 0.06 * ref[0].abs().max().item()
            rel = ((lv.double().cpu() - rv).norm() / rv.norm()).item()
            print(f"[fp8 build] E={E} C={C} {H}x{W} level max err {err:.3e} (bound {bound:.3e}), rel L2 {rel:.3e}")
            assert err <= bound and rel < 0.05, (err, rel)
        else:
            assert err <= tol, err


def test_altcorr_forward_backward_and_indexed():
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(2)
    B, N, H, W, C, r = 2, 3, 8, 10, 64, 3
    f1 = torch.randn(B, H, W, C, generator=g) / math.sqrt(C)
    f2 = torch.randn(B, H, W, C, generator=g) / math.sqrt(C)
    coords = _coords(B * N, H, W, H, W, g).view(B, N, 2, H, W).permute(0, 1, 3, 4, 2).contiguous()
    (out,) = D.altcorr_forward(f1.cuda(), f2.cuda(), coords.cuda(), r)
    ref = OC.altcorr_forward(f1.double(), f2.double(), coords.double(), r)
    assert (out.double().cpu() - ref).abs().max().item() < 1e-5
    cg = torch.randn(ref.shape, generator=g)
    g1, g2, gc = D.altcorr_backward(f1.cuda(), f2.cuda(), coords.cuda(), cg.cuda(), r)
    r1, r2, _ = OC.altcorr_backward(f1.double(), f2.double(), coords.double(), cg.double(), r)
    assert (g1.double().cpu() - r1).abs().max().item() < 1e-4
    assert (g2.double().cpu() - r2).abs().max().item() < 1e-4
    assert torch.count_nonzero(gc) == 0

    # indexed multi-level lookup (test_altcorr_indexed.py construction, FP16 features)
    Nf, M = 5, 4
    fmaps = torch.randn(1, Nf, C, H * 2, W * 2, generator=g).half()
    pyr = OC.altcorr_pyramid(fmaps.double(), 4)
    pyr16 = [p.half().contiguous() for p in pyr]
    ii = torch.tensor([0, 1, 2, 3])
    jj = torch.tensor([1, 2, 3, 4])
    yy, xx = torch.meshgrid(torch.arange(2 * H), torch.arange(2 * W), indexing="ij")
    cc = torch.stack([xx, yy], -1).float()[None, None].repeat(1, M, 1, 1, 1)
    cc = cc + 0.4 * torch.randn(cc.shape, generator=g)
    (idx,) = D.altcorr_index_forward(pyr16[0].cuda(), [p.cuda() for p in pyr16], cc.cuda(), ii.cuda(), jj.cuda(), r)
    iref = OC.altcorr_index_forward(pyr16[0].double(), [p.double() for p in pyr16], cc.double(), ii, jj, r)
    assert idx.dtype == torch.float32
    assert (idx.double().cpu() - iref).abs().max().item() < 2e-4


@pytest.mark.parametrize("layout", ["planar", "tiled"])
@pytest.mark.parametrize("vol_dt,tol", [(torch.float32, 1e-5), (torch.float16, 2e-3)])
@pytest.mark.parametrize("H,W", [(12, 16), (13, 20), (12, 21)])
def test_fused_pyramid_lookup_matches_per_level_and_oracle(layout, vol_dt, tol, H, W):
    from vipe_amd.corr import CorrBlock

    g = torch.Generator().manual_seed(5)
    B, N, C = 1, 3, 32
    f1 = torch.randn(B, N, C, H, W, generator=g) / math.sqrt(C)
    f2 = torch.randn(B, N, C, H, W, generator=g) / math.sqrt(C)
    coords = torch.rand(B, N, H, W, 2, generator=g) * torch.tensor([W + 4.0, H + 4.0]) - 2.0
    blk = CorrBlock(f1.cuda(), f2.cuda(), num_levels=4, radius=3, volume_dtype=vol_dt, layout=layout)
    assert blk.layout == layout
    fused = blk(coords.cuda())
    per_level = blk.lookup_per_level(coords.cuda())
    assert torch.equal(fused, per_level)
    ref = OC.corr_block_lookup(OC.corr_pyramid(OC.corr_volume(f1.double(), f2.double()), 4), coords.double(), 3)
    assert (fused.double().cpu() - ref).abs().max().item() < tol
    # backward through the fused op equals backward through the per-level op
    lv = [t.detach().clone().requires_grad_(True) for t in blk.corr_pyramid]
    blk.corr_pyramid = lv
    go = torch.randn(fused.shape, generator=g).to(vol_dt).cuda()
    (blk(coords.cuda()) * go).sum().backward()
    g_fused = [t.grad.clone() for t in lv]
    for t in lv:
        t.grad = None
    (blk.lookup_per_level(coords.cuda()) * go).sum().backward()
    for a, b in zip(g_fused, lv):
        assert torch.equal(a, b.grad)


@pytest.mark.parametrize("out_dt", [torch.float32, torch.float16])
@pytest.mark.parametrize("E,C,H,W", [(2, 40, 13, 20), (1, 128, 24, 40), (2, 32, 12, 21), (1, 160, 12, 35)])
def test_tiled_build_equals_planar_build(out_dt, E, C, H, W):
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(3)
    f1 = (torch.randn(E, C, H, W, generator=g) / math.sqrt(C)).cuda()
    f2 = (torch.randn(E, C, H, W, generator=g) / math.sqrt(C)).cuda()
    planar = D.corr_build(f1, f2, 4, out_dtype=out_dt)
    tiled = D.corr_build(f1, f2, 4, out_dtype=out_dt, layout="tiled")
    for lvl, (p, t) in enumerate(zip(planar, tiled)):
        assert torch.equal(D.untile(t, H >> lvl, W >> lvl), p)
        assert torch.equal(t, D.tile(p))  # padding is zero


def test_tiled_layout_needs_whole_plane_groups():
    """Tiled levels group neighbouring source planes by 4, so an edge's H*W must be a
    multiple of 4: the build refuses other counts and CorrBlock keeps such volumes
    planar."""
    from vipe_amd.corr import CorrBlock
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(4)
    f1 = torch.randn(1, 2, 32, 11, 21, generator=g)
    f2 = torch.randn(1, 2, 32, 11, 21, generator=g)
    with pytest.raises(RuntimeError, match="group"):
        D.corr_build(f1[0].cuda(), f2[0].cuda(), 4, layout="tiled")
    blk = CorrBlock(f1.cuda(), f2.cuda(), num_levels=4, radius=3)
    assert blk.layout == "planar"
    coords = torch.rand(1, 2, 11, 21, 2, generator=g) * torch.tensor([21.0, 11.0])
    ref = OC.corr_block_lookup(OC.corr_pyramid(OC.corr_volume(f1.double(), f2.double()), 4), coords.double(), 3)
    assert (blk(coords.cuda()).double().cpu() - ref).abs().max().item() < 1e-5


def test_corr_lookup_gradients_to_both_feature_maps():
    """CorrBlock lookup backward to fmap1 and fmap2 against FP64 autograd through
    the oracle's volume + avg_pool2d + bilinear lookup, and against central
    differences (FP64, eps=1e-6) of the oracle on 64 random directions."""
    from vipe_amd.corr import CorrBlock

    g = torch.Generator().manual_seed(7)
    B, N, C, H, W = 1, 3, 32, 12, 16
    f1 = torch.randn(B, N, C, H, W, generator=g) / math.sqrt(C)
    f2 = torch.randn(B, N, C, H, W, generator=g) / math.sqrt(C)
    coords = torch.rand(B, N, H, W, 2, generator=g) * torch.tensor([W + 2.0, H + 2.0]) - 1.0
    go = torch.randn(B, N, 4 * 49, H, W, generator=g)

    a1, a2 = f1.cuda().requires_grad_(True), f2.cuda().requires_grad_(True)
    out = CorrBlock(a1, a2, num_levels=4, radius=3)(coords.cuda())
    (out * go.cuda()).sum().backward()
    d1, d2 = a1.grad.double().cpu(), a2.grad.double().cpu()

    def loss(x1, x2):
        o = OC.corr_block_lookup(OC.corr_pyramid(OC.corr_volume(x1, x2), 4), coords.double(), 3)
        return (o * go.double()).sum()

    r1, r2 = f1.double().requires_grad_(True), f2.double().requires_grad_(True)
    loss(r1, r2).backward()
    scale = max(r1.grad.abs().max().item(), r2.grad.abs().max().item())
    assert (d1 - r1.grad).abs().max().item() < 1e-5 * scale
    assert (d2 - r2.grad).abs().max().item() < 1e-5 * scale

    eps = 1e-6
    f1d, f2d = f1.double(), f2.double()
    errs = []
    for k in range(64):
        v1 = torch.randn(f1.shape, generator=g, dtype=torch.float64)
        v2 = torch.randn(f2.shape, generator=g, dtype=torch.float64)
        with torch.no_grad():
            fd = (loss(f1d + eps * v1, f2d + eps * v2) - loss(f1d - eps * v1, f2d - eps * v2)).item() / (2 * eps)
        an = ((d1 * v1).sum() + (d2 * v2).sum()).item()
        # error relative to |grad| |v| (the bound of any directional derivative):
        # FP32 gradients with atomically accumulated fmap2 terms sit near 1e-7
        gv = (d1.norm() ** 2 + d2.norm() ** 2).sqrt().item() * (v1.norm() ** 2 + v2.norm() ** 2).sqrt().item()
        errs.append(abs(fd - an) / gv)
    print(f"[corr feature grads] max normalised directional error over 64 directions: {max(errs):.3e}")
    assert max(errs) < 1e-5


@pytest.mark.parametrize("C", [128, 96, 40])
@pytest.mark.parametrize("dt", [torch.float32, torch.float16])
def test_altcorr_index_vector_and_scalar_paths(C, dt):
    """altcorr_index_forward on each of its paths against the FP64 oracle: one
    wave per pixel (C = 128; C = 96 in FP32), one lane per pixel with 16-byte
    vectors (C = 96 in FP16, not a whole 64-channel step) and per element
    (C = 40, not whole 32-channel chunks); some windows lie partly or wholly
    outside the target image."""
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(5 + C)
    Nf, M, H, W, r = 4, 5, 12, 14, 3
    fmaps = (torch.randn(1, Nf, C, H, W, generator=g) / math.sqrt(C)).to(dt)
    pyr = OC.altcorr_pyramid(fmaps.double(), 4)
    pyr_t = [p.to(dt).contiguous() for p in pyr]
    ii = torch.tensor([0, 1, 2, 3, 0])
    jj = torch.tensor([1, 2, 3, 0, 2])
    yy, xx = torch.meshgrid(torch.arange(H), torch.arange(W), indexing="ij")
    cc = torch.stack([xx, yy], -1).float()[None, None].repeat(1, M, 1, 1, 1)
    cc = cc + 1.5 * torch.randn(cc.shape, generator=g)
    cc[0, 1, :3] += 30.0   # windows wholly outside
    cc[0, 2, :, :2] -= 5.0  # windows across the left border
    (out,) = D.altcorr_index_forward(pyr_t[0].cuda(), [p.cuda() for p in pyr_t], cc.cuda(), ii.cuda(), jj.cuda(), r)
    ref = OC.altcorr_index_forward(pyr_t[0].double(), [p.double() for p in pyr_t], cc.double(), ii, jj, r)
    assert out.dtype == torch.float32 and out.shape == ref.shape
    assert (out.double().cpu() - ref).abs().max().item() < 2e-5 * max(1.0, ref.abs().max().item())


@pytest.mark.parametrize("C", [128, 96, 40])
@pytest.mark.parametrize("dt", [torch.float32, torch.float16])
def test_altcorr_forward_paths(C, dt):
    """altcorr_forward (CorrLayer) on each of its paths against the FP64 oracle:
    one wave per pixel (C = 128; C = 96 in FP32), one lane per pixel with 16-byte
    vectors (C = 96 in FP16) and per element (C = 40).  FP16 outputs are compared
    at one FP16 rounding of the largest output."""
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(17 + C)
    B, N, H, W, r = 2, 2, 9, 13, 3
    f1 = (torch.randn(B, H, W, C, generator=g) / math.sqrt(C)).to(dt)
    f2 = (torch.randn(B, H, W, C, generator=g) / math.sqrt(C)).to(dt)
    coords = _coords(B * N, H, W, H, W, g).view(B, N, 2, H, W).permute(0, 1, 3, 4, 2).contiguous()
    coords[0, 1, :2] += 20.0  # windows wholly outside
    (out,) = D.altcorr_forward(f1.cuda(), f2.cuda(), coords.cuda(), r)
    ref = OC.altcorr_forward(f1.double(), f2.double(), coords.double(), r)
    assert out.dtype == dt and out.shape == ref.shape
    tol = 1e-5 if dt == torch.float32 else 2 ** -10 * max(1.0, ref.abs().max().item())
    assert (out.double().cpu() - ref).abs().max().item() < tol


@pytest.mark.parametrize("C", [128, 96])
@pytest.mark.parametrize("down", [2, 8])
def test_altcorr_backward_production_shapes(C, down):
    """altcorr_backward as AltCorrBlock's pooled levels use it: C = 128 (two
    passes of the 64-lane channel loop) and C = 96 (a partial second pass),
    fmap2 pooled to H/2 or H/8, N = 2 coordinate sets, some windows partly or
    wholly outside the target; both gradients against the FP64 oracle."""
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(40 + C + down)
    B, N, H, W, r = 2, 2, 16, 24, 3
    H2, W2 = H // down, W // down
    f1 = torch.randn(B, H, W, C, generator=g) / math.sqrt(C)
    f2 = torch.randn(B, H2, W2, C, generator=g) / math.sqrt(C)
    coords = _coords(B * N, H, W, H2, W2, g, spread=1.5).view(B, N, 2, H, W).permute(0, 1, 3, 4, 2).contiguous()
    coords[0, 1, :3] += 40.0   # wholly outside
    coords[1, 0, :, :4] -= 3.0  # across the left border
    cg = torch.randn(B, N, 49, H, W, generator=g)
    g1, g2, gc = D.altcorr_backward(f1.cuda(), f2.cuda(), coords.cuda(), cg.cuda(), r)
    r1, r2, _ = OC.altcorr_backward(f1.double(), f2.double(), coords.double(), cg.double(), r)
    for got, ref in ((g1, r1), (g2, r2)):
        err = (got.double().cpu() - ref).abs().max().item()
        assert err <= 1e-5 * max(1.0, ref.abs().max().item()), err
    assert torch.count_nonzero(gc) == 0


def test_altcorr_backward_with_no_coordinate_sets_writes_zeros():
    from vipe_amd.ext import droid_net_ext as D

    B, H, W, C = 1, 6, 8, 64
    f1 = torch.randn(B, H, W, C).cuda()
    f2 = torch.randn(B, H, W, C).cuda()
    coords = torch.zeros(B, 0, H, W, 2).cuda()
    cg = torch.zeros(B, 0, 49, H, W).cuda()
    g1, g2, _ = D.altcorr_backward(f1, f2, coords, cg, 3)
    assert torch.count_nonzero(g1) == 0 and torch.count_nonzero(g2) == 0


@pytest.mark.parametrize("C", [128, 40])
def test_altcorr_float64_forwards(C):
    """float64 features (the reference dispatches floating types and half):
    altcorr_forward returns float64, altcorr_index_forward float32."""
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(60 + C)
    B, N, H, W, r = 1, 2, 9, 13, 3
    f1 = torch.randn(B, H, W, C, generator=g, dtype=torch.float64) / math.sqrt(C)
    f2 = torch.randn(B, H, W, C, generator=g, dtype=torch.float64) / math.sqrt(C)
    coords = _coords(B * N, H, W, H, W, g).view(B, N, 2, H, W).permute(0, 1, 3, 4, 2).contiguous()
    (out,) = D.altcorr_forward(f1.cuda(), f2.cuda(), coords.cuda(), r)
    ref = OC.altcorr_forward(f1, f2, coords.double(), r)
    assert out.dtype == torch.float64
    # bilinear weights are formed from float coordinates, as in the reference
    assert (out.cpu() - ref).abs().max().item() < 1e-7

    Nf, M = 3, 2
    fmaps = torch.randn(1, Nf, C, H, W, generator=g, dtype=torch.float64) / math.sqrt(C)
    pyr = [p.contiguous() for p in OC.altcorr_pyramid(fmaps, 3)]
    ii, jj = torch.tensor([0, 2]), torch.tensor([1, 0])
    yy, xx = torch.meshgrid(torch.arange(H), torch.arange(W), indexing="ij")
    cc = (torch.stack([xx, yy], -1).float()[None, None].repeat(1, M, 1, 1, 1)
          + torch.randn(1, M, H, W, 2, generator=g)).contiguous()
    (idx,) = D.altcorr_index_forward(pyr[0].cuda(), [p.cuda() for p in pyr], cc.cuda(), ii.cuda(), jj.cuda(), r)
    iref = OC.altcorr_index_forward(pyr[0], pyr, cc.double(), ii, jj, r)
    assert idx.dtype == torch.float32
    assert (idx.double().cpu() - iref).abs().max().item() < 1e-6 * max(1.0, iref.abs().max().item())


@pytest.mark.parametrize("C", [64, 128, 256])
def test_altcorr_index_matrix_core_path(C):
    """FP16 altcorr_index_forward on the matrix cores (32-pixel groups, their
    union of window rows as 32 x 64 MFMA blocks) against the FP64 oracle:
    smooth flow (the MFMA path), a row whose coordinates diverge (the per-pixel
    fallback), windows past every border, rows of 80 pixels (a partial last
    group), all 4 levels."""
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(9 + C)
    Nf, M, H, W, r = 3, 4, 10, 80, 3
    fmaps = (torch.randn(1, Nf, C, H, W, generator=g) / math.sqrt(C)).half()
    pyr = OC.altcorr_pyramid(fmaps.double(), 4)
    pyr_t = [p.half().contiguous() for p in pyr]
    ii = torch.tensor([0, 1, 2, 0])
    jj = torch.tensor([1, 2, 0, 2])
    yy, xx = torch.meshgrid(torch.arange(H, dtype=torch.float32), torch.arange(W, dtype=torch.float32), indexing="ij")
    flow = torch.stack([3.0 * torch.sin(0.05 * xx + 0.3 * yy), 2.0 * torch.cos(0.04 * xx - 0.2 * yy)], -1)
    cc = (torch.stack([xx, yy], -1) + flow)[None, None].repeat(1, M, 1, 1, 1)
    cc = cc + 0.3 * torch.randn(cc.shape, generator=g)
    cc[0, 1, 4] = torch.rand(W, 2, generator=g) * torch.tensor([160.0, 40.0]) - 40.0  # diverging row
    cc[0, 2, :2] -= torch.tensor([0.0, 12.0])   # above the top
    cc[0, 3, :, :20] -= torch.tensor([30.0, 0.0])  # left of the left border
    cc[0, 3, :, 60:] += torch.tensor([25.0, 0.0])  # right of the right border
    (out,) = D.altcorr_index_forward(pyr_t[0].cuda(), [p.cuda() for p in pyr_t], cc.cuda(), ii.cuda(), jj.cuda(), r)
    ref = OC.altcorr_index_forward(pyr_t[0].double(), [p.double() for p in pyr_t], cc.double(), ii, jj, r)
    assert out.dtype == torch.float32 and out.shape == ref.shape
    err = (out.double().cpu() - ref).abs().max().item()
    print(f"[altcorr mma C={C}] max err {err:.3e} (max |ref| {ref.abs().max().item():.3e})")
    assert err < 2e-5 * max(1.0, ref.abs().max().item())


@pytest.mark.parametrize("C,W,conc,spread", [(128, 80, False, 2.0), (96, 24, True, 2.0), (36, 20, False, 2.0),
                                             (64, 40, False, 0.2), (28, 21, False, 0.2)])
def test_altcorr_backward_gather_is_deterministic(C, W, conc, spread):
    """fmap2_grad gathered by target rows (no atomics): bit-identical on repeated
    calls, equal to the FP64 oracle, and close to the atomic path (workspace
    NULL) of the same entry point; with every window crowded onto a few target
    rows (conc), with channel counts of 128, 96, 64, 36 and 28, with smooth
    coordinates (four neighbouring windows share one bounding box) and with
    scattered ones (each window walked alone), and with H * W not a multiple
    of four (one pixel per wave)."""
    from vipe_amd import _lib
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(70 + C)
    B, N, H, r = 2, 3, 12 if W % 2 == 0 else 13, 3
    H2, W2 = H, W
    f1 = torch.randn(B, H, W, C, generator=g) / math.sqrt(C)
    f2 = torch.randn(B, H2, W2, C, generator=g) / math.sqrt(C)
    coords = _coords(B * N, H, W, H2, W2, g, spread=spread).view(B, N, 2, H, W).permute(0, 1, 3, 4, 2).contiguous()
    if conc:
        coords[..., 1] = 4.0 + 0.5 * torch.rand(coords[..., 1].shape, generator=g)
    cg = torch.randn(B, N, 49, H, W, generator=g)
    a1, a2, ac, acg = f1.cuda(), f2.cuda(), coords.cuda(), cg.cuda()
    g1, g2, _ = D.altcorr_backward(a1, a2, ac, acg, r)
    h1, h2, _ = D.altcorr_backward(a1, a2, ac, acg, r)
    assert torch.equal(g1, h1) and torch.equal(g2, h2)
    r1, r2, _ = OC.altcorr_backward(f1.double(), f2.double(), coords.double(), cg.double(), r)
    for got, ref in ((g1, r1), (g2, r2)):
        err = (got.double().cpu() - ref).abs().max().item()
        assert err <= 1e-5 * max(1.0, ref.abs().max().item()), err
    # the atomic path of the same entry point
    k1 = torch.empty_like(g1)
    k2 = torch.empty_like(g2)
    lib = _lib.load()
    rc = lib.vh_altcorr_backward(_lib.VH_F32, B, N, H, W, H2, W2, C, _lib.ptr(a1), _lib.ptr(a2), _lib.ptr(ac),
                                 _lib.ptr(acg), r, _lib.ptr(k1), _lib.ptr(k2), None, _lib.stream_of(a1))
    _lib.check(rc, "altcorr_backward")
    torch.cuda.synchronize()
    assert (k1 - g1).abs().max().item() <= 1e-6 * max(1.0, g1.abs().max().item())
    assert (k2 - g2).abs().max().item() <= 1e-5 * max(1.0, g2.abs().max().item())


@pytest.mark.parametrize("dt", [torch.float32, torch.float16, torch.bfloat16, torch.float64])
def test_feature_maps_channels_last_and_pooled(dt):
    """feat_to_nhwc / pool2x2_nhwc against torch: the scaled channels-last copy is
    exact; the 2x2 means match avg_pool2d to the map type's rounding."""
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(11)
    f = torch.randn(3, 40, 13, 19, generator=g).to(dt)
    for out_dt in (torch.float32, dt):
        x = D.feat_to_nhwc(f.cuda(), 0.25, out_dtype=out_dt).cpu()
        assert x.dtype == out_dt
        assert torch.equal(x, (f.to(out_dt) * 0.25).permute(0, 2, 3, 1))
    lvl = D.feat_to_nhwc(f.cuda(), 0.25, out_dtype=dt)
    ref = f.double() * 0.25
    tol = {torch.float32: 1e-6, torch.float64: 1e-12, torch.float16: 2e-3, torch.bfloat16: 1.6e-2}[dt]
    for _ in range(3):
        lvl = D.pool2x2_nhwc(lvl)
        ref = torch.nn.functional.avg_pool2d(ref, 2, stride=2)
        assert lvl.dtype == dt and lvl.shape == ref.permute(0, 2, 3, 1).shape
        assert (lvl.cpu().double() - ref.permute(0, 2, 3, 1)).abs().max().item() <= tol


@pytest.mark.parametrize("L,H,W", [(4, 12, 16), (3, 13, 19), (1, 8, 8)])
def test_feature_grad_combination(L, H, W):
    """corr_feature_grad against its torch restatement: df1 = s * sum_l g1_l and
    df2 = s * sum_l 4^-l g2_l spread over the 2^l x 2^l blocks (the level's
    floor-dropped rows and columns get nothing), both NCHW."""
    from vipe_amd.ext import droid_net_ext as D

    g = torch.Generator().manual_seed(L * H)
    E, C = 2, 36
    g1 = [torch.randn(E, H, W, C, generator=g) for _ in range(L)]
    g2 = [torch.randn(E, H >> l, W >> l, C, generator=g) for l in range(L)]
    for dt in (torch.float32, torch.float16):
        df1, df2 = D.corr_feature_grad([t.cuda() for t in g1], [t.cuda() for t in g2], 0.25, dt)
        r1 = sum(g1).permute(0, 3, 1, 2) * 0.25
        r2 = torch.zeros(E, C, H, W)
        for lv in range(L):
            s = 2 ** lv
            up = (g2[lv].permute(0, 3, 1, 2) / 4.0 ** lv).repeat_interleave(s, 2).repeat_interleave(s, 3)
            r2[:, :, :up.shape[2], :up.shape[3]] += up
        r2 *= 0.25
        tol = 1e-6 if dt == torch.float32 else 4e-3
        assert df1.dtype == dt and (df1.cpu().float() - r1).abs().max().item() <= tol * max(1.0, r1.abs().max().item())
        assert (df2.cpu().float() - r2).abs().max().item() <= tol * max(1.0, r2.abs().max().item())


def test_altcorr_block_pyramid_keeps_the_feature_dtype():
    from vipe_amd.corr import AltCorrBlock

    g = torch.Generator().manual_seed(12)
    f = torch.randn(1, 3, 32, 12, 16, generator=g)
    for dt in (torch.float32, torch.float16):
        blk = AltCorrBlock(f.to(dt).cuda(), num_levels=4, radius=3)
        ref = f.reshape(3, 32, 12, 16).to(dt).double() / 4.0
        for i, lvl in enumerate(blk.pyramid):
            assert lvl.dtype == dt and lvl.shape == (1, 3, 12 >> i, 16 >> i, 32)
            if i > 0:
                ref = torch.nn.functional.avg_pool2d(ref, 2, stride=2)
            tol = 1e-6 if dt == torch.float32 else 2e-3
            assert (lvl.cpu().double().reshape(3, 12 >> i, 16 >> i, 32) - ref.permute(0, 2, 3, 1)).abs().max() <= tol
