"""gfx950 scatter_ext reductions vs the oracle.

Integers (wrap-around sums and products, min / max, floor means) and every
min / max / arg_out are bit-exact.  Float sums, products and means are exact
up to the order of the atomics: each of the n terms of a bucket adds at most
one rounding of the output type, so the bound is n * eps(type) * (sum of |terms|)
for sums (n * eps * |product| for products), with n the bucket's term count."""
import zlib

import numpy as np
import pytest
import torch

from oracle import scatter as OSc

pytestmark = pytest.mark.gpu

FLOATS = [torch.float32, torch.float64, torch.float16, torch.bfloat16]
INTS = [torch.int64, torch.int32, torch.int16, torch.int8, torch.uint8]
EPS = {torch.float32: 2 ** -24, torch.float64: 2 ** -53, torch.float16: 2 ** -11, torch.bfloat16: 2 ** -8}


def _np(t):
    return t.cpu().numpy() if t.dtype != torch.bfloat16 else t.cpu().float().numpy()


def _src(shape, dtype, g, reduce):
    if dtype.is_floating_point:
        lo, hi = (0.5, 1.5) if reduce == "mul" else (-2.0, 2.0)
        return (torch.rand(shape, generator=g, dtype=torch.float64) * (hi - lo) + lo).to(dtype)
    lo, hi = (0, 200) if dtype == torch.uint8 else (-100, 100)
    if reduce == "mul":
        lo, hi = (1, 9) if dtype == torch.uint8 else (-9, 9)
    return torch.randint(lo, hi, shape, generator=g).to(dtype)


def _check(reduce, dtype, got, ref, src_np, idx, dim, dim_size, out_np=None):
    got = _np(got).astype(np.float64) if dtype.is_floating_point else _np(got)
    if not dtype.is_floating_point or reduce in ("min", "max"):
        assert np.array_equal(got, ref.astype(got.dtype), equal_nan=True)
        return
    cnt, _ = OSc.scatter(np.ones(src_np.shape), idx, dim, None, dim_size, "sum")
    n = cnt.max() + (out_np is not None) + 1
    if reduce == "mul":
        bound = n * EPS[dtype] * np.abs(ref)
    else:
        mag, _ = OSc.scatter(np.abs(src_np), idx, dim, None if out_np is None else np.abs(out_np), dim_size, "sum")
        bound = n * EPS[dtype] * mag + (EPS[dtype] * np.abs(ref) if reduce == "mean" else 0)
    assert np.all(np.abs(got - ref) <= bound + 1e-300), np.abs(got - ref).max()


@pytest.mark.parametrize("K", [17, 40, 36])   # 17: atomics; 40, 36: the segmented path (36: 8-wide types unvectorised)
@pytest.mark.parametrize("reduce", ["sum", "mul", "min", "max", "mean"])
@pytest.mark.parametrize("dtype", FLOATS + INTS)
def test_broadcast_index_matches_oracle(reduce, dtype, K):
    from vipe_amd.ext import scatter

    g = torch.Generator().manual_seed(zlib.crc32(f"{reduce}{dtype}{K}".encode()) % 1000)
    src = _src((3, 40, K), dtype, g, reduce)
    idx = torch.randint(0, 9, (40,), generator=g)
    fn = getattr(scatter, f"scatter_{reduce}")
    res = fn(src.cuda(), idx.cuda(), 1, None, 11)
    src_np = _np(src)
    ref, arg = OSc.scatter(src_np, idx.numpy(), 1, None, 11, reduce)
    if reduce in ("min", "max"):
        res, got_arg = res
        assert np.array_equal(got_arg.cpu().numpy(), arg)
    assert res.dtype == dtype and tuple(res.shape) == (3, 11, K)
    _check(reduce, dtype, res, ref, src_np.astype(np.float64) if dtype.is_floating_point else src_np,
           idx.numpy(), 1, 11)


@pytest.mark.parametrize("reduce", ["sum", "mul", "min", "max"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.int32, torch.int8])
def test_full_index_and_given_out(reduce, dtype):
    """An index per element (no broadcast), last-dim scatter, and an out that
    takes part in the reduction and is updated in place."""
    from vipe_amd.ext import scatter_ext

    g = torch.Generator().manual_seed(17)
    src = _src((5, 6, 33), dtype, g, reduce)
    idx = torch.randint(0, 7, (5, 6, 33), generator=g)
    out = _src((5, 6, 7), dtype, g, reduce).cuda()
    out0 = _np(out)
    fn = getattr(scatter_ext, f"scatter_{reduce}")
    res = fn(src.cuda(), idx.cuda(), -1, out, None)
    if reduce in ("min", "max"):
        res, got_arg = res
    assert res.data_ptr() == out.data_ptr()
    src_np = _np(src)
    ref, arg = OSc.scatter(src_np, idx.numpy(), 2, out0, None, reduce)
    if reduce in ("min", "max"):
        assert np.array_equal(got_arg.cpu().numpy(), arg)
    _check(reduce, dtype, out, ref, src_np.astype(np.float64) if dtype.is_floating_point else src_np,
           idx.numpy(), 2, 7, out_np=out0)


def test_partial_index_rank_and_large_sum():
    """index [B, E] broadcast over a trailing K; a 2M-element float32 sum."""
    from vipe_amd.ext import scatter

    g = torch.Generator().manual_seed(3)
    src = torch.randn(4, 500, 8, generator=g)
    idx = torch.randint(0, 50, (4, 500), generator=g)
    res = scatter.scatter_add(src.cuda(), idx.cuda(), 1)
    ref, _ = OSc.scatter(src.numpy(), idx.numpy(), 1, reduce="sum")
    assert np.abs(res.cpu().numpy() - ref).max() < 1e-4
    big = torch.randn(2_000_000, generator=g)
    bidx = torch.randint(0, 1000, (2_000_000,), generator=g)
    res = scatter.scatter(big.cuda(), bidx.cuda(), 0, reduce="sum")
    ref, _ = OSc.scatter(big.numpy(), bidx.numpy(), 0, reduce="sum")
    mag, _ = OSc.scatter(np.abs(big.numpy()), bidx.numpy(), 0, reduce="sum")
    assert np.all(np.abs(res.cpu().numpy() - ref) <= 4000 * 2 ** -24 * mag)


def test_min_max_edge_cases_are_deterministic():
    from vipe_amd.ext import scatter_ext

    nan = float("nan")
    fmax = torch.finfo(torch.float32).max
    src = torch.tensor([3.0, -0.0, 0.0, nan, 5.0, 5.0, fmax, -1.0], device="cuda")
    idx = torch.tensor([0, 1, 1, 2, 3, 3, 4, 6], device="cuda")
    mn, amn = scatter_ext.scatter_min(src, idx, 0, None, 8)
    assert mn.cpu().tolist() == [3.0, 0.0, 0.0, 5.0, 0.0, 0.0, -1.0, 0.0] and torch.signbit(mn[1]).item()
    # NaN only: empty (0, arg E); FLT_MAX reads as the identity and is zeroed
    assert amn.cpu().tolist() == [0, 1, 8, 4, 8, 8, 7, 8]
    mx, amx = scatter_ext.scatter_max(src, idx, 0, None, 8)
    assert mx.cpu().tolist() == [3.0, 0.0, 0.0, 5.0, fmax, 0.0, -1.0, 0.0] and not torch.signbit(mx[1]).item()
    # -0.0 == +0.0 as values, so the smaller position names the +0.0 maximum
    assert amx.cpu().tolist() == [0, 1, 8, 4, 6, 8, 7, 8]
    # a NaN already in a given out gives way to the sources
    out = torch.tensor([nan, nan], device="cuda")
    r, a = scatter_ext.scatter_max(torch.tensor([1.0, 2.0], device="cuda"), torch.tensor([0, 0], device="cuda"),
                                   0, out, None)
    assert torch.isnan(r[1]).item() and r[0].item() == 2.0 and a.cpu().tolist() == [1, 2]
    # repeated runs give identical bits
    big = torch.randint(-5, 5, (100000,), device="cuda").float()
    bidx = torch.randint(0, 100, (100000,), device="cuda")
    r1 = scatter_ext.scatter_max(big, bidx, 0)
    r2 = scatter_ext.scatter_max(big, bidx, 0)
    assert torch.equal(r1[0], r2[0]) and torch.equal(r1[1], r2[1])


@pytest.mark.parametrize("reduce", ["min", "max", "sum"])
def test_segmented_path_equals_atomic_path(reduce):
    """The same reduction through a broadcast 1-D index (segmented path) and
    through the materialised index (atomics): min / max / arg_out bit-equal,
    including -0.0 / +0.0, NaN, the identity value and a given out holding NaN;
    the segmented sum is bit-identical from run to run."""
    from vipe_amd.ext import scatter_ext

    nan = float("nan")
    fmax = torch.finfo(torch.float32).max
    col = torch.tensor([3.0, -0.0, 0.0, nan, 5.0, 5.0, fmax, -1.0, 2.0, -7.0])
    idx = torch.tensor([0, 1, 1, 2, 3, 3, 4, 6, 0, 6], device="cuda")
    g = torch.Generator().manual_seed(1)
    src = (col[:, None] * torch.ones(1, 64)).clone()
    src[8:] = torch.randn(2, 64, generator=g)
    src = src.cuda()
    full = idx[:, None].expand(10, 64).contiguous()
    fn = getattr(scatter_ext, f"scatter_{reduce}")
    for out0 in (None, torch.full((8, 64), nan, device="cuda")):
        a = fn(src, idx, 0, None if out0 is None else out0.clone(), 8)
        b = fn(src, full, 0, None if out0 is None else out0.clone(), 8)
        if reduce == "sum":
            assert torch.allclose(a, b, equal_nan=True, atol=1e-6)
            r1, r2 = fn(src, idx, 0, None, 8), fn(src, idx, 0, None, 8)
            assert torch.equal(r1.view(torch.int32), r2.view(torch.int32))   # bits: bucket 2 holds a NaN
        else:
            assert torch.equal(a[0].view(torch.int32), b[0].view(torch.int32))
            assert torch.equal(a[1], b[1])


@pytest.mark.parametrize("reduce", ["sum", "max"])
def test_segmented_path_many_sources(reduce):
    """E = 5000 sources: the placement pass reads the first 4096 indices from
    its LDS copy and the rest from memory; out-of-range indices are skipped."""
    from vipe_amd.ext import scatter_ext

    g = torch.Generator().manual_seed(8)
    src = torch.randn(2, 5000, 32, generator=g)
    idx = torch.randint(-3, 300, (5000,), generator=g)
    fn = getattr(scatter_ext, f"scatter_{reduce}")
    res = fn(src.cuda(), idx.cuda(), 1, None, 300)
    ref, arg = OSc.scatter(src.numpy(), idx.numpy(), 1, None, 300, reduce)
    if reduce == "max":
        assert np.array_equal(res[0].cpu().numpy(), ref.astype(np.float32))
        assert np.array_equal(res[1].cpu().numpy(), arg)
    else:
        mag, _ = OSc.scatter(np.abs(src.numpy()), idx.numpy(), 1, None, 300, "sum")
        assert np.all(np.abs(res.cpu().numpy() - ref) <= 64 * 2 ** -24 * mag + 1e-30)


def test_out_of_range_indices_are_skipped_and_empty_src():
    from vipe_amd.ext import scatter_ext

    src = torch.tensor([1.0, 2.0, 4.0], device="cuda")
    idx = torch.tensor([0, 9, -1], device="cuda")
    assert scatter_ext.scatter_sum(src, idx, 0, None, 2).cpu().tolist() == [1.0, 0.0]
    e = torch.zeros(0, 4, device="cuda")
    r, a = scatter_ext.scatter_max(e, torch.zeros(0, dtype=torch.long, device="cuda"), 0, None, 3)
    assert r.shape == (3, 4) and torch.all(r == 0) and torch.all(a == 0)


def test_autograd_matches_reference_formulas():
    from vipe_amd.ext import scatter

    g = torch.Generator().manual_seed(4)
    src = (torch.rand(6, 20, 3, generator=g, dtype=torch.float64) + 0.5).cuda().requires_grad_()
    idx = torch.randint(0, 5, (20,), generator=g).cuda()
    w = torch.randn(6, 5, 3, generator=g, dtype=torch.float64).cuda()
    for name in ("sum", "mean", "mul"):
        fn = getattr(scatter, f"scatter_{name}")
        assert torch.autograd.gradcheck(lambda s: fn(s, idx, 1, None, 5), (src,))
    out, arg = scatter.scatter_max(src, idx, 1, None, 5)
    (out * w).sum().backward()
    expect = torch.zeros(6, 21, 3, dtype=torch.float64, device="cuda")
    expect.scatter_(1, arg, w)
    assert torch.equal(src.grad, expect[:, :20])


@pytest.mark.parametrize("reduce", ["sum", "min", "max"])
def test_segmented_path_with_given_out(reduce):
    """A broadcast 1-D index (segmented path) into a given out that takes part
    in the reduction, against the oracle."""
    from vipe_amd.ext import scatter_ext

    g = torch.Generator().manual_seed(12)
    src = torch.randn(2, 50, 64, generator=g)
    idx = torch.randint(0, 9, (50,), generator=g)
    out = torch.randn(2, 11, 64, generator=g).cuda()
    out0 = out.cpu().numpy()
    fn = getattr(scatter_ext, f"scatter_{reduce}")
    res = fn(src.cuda(), idx.cuda(), 1, out, None)
    ref, arg = OSc.scatter(src.numpy(), idx.numpy(), 1, out0, None, reduce)
    if reduce == "sum":
        assert res.data_ptr() == out.data_ptr()
        mag, _ = OSc.scatter(np.abs(src.numpy()), idx.numpy(), 1, np.abs(out0), None, "sum")
        assert np.all(np.abs(out.cpu().numpy() - ref) <= 16 * 2 ** -24 * mag + 1e-30)
    else:
        assert res[0].data_ptr() == out.data_ptr()
        assert np.array_equal(out.cpu().numpy(), ref.astype(np.float32))
        assert np.array_equal(res[1].cpu().numpy(), arg)
