"""Drop-in for ``vipe.ext.scatter_ext`` (csrc/scatter_ext/scatter.cpp:54-237):
``scatter_sum``, ``scatter_mul``, ``scatter_mean``, ``scatter_min``,
``scatter_max`` with signature ``(src, index, dim, out, dim_size)``, as
autograd functions, backed by the gfx950 ``vh_scatter`` kernels.

Semantics kept from the reference (cuda/scatter_cuda.cu:56-129):
  * ``index`` is broadcast to ``src`` (1-D indices are placed at ``dim``);
  * without ``out`` the result has ``index.max() + 1`` entries along ``dim``
    (``dim_size`` if given, 0 for an empty index), starts from the reduction's
    identity, and min / max entries that no source reached are 0;
  * a given ``out`` takes part in the reduction and is updated in place (a
    non-contiguous ``out`` is reduced into a contiguous copy, which is returned);
  * min / max also return ``arg_out`` (int64), the source position along
    ``dim`` of the result, ``src.size(dim)`` where there is none; NaN sources
    are ignored.
A 1-D index broadcast over the other axes (K >= 32 trailing elements) runs
atomic-free over a stable sort of the index: the same bits on every run, sums
and products taken in ascending source order, 8 / 16-bit floats accumulated in
FP32 and rounded once.  Other index shapes use atomics.
Differences, by design: ties in ``arg_out`` resolve to the smallest position
(the reference keeps whichever write lands last); min / max values do not
depend on the order of the atomics; indices outside ``[0, out.size(dim))``
are skipped rather than written out of bounds; 8-bit products wrap within
their own byte.
"""
from __future__ import annotations

import torch

from .. import _lib

SUM, MUL, MIN, MAX = 0, 1, 2, 3


def broadcast(src, other, dim):
    """The reference's index broadcast (scatter.cpp:20-26)."""
    if dim < 0:
        dim = other.dim() + dim
    if src.dim() == 1:
        for _ in range(dim):
            src = src.unsqueeze(0)
    for _ in range(src.dim(), other.dim()):
        src = src.unsqueeze(-1)
    return src.expand(other.size())


def _collapse(sizes, strides):
    """One stride s with offset = flat_index * s over these dims, or None."""
    dims = [(n, s) for n, s in zip(sizes, strides) if n != 1]
    if not dims:
        return 0
    s = dims[-1][1]
    span = 1
    for n, st in reversed(dims):
        if st != s * span:
            return None
        span *= n
    return s


def _index_strides(index, dim):
    sz, st = list(index.shape), list(index.stride())
    sB = _collapse(sz[:dim], st[:dim])
    sK = _collapse(sz[dim + 1:], st[dim + 1:])
    if sB is None or sK is None:
        index = index.contiguous()
        sz, st = list(index.shape), list(index.stride())
        sB = _collapse(sz[:dim], st[:dim])
        sK = _collapse(sz[dim + 1:], st[dim + 1:])
    return index, sB, st[dim], sK


def _scatter(src, index, dim, out, dim_size, reduce):
    """scatter_cuda (scatter_cuda.cu:56-129) on the gfx950 kernels."""
    if src.device.type != "cuda":
        raise RuntimeError("src must be CUDA tensor")
    if index.device.type != "cuda":
        raise RuntimeError("index must be CUDA tensor")
    if out is not None and out.device.type != "cuda":
        raise RuntimeError("optional_out must be CUDA tensor")
    if src.dim() != index.dim() or any(src.size(i) < index.size(i) for i in range(index.dim() - 1)):
        raise RuntimeError("Input mismatch")
    if index.dtype != torch.int64:
        raise RuntimeError(f"index must be int64 (got {index.dtype})")
    code = _lib.dtype_code(src)
    if code < 0 or code == _lib.VH_FP8E4M3:
        raise RuntimeError(f"scatter: unsupported dtype {src.dtype}")
    src = src.contiguous()
    out_given = out is not None
    if out_given:
        if out.dtype != src.dtype:
            raise RuntimeError("out must have the dtype of src")
        out = out.contiguous()
        if out.dim() != src.dim() or any(src.size(i) != out.size(i) for i in range(out.dim()) if i != dim):
            raise RuntimeError("Input mismatch")
    else:
        sizes = list(src.shape)
        if dim_size is not None:
            sizes[dim] = int(dim_size)
        elif index.numel() == 0:
            sizes[dim] = 0
        else:
            sizes[dim] = int(index.max().item()) + 1
        out = torch.empty(sizes, dtype=src.dtype, device=src.device)
    arg_out = None
    if reduce in (MIN, MAX):
        arg_out = torch.empty(out.shape, dtype=torch.int64, device=src.device)
    if src.numel() == 0:
        if not out_given:
            out.zero_()
        if arg_out is not None:
            arg_out.fill_(src.size(dim))
        return out, arg_out
    B = 1
    for i in range(dim):
        B *= src.size(i)
    E = src.size(dim)
    K = src.numel() // (B * E)
    N = out.size(dim)
    index, sB, sE, sK = _index_strides(index, dim)
    lib = _lib.load()
    ws_bytes = lib.vh_scatter_workspace_size(B, E, K, sB, sE, sK, N)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=src.device) if ws_bytes else None
    with torch.cuda.device(src.device):
        rc = lib.vh_scatter(reduce, code, _lib.ptr(src), _lib.ptr(index), B, E, K, sB, sE, sK, N, _lib.ptr(out),
                            _lib.ptr(arg_out), int(out_given), _lib.ptr(ws), ws_bytes, _lib.stream_of(src))
    _lib.check(rc, "scatter")
    return out, arg_out


def _norm_dim(dim, src):
    return dim + src.dim() if dim < 0 else dim


def _dirty(ctx, out, result):
    if out is not None and result.data_ptr() == out.data_ptr():
        ctx.mark_dirty(out)
        return out
    return result


class ScatterSum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, src, index, dim, out, dim_size):
        dim = _norm_dim(dim, src)
        index = broadcast(index, src, dim)
        res, _ = _scatter(src, index, dim, out, dim_size, SUM)
        ctx.dim = dim
        ctx.save_for_backward(index)
        return _dirty(ctx, out, res)

    @staticmethod
    def backward(ctx, grad_out):
        (index,) = ctx.saved_tensors
        return torch.gather(grad_out, ctx.dim, index), None, None, None, None


class ScatterMul(torch.autograd.Function):
    @staticmethod
    def forward(ctx, src, index, dim, out, dim_size):
        dim = _norm_dim(dim, src)
        index = broadcast(index, src, dim)
        res, _ = _scatter(src, index, dim, out, dim_size, MUL)
        ctx.dim = dim
        ctx.save_for_backward(src, index, res)
        return _dirty(ctx, out, res)

    @staticmethod
    def backward(ctx, grad_out):
        src, index, res = ctx.saved_tensors
        g = torch.gather(grad_out * res, ctx.dim, index).div_(src)
        g.masked_fill_(g.isnan(), 0)
        return g, None, None, None, None


class ScatterMean(torch.autograd.Function):
    @staticmethod
    def forward(ctx, src, index, dim, out, dim_size):
        dim = _norm_dim(dim, src)
        old_index = index
        index = broadcast(index, src, dim)
        res, _ = _scatter(src, index, dim, out, dim_size, SUM)
        ones = torch.ones(old_index.shape, dtype=src.dtype, device=src.device)
        cdim = old_index.dim() - 1 if old_index.dim() <= dim else dim
        count, _ = _scatter(ones, old_index, cdim, None, res.size(dim), SUM)
        count.masked_fill_(count < 1, 1)
        count = broadcast(count, res, dim)
        if res.is_floating_point():
            res.true_divide_(count)
        else:
            res.div_(count, rounding_mode="floor")
        ctx.dim = dim
        ctx.save_for_backward(index, count)
        return _dirty(ctx, out, res)

    @staticmethod
    def backward(ctx, grad_out):
        index, count = ctx.saved_tensors
        count = torch.gather(count, ctx.dim, index)
        g = torch.gather(grad_out, ctx.dim, index)
        g.true_divide_(count)
        return g, None, None, None, None


class _ScatterArg(torch.autograd.Function):
    REDUCE = MIN

    @classmethod
    def _fwd(cls, ctx, src, index, dim, out, dim_size):
        dim = _norm_dim(dim, src)
        index = broadcast(index, src, dim)
        res, arg = _scatter(src, index, dim, out, dim_size, cls.REDUCE)
        ctx.dim = dim
        ctx.src_shape = list(src.shape)
        ctx.save_for_backward(index, arg)
        ctx.mark_non_differentiable(arg)
        return _dirty(ctx, out, res), arg

    @staticmethod
    def backward(ctx, grad_out, grad_arg):
        index, arg = ctx.saved_tensors
        shape = list(ctx.src_shape)
        shape[ctx.dim] += 1
        g = torch.zeros(shape, dtype=grad_out.dtype, device=grad_out.device)
        g.scatter_(ctx.dim, arg, grad_out)
        return g.narrow(ctx.dim, 0, shape[ctx.dim] - 1), None, None, None, None


class ScatterMin(_ScatterArg):
    REDUCE = MIN

    @staticmethod
    def forward(ctx, src, index, dim, out, dim_size):
        return ScatterMin._fwd(ctx, src, index, dim, out, dim_size)


class ScatterMax(_ScatterArg):
    REDUCE = MAX

    @staticmethod
    def forward(ctx, src, index, dim, out, dim_size):
        return ScatterMax._fwd(ctx, src, index, dim, out, dim_size)


def scatter_sum(src, index, dim, out=None, dim_size=None):
    return ScatterSum.apply(src, index, dim, out, dim_size)


def scatter_mul(src, index, dim, out=None, dim_size=None):
    return ScatterMul.apply(src, index, dim, out, dim_size)


def scatter_mean(src, index, dim, out=None, dim_size=None):
    return ScatterMean.apply(src, index, dim, out, dim_size)


def scatter_min(src, index, dim, out=None, dim_size=None):
    res, arg = ScatterMin.apply(src, index, dim, out, dim_size)
    return res, arg


def scatter_max(src, index, dim, out=None, dim_size=None):
    res, arg = ScatterMax.apply(src, index, dim, out, dim_size)
    return res, arg
