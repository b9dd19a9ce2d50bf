"""Restatement of scatter_ext's reductions (oracle; test-only).

Follows csrc/scatter_ext/cuda/scatter_cuda.cu:56-129 and reducer.cuh:
  src viewed as [B, E, K] around dim, out as [B, N, K]; out starts from the
  reduction's identity (0, 1, dtype max for min, dtype lowest for max) unless
  given; every (b, e, k) combines src into out[b, index, k]; min / max without a
  given out turn entries equal to the identity into 0; arg_out is an e whose
  src equals the result (this restatement takes the smallest), else E.
  scatter_mean: scatter.cpp:107-145 (sum, then a count scattered from the
  un-broadcast index, clamped to >= 1, true division for floats, floor
  division for integers).
Floats are reduced in float64 (callers compare with a tolerance for sums and
products; min / max select an input exactly); integers in their own dtype
with wrap-around.  NaN sources are ignored by min / max (fmin / fmax).
Pinned by tests/test_oracle_scatter.py against torch.Tensor.scatter_reduce_
(sum / prod / amin / amax / mean), the reference docstring's shape example,
and first-occurrence argmin / argmax per segment.
"""
from __future__ import annotations

import numpy as np


def broadcast(index, shape, dim):
    """vipe/ext/scatter.py:12-21: a 1-D index goes to dim, trailing dims appended."""
    index = np.asarray(index)
    nd = len(shape)
    if dim < 0:
        dim += nd
    if index.ndim == 1:
        index = index.reshape((1,) * dim + index.shape)
    index = index.reshape(index.shape + (1,) * (nd - index.ndim))
    return np.broadcast_to(index, shape)


def identity(dtype, reduce):
    dtype = np.dtype(dtype)
    if reduce in ("sum", "mean"):
        return dtype.type(0)
    if reduce == "mul":
        return dtype.type(1)
    info = np.finfo(dtype) if dtype.kind == "f" else np.iinfo(dtype)
    return dtype.type(info.max if reduce == "min" else info.min)


def scatter(src, index, dim, out=None, dim_size=None, reduce="sum", compute_dtype=None):
    """(out, arg_out or None).  compute_dtype: the dtype of the reduction
    (float64 for float inputs unless given)."""
    src = np.asarray(src)
    nd = src.ndim
    dim = dim + nd if dim < 0 else dim
    idx = broadcast(index, src.shape, dim)
    cdt = compute_dtype or (np.float64 if src.dtype.kind == "f" else src.dtype)
    if out is None:
        size = list(src.shape)
        size[dim] = dim_size if dim_size is not None else (0 if idx.size == 0 else int(idx.max()) + 1)
        res = np.full(size, identity(src.dtype, reduce), dtype=src.dtype).astype(cdt)
        given = False
    else:
        res = np.array(out, dtype=cdt, copy=True)
        given = True
    B = int(np.prod(src.shape[:dim], dtype=np.int64))
    E = src.shape[dim]
    K = int(np.prod(src.shape[dim + 1:], dtype=np.int64))
    N = res.shape[dim]
    s3 = src.astype(cdt).reshape(B, E, K)
    i3 = np.ascontiguousarray(idx).reshape(B, E, K)
    r3 = res.reshape(B, N, K)
    bb, ee, kk = np.meshgrid(np.arange(B), np.arange(E), np.arange(K), indexing="ij")
    ok = (i3 >= 0) & (i3 < N)
    at = (bb[ok], i3[ok], kk[ok])
    vals = s3[ok]
    with np.errstate(over="ignore", invalid="ignore"):
        if reduce in ("sum", "mean"):
            np.add.at(r3, at, vals)
        elif reduce == "mul":
            np.multiply.at(r3, at, vals)
        elif reduce == "min":
            np.fmin.at(r3, at, vals)
        elif reduce == "max":
            np.fmax.at(r3, at, vals)
        else:
            raise ValueError(reduce)
    arg = None
    if reduce in ("min", "max"):
        if not given:
            r3[r3 == identity(src.dtype, reduce)] = 0
        a3 = np.full((B, N, K), E, dtype=np.int64)
        hit = ok & (s3 == r3[bb, np.clip(i3, 0, max(N - 1, 0)), kk]) if N > 0 else ok & False
        np.minimum.at(a3, (bb[hit], i3[hit], kk[hit]), ee[hit])
        arg = a3.reshape(res.shape)
    if reduce == "mean":
        ix = np.asarray(index)
        cdim = ix.ndim - 1 if ix.ndim <= dim else dim
        cnt, _ = scatter(np.ones(ix.shape, dtype=src.dtype), ix, cdim, None, N, "sum")
        cnt = np.maximum(cnt, 1)
        cnt = broadcast(cnt, res.shape, dim)
        if src.dtype.kind == "f":
            res = res / cnt
        else:
            res = np.floor_divide(res, cnt)
    return res.reshape(res.shape), arg
