"""Drop-in for ``vipe.ext.lietorch_ext`` (csrc/lietorch_ext/lietorch.cpp:23-335).

Same function names, positional arguments and return conventions:
forward ops return a tensor, ``*_backward`` return a list of gradients,
``TORCH_CHECK(x.is_contiguous())`` becomes ``RuntimeError("x must be
contiguous")``, and an unknown group id or dtype gives what the reference's launchers give:
``None`` for expm / logm / projector (an undefined tensor) and the
preallocated zeros for every other op (lietorch_gpu.cu:302-562,
dispatch.h:40-50).  CUDA(HIP) tensors
run on gfx950 kernels on the current stream; CPU tensors run the same element
functors on the host (the reference's lietorch_cpu.cpp path).
"""
from __future__ import annotations

import contextlib

import torch

from .. import _lib

# group id -> (embedding N, tangent K)
DIMS = {1: (4, 3), 2: (5, 4), 3: (7, 6), 4: (8, 7)}


def _contig(**named):
    for k, v in named.items():
        if not v.is_contiguous():
            raise RuntimeError(f"{k} must be contiguous")


def _dev(t):
    return torch.cuda.device(t.device) if t.device.type == "cuda" else contextlib.nullcontext()


def _dims(group_index):
    return DIMS.get(int(group_index), (None, None))


def _supported(group_index, t):
    return DIMS.get(int(group_index)) is not None and t.dtype in (torch.float32, torch.float64)


def _call(fn_name, group_index, ref, *ptrs):
    """Run one op; returns False when the reference would dispatch to nothing
    (unknown group id or dtype), leaving the preallocated outputs untouched."""
    if not _supported(group_index, ref):
        return False
    lib = _lib.load()
    with _dev(ref):
        rc = getattr(lib, fn_name)(int(group_index), _lib.dtype_code(ref), ref.shape[0], *ptrs,
                                   _lib.place_of(ref), _lib.stream_of(ref))
    _lib.check(rc, fn_name)
    return True


def _z(shape, like):
    return torch.zeros(tuple(shape), dtype=like.dtype, device=like.device)


# ---- forward ops (lietorch_gpu.cu:302-562 allocation conventions) -----------

def expm(group_index, a):
    _contig(a=a)
    N, K = _dims(group_index)
    if not _supported(group_index, a):
        return None  # undefined tensor in the reference
    X = _z((a.shape[0], N), a)
    _call("vh_lie_expm", group_index, a, _lib.ptr(a), _lib.ptr(X))
    return X


def logm(group_index, X):
    _contig(X=X)
    N, K = _dims(group_index)
    if not _supported(group_index, X):
        return None
    a = _z((X.shape[0], K), X)
    _call("vh_lie_logm", group_index, X, _lib.ptr(X), _lib.ptr(a))
    return a


def projector(group_index, X):
    _contig(X=X)
    N, K = _dims(group_index)
    if not _supported(group_index, X):
        return None
    P = _z((X.shape[0], N, N), X)
    _call("vh_lie_projector", group_index, X, _lib.ptr(X), _lib.ptr(P))
    return P


def inv(group_index, X):
    _contig(X=X)
    Y = torch.zeros_like(X)
    _call("vh_lie_inv", group_index, X, _lib.ptr(X), _lib.ptr(Y))
    return Y


def as_matrix(group_index, X):
    _contig(X=X)
    T = _z((X.shape[0], 4, 4), X)
    _call("vh_lie_as_matrix", group_index, X, _lib.ptr(X), _lib.ptr(T))
    return T


def mul(group_index, X, Y):
    _contig(X=X, Y=Y)
    Z = torch.zeros_like(X)
    _call("vh_lie_mul", group_index, X, _lib.ptr(X), _lib.ptr(Y), _lib.ptr(Z))
    return Z


def _xv(fn_name, group_index, X, v):
    out = _z(v.shape, v)
    _call(fn_name, group_index, X, _lib.ptr(X), _lib.ptr(v), _lib.ptr(out))
    return out


def adj(group_index, X, a):
    _contig(X=X, a=a)
    return _xv("vh_lie_adj", group_index, X, a)


def adjT(group_index, X, a):
    _contig(X=X, a=a)
    return _xv("vh_lie_adjT", group_index, X, a)


def act(group_index, X, p):
    _contig(X=X, p=p)
    return _xv("vh_lie_act", group_index, X, p)


def act4(group_index, X, p):
    _contig(X=X, p=p)
    return _xv("vh_lie_act4", group_index, X, p)


def Jinv(group_index, X, a):
    _contig(X=X, a=a)
    return _xv("vh_lie_Jinv", group_index, X, a)


# ---- backward ops -----------------------------------------------------------

def _zg(shape, ref, grad):
    return torch.zeros(tuple(shape), dtype=grad.dtype, device=ref.device)


def expm_backward(group_index, grad, a):
    _contig(a=a, grad=grad)
    da = _zg(a.shape, a, grad)
    _call("vh_lie_expm_backward", group_index, a, _lib.ptr(grad), _lib.ptr(a), _lib.ptr(da))
    return [da]


def logm_backward(group_index, grad, X):
    _contig(X=X, grad=grad)
    dX = _zg(X.shape, X, grad)
    _call("vh_lie_logm_backward", group_index, X, _lib.ptr(grad), _lib.ptr(X), _lib.ptr(dX))
    return [dX]


def inv_backward(group_index, grad, X):
    _contig(X=X, grad=grad)
    dX = _zg(X.shape, X, grad)
    _call("vh_lie_inv_backward", group_index, X, _lib.ptr(grad), _lib.ptr(X), _lib.ptr(dX))
    return [dX]


def mul_backward(group_index, grad, X, Y):
    _contig(X=X, Y=Y, grad=grad)
    dX = _zg(X.shape, X, grad)
    dY = _zg(Y.shape, X, grad)
    _call("vh_lie_mul_backward", group_index, X, _lib.ptr(grad), _lib.ptr(X), _lib.ptr(Y), _lib.ptr(dX),
          _lib.ptr(dY))
    return [dX, dY]


def _bwd2(fn_name, group_index, grad, X, v):
    dX = _zg(X.shape, X, grad)
    dv = _zg(v.shape, X, grad)
    _call(fn_name, group_index, X, _lib.ptr(grad), _lib.ptr(X), _lib.ptr(v), _lib.ptr(dX), _lib.ptr(dv))
    return [dX, dv]


def adj_backward(group_index, grad, X, a):
    _contig(X=X, a=a, grad=grad)
    return _bwd2("vh_lie_adj_backward", group_index, grad, X, a)


def adjT_backward(group_index, grad, X, a):
    _contig(X=X, a=a, grad=grad)
    return _bwd2("vh_lie_adjT_backward", group_index, grad, X, a)


def act_backward(group_index, grad, X, p):
    _contig(X=X, p=p, grad=grad)
    return _bwd2("vh_lie_act_backward", group_index, grad, X, p)


def act4_backward(group_index, grad, X, p):
    _contig(X=X, p=p, grad=grad)
    return _bwd2("vh_lie_act4_backward", group_index, grad, X, p)
