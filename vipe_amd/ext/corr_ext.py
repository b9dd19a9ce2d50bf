"""Drop-in for ``vipe.ext.corr_ext``: the spatial correlation sampler
(csrc/corr_ext/correlation_sampler.cpp:41-85), backed by gfx950 kernels in
libvipe_hip.so.

``forward(input1, input2, kH, kW, patchH, patchW, padH, padW, dilationH,
dilationW, dilation_patchH, dilation_patchW, dH, dW)`` -> [N, patchH, patchW, oH, oW]
``backward(input1, input2, grad_output, <same ints>)`` -> [grad_input1, grad_input2]

Inputs are [N, C, H, W] float32 / float64 / float16 / bfloat16 device tensors.
The same checks as the reference raise RuntimeError with its messages
("input1 must be a CUDA tensor", "input1 must be contiguous", "input1 is not
on same device as input2").  The reference's CPU path has no counterpart
here: host tensors raise.  Sums accumulate in FP32 (FP64 for float64 inputs),
where the reference accumulates half inputs in half.  Kernel size 1 with
stride 1 (the local-attention form) runs on the matrix cores, forward and
backward, over channels-last copies of the inputs (a workspace allocated per
call).
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _lib


def _require(name, t):
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _same_device(a_name, a, b_name, b):
    if a.device != b.device:
        raise RuntimeError(f"{a_name} is not on same device as {b_name}")


def _dtype(t):
    code = _lib.dtype_code(t)
    if code < 0:
        raise RuntimeError(f"corr_ext: unsupported dtype {t.dtype}")
    return code


def output_size(H, W, kH, kW, padH, padW, dilationH, dilationW, dH, dW):
    oH, oW = C.c_int(0), C.c_int(0)
    rc = _lib.load().vh_spatial_corr_out_size(H, W, kH, kW, padH, padW, dilationH, dilationW, dH, dW,
                                               C.byref(oH), C.byref(oW))
    _lib.check(rc, "corr_ext.output_size")
    return oH.value, oW.value


def _inputs(input1, input2):
    _require("input1", input1)
    _require("input2", input2)
    _same_device("input1", input1, "input2", input2)
    if input1.dim() != 4 or input1.shape != input2.shape:
        raise RuntimeError(f"input1 and input2 must both be [N,C,H,W] of one shape "
                           f"(got {tuple(input1.shape)} and {tuple(input2.shape)})")
    if input1.dtype != input2.dtype:
        raise RuntimeError("input1 and input2 must have the same dtype")


def forward(input1, input2, kH, kW, patchH, patchW, padH, padW, dilationH, dilationW, dilation_patchH,
            dilation_patchW, dH, dW):
    """Correlation of every output position's kH x kW window of input1 with the
    same window of input2 shifted over a patchH x patchW grid of offsets."""
    _inputs(input1, input2)
    N, Cc, H, W = input1.shape
    oH, oW = output_size(H, W, kH, kW, padH, padW, dilationH, dilationW, dH, dW)
    if oH <= 0 or oW <= 0:
        raise RuntimeError(f"corr_ext.forward: empty output size ({oH} x {oW})")
    out = torch.empty((N, patchH, patchW, oH, oW), dtype=input1.dtype, device=input1.device)
    if out.numel() == 0:
        return out
    lib = _lib.load()
    code = _dtype(input1)
    ints = (N, Cc, H, W, kH, kW, patchH, patchW, padH, padW, dilationH, dilationW, dilation_patchH, dilation_patchW,
            dH, dW)
    ws_bytes = lib.vh_spatial_corr_workspace_size(code, *ints)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=input1.device) if ws_bytes else None
    with torch.cuda.device(input1.device):
        rc = lib.vh_spatial_corr_forward(code, _lib.ptr(input1), _lib.ptr(input2), *ints, _lib.ptr(out),
                                         _lib.ptr(ws), ws_bytes, _lib.stream_of(input1))
    _lib.check(rc, "corr_ext.forward")
    return out


def backward(input1, input2, grad_output, kH, kW, patchH, patchW, padH, padW, dilationH, dilationW,
             dilation_patchH, dilation_patchW, dH, dW):
    """[grad_input1, grad_input2] of the forward's sum for grad_output
    [N, patchH, patchW, oH, oW]; the patch is centred at
    ((patch-1)/2) * dilation_patch as in correlation_cuda_kernel.cu:96-97."""
    _inputs(input1, input2)
    if grad_output.device.type != "cuda":
        raise RuntimeError("grad_output must be a CUDA tensor")
    _same_device("input1", input1, "grad_output", grad_output)
    N, Cc, H, W = input1.shape
    oH, oW = output_size(H, W, kH, kW, padH, padW, dilationH, dilationW, dH, dW)
    if tuple(grad_output.shape) != (N, patchH, patchW, oH, oW):
        raise RuntimeError(f"grad_output must be {(N, patchH, patchW, oH, oW)} (got {tuple(grad_output.shape)})")
    go = grad_output.to(input1.dtype).contiguous()
    g1 = torch.empty_like(input1)
    g2 = torch.empty_like(input2)
    if g1.numel() == 0:
        return [g1, g2]
    lib = _lib.load()
    code = _dtype(input1)
    ints = (N, Cc, H, W, kH, kW, patchH, patchW, padH, padW, dilationH, dilationW, dilation_patchH, dilation_patchW,
            dH, dW)
    ws_bytes = lib.vh_spatial_corr_backward_workspace_size(code, *ints)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=input1.device) if ws_bytes else None
    with torch.cuda.device(input1.device):
        rc = lib.vh_spatial_corr_backward(code, _lib.ptr(input1), _lib.ptr(input2), _lib.ptr(go), *ints,
                                          _lib.ptr(g1), _lib.ptr(g2), _lib.ptr(ws), ws_bytes, _lib.stream_of(input1))
    _lib.check(rc, "corr_ext.backward")
    return [g1, g2]
