"""FP64 restatement of corr_ext's spatial correlation sampler (oracle; test-only).

Follows csrc/corr_ext/correlation_cuda_kernel.cu:
  forward   :27-81, :217-247   out [N, patchH, patchW, oH, oW];
            oH = (H + 2 padH - ((kH-1) dilH + 1)) / dH + 1;
            out[n,ph,pw,h,w] = sum_{c,i,j} in1[n,c,y1,x1] in2[n,c,y1+sh,x1+sw],
            y1 = h dH - padH + i dilH, x1 = w dW - padW + j dilW,
            sh = ph dpH - (dpH (patchH-1)) / 2 (integer division), a term only
            where both positions are inside the map
  backward  :83-214, :249-286  gathers of the same products; its patch centre
            is ((patchH-1)/2) dpH, which equals the forward's for odd patches
            and for dilation_patch 1, and differs otherwise
Vectorised over batch, channels and output pixels with numpy slicing; one loop
over taps and patch shifts.  Pinned by tests/test_oracle_spatial_corr.py:
a zero-padded conv2d / unfold formulation of the same sum (torch library
routines), autograd of that formulation for the backward, and closed forms.
"""
from __future__ import annotations

import numpy as np


def _cdiv_trunc(a, b):
    """C integer division (truncating toward zero), as the reference writes it."""
    return int(a) // int(b) if a >= 0 else -((-int(a)) // int(b))


def out_size(H, W, kH, kW, padH, padW, dilH, dilW, dH, dW):
    """correlation_cuda_kernel.cu:223-227; a window one stride too large still
    gives one output row (C division truncates toward zero)."""
    oH = _cdiv_trunc(H + 2 * padH - ((kH - 1) * dilH + 1), dH) + 1
    oW = _cdiv_trunc(W + 2 * padW - ((kW - 1) * dilW + 1), dW) + 1
    return oH, oW


def _span(o, stride, pad, tap, dil, shift, size):
    """Output indices o (1-D) whose input position y1 = o*stride - pad + tap*dil and
    y1 + shift both lie in [0, size); returns (mask, y1)."""
    y1 = o * stride - pad + tap * dil
    y2 = y1 + shift
    return (y1 >= 0) & (y1 < size) & (y2 >= 0) & (y2 < size), y1


def forward(in1, in2, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW):
    a = np.asarray(in1, dtype=np.float64)
    b = np.asarray(in2, dtype=np.float64)
    N, C, H, W = a.shape
    oH, oW = out_size(H, W, kH, kW, padH, padW, dilH, dilW, dH, dW)
    out = np.zeros((N, patchH, patchW, max(oH, 0), max(oW, 0)))
    if oH <= 0 or oW <= 0:
        return out
    radH = _cdiv_trunc(dpH * (patchH - 1), 2)
    radW = _cdiv_trunc(dpW * (patchW - 1), 2)
    oh, ow = np.arange(oH), np.arange(oW)
    for ph in range(patchH):
        sh = ph * dpH - radH
        for pw in range(patchW):
            sw = pw * dpW - radW
            for i in range(kH):
                mh, y1 = _span(oh, dH, padH, i, dilH, sh, H)
                if not mh.any():
                    continue
                for j in range(kW):
                    mw, x1 = _span(ow, dW, padW, j, dilW, sw, W)
                    if not mw.any():
                        continue
                    hs, ws = oh[mh], ow[mw]
                    p1 = a[:, :, y1[mh]][:, :, :, x1[mw]]
                    p2 = b[:, :, y1[mh] + sh][:, :, :, x1[mw] + sw]
                    out[:, ph, pw, hs[:, None], ws[None, :]] += (p1 * p2).sum(1)
    return out


def backward(in1, in2, grad_out, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW):
    """(grad_input1, grad_input2) with the reference backward's patch centre."""
    a = np.asarray(in1, dtype=np.float64)
    b = np.asarray(in2, dtype=np.float64)
    g = np.asarray(grad_out, dtype=np.float64)
    N, C, H, W = a.shape
    oH, oW = out_size(H, W, kH, kW, padH, padW, dilH, dilW, dH, dW)
    g1 = np.zeros_like(a)
    g2 = np.zeros_like(b)
    radH = (patchH - 1) // 2 * dpH
    radW = (patchW - 1) // 2 * dpW
    oh, ow = np.arange(oH), np.arange(oW)
    for ph in range(patchH):
        sh = ph * dpH - radH
        for pw in range(patchW):
            sw = pw * dpW - radW
            for i in range(kH):
                mh, y1 = _span(oh, dH, padH, i, dilH, sh, H)
                if not mh.any():
                    continue
                for j in range(kW):
                    mw, x1 = _span(ow, dW, padW, j, dilW, sw, W)
                    if not mw.any():
                        continue
                    # (h, w) -> (y1, x1) is one-to-one for a fixed tap, so += by fancy index is exact
                    ys, xs = y1[mh], x1[mw]
                    go = g[:, ph, pw][:, oh[mh]][:, :, ow[mw]][:, None]       # [N, 1, h, w]
                    Y1, X1 = ys[:, None], xs[None, :]
                    g1[:, :, Y1, X1] += go * b[:, :, Y1 + sh, X1 + sw]
                    g2[:, :, Y1 + sh, X1 + sw] += go * a[:, :, Y1, X1]
    return g1, g2
