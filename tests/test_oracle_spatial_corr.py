"""Pins oracle/spatial_corr.py (host only): the correlation sum written as a
zero-padded conv2d of shifted products (torch library routines), the unfold
form that the local-attention callers use for kernel_size 1, autograd of the
conv2d form for the backward, and closed forms."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import spatial_corr as OS


def _shift(x, sh, sw):
    """y[..., r, c] = x[..., r + sh, c + sw], zero outside the map."""
    H, W = x.shape[-2:]
    y = torch.zeros_like(x)
    r0, r1 = max(0, -sh), min(H, H - sh)
    c0, c1 = max(0, -sw), min(W, W - sw)
    if r0 < r1 and c0 < c1:
        y[..., r0:r1, c0:c1] = x[..., r0 + sh:r1 + sh, c0 + sw:c1 + sw]
    return y


def _conv_form(in1, in2, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW, radH=None, radW=None):
    """Library form: for each shift, conv2d of in1 * shift(in2) with a ones kernel
    (zero padding) sums exactly the in-bounds products of the definition."""
    N, C, H, W = in1.shape
    radH = (dpH * (patchH - 1)) // 2 if radH is None else radH
    radW = (dpW * (patchW - 1)) // 2 if radW is None else radW
    ones = torch.ones(1, C, kH, kW, dtype=in1.dtype)
    rows = []
    for ph in range(patchH):
        cols = []
        for pw in range(patchW):
            prod = in1 * _shift(in2, ph * dpH - radH, pw * dpW - radW)
            cols.append(F.conv2d(prod, ones, stride=(dH, dW), padding=(padH, padW), dilation=(dilH, dilW))[:, 0])
        rows.append(torch.stack(cols, 1))
    return torch.stack(rows, 1)


CASES = [
    # kH kW pH pW padH padW dilH dilW dpH dpW dH dW
    (1, 1, 7, 7, 0, 0, 1, 1, 1, 1, 1, 1),
    (3, 3, 5, 3, 1, 1, 1, 1, 1, 1, 1, 1),
    (3, 2, 3, 5, 2, 0, 2, 1, 2, 1, 2, 1),
    (2, 3, 4, 2, 1, 2, 1, 2, 3, 2, 1, 3),   # even patches with dilation_patch > 1
    (1, 1, 9, 9, 0, 0, 1, 1, 2, 2, 1, 1),
    (5, 1, 1, 3, 0, 3, 1, 1, 1, 1, 3, 2),
]


@pytest.mark.parametrize("case", CASES)
def test_forward_equals_conv2d_form(case):
    g = torch.Generator().manual_seed(sum(case))
    in1 = torch.randn(2, 5, 11, 13, generator=g, dtype=torch.float64)
    in2 = torch.randn(2, 5, 11, 13, generator=g, dtype=torch.float64)
    ref = _conv_form(in1, in2, *case)
    out = OS.forward(in1.numpy(), in2.numpy(), *case)
    assert out.shape == tuple(ref.shape)
    np.testing.assert_allclose(out, ref.numpy(), rtol=1e-12, atol=1e-12)


def test_forward_kernel1_equals_unfold_form():
    """The local-attention callers' fallback: pad k by the patch radius, unfold the
    window and contract channels with q."""
    g = torch.Generator().manual_seed(3)
    q = torch.randn(3, 8, 9, 10, generator=g, dtype=torch.float64)
    k = torch.randn(3, 8, 9, 10, generator=g, dtype=torch.float64)
    P, d = 7, 2
    rad = d * (P - 1) // 2
    unf = F.unfold(F.pad(k, (rad, rad, rad, rad)), P, dilation=d).view(3, 8, P * P, 9, 10)
    ref = (q[:, :, None] * unf).sum(1).view(3, P, P, 9, 10)
    out = OS.forward(q.numpy(), k.numpy(), 1, 1, P, P, 0, 0, 1, 1, d, d, 1, 1)
    np.testing.assert_allclose(out, ref.numpy(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("case", CASES)
def test_backward_equals_autograd(case):
    kH, kW, pH, pW, padH, padW, dilH, dilW, dpH, dpW, dH, dW = case
    g = torch.Generator().manual_seed(7 + sum(case))
    in1 = torch.randn(2, 4, 10, 12, generator=g, dtype=torch.float64, requires_grad=True)
    in2 = torch.randn(2, 4, 10, 12, generator=g, dtype=torch.float64, requires_grad=True)
    # the reference backward centres the patch at ((patch-1)/2) * dilation_patch
    radH, radW = (pH - 1) // 2 * dpH, (pW - 1) // 2 * dpW
    out = _conv_form(in1, in2, *case, radH=radH, radW=radW)
    go = torch.randn(out.shape, generator=g, dtype=torch.float64)
    r1, r2 = torch.autograd.grad(out, (in1, in2), go)
    g1, g2 = OS.backward(in1.detach().numpy(), in2.detach().numpy(), go.numpy(), *case)
    np.testing.assert_allclose(g1, r1.numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g2, r2.numpy(), rtol=1e-12, atol=1e-12)
    if (dpH * (pH - 1)) // 2 == radH and (dpW * (pW - 1)) // 2 == radW:
        # same centre as the forward: the backward is the forward's adjoint
        y = OS.forward(in1.detach().numpy(), in2.detach().numpy(), *case)
        lhs = float((y * go.numpy()).sum())
        assert abs(lhs - float((g1 * in1.detach().numpy()).sum())) < 1e-9 * max(1.0, abs(lhs))
        assert abs(lhs - float((g2 * in2.detach().numpy()).sum())) < 1e-9 * max(1.0, abs(lhs))


def test_closed_forms():
    # patch 1, kernel 1: the channel dot product at every pixel
    a = np.arange(2 * 3 * 4 * 5, dtype=np.float64).reshape(2, 3, 4, 5) / 10
    b = np.cos(a)
    out = OS.forward(a, b, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 1)
    np.testing.assert_allclose(out[:, 0, 0], (a * b).sum(1), rtol=1e-14)
    # constant maps: each output counts its in-bounds (tap, shift) pairs
    ones = np.ones((1, 2, 6, 6))
    out = OS.forward(ones, ones, 1, 1, 3, 3, 0, 0, 1, 1, 1, 1, 1, 1)
    assert out[0, 1, 1].min() == 2.0                   # zero shift: every pixel, C = 2
    assert out[0, 0, 0, 0, 0] == 0.0 and out[0, 2, 2, 5, 5] == 0.0   # shifted off the map
    assert out[0, 0, 0, 1, 1] == 2.0
    # output size
    assert OS.out_size(11, 13, 3, 2, 2, 0, 2, 1, 2, 1) == (6, 12)
    assert OS.forward(ones, ones, 7, 7, 1, 1, 0, 0, 1, 1, 1, 1, 1, 1).shape == (1, 1, 1, 0, 0)
