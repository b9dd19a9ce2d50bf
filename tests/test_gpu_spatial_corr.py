"""gfx950 spatial correlation sampler (corr_ext) vs the FP64 oracle.

FP32 / FP16 / BF16 inputs accumulate in FP32; the tolerance is relative to the
largest |sum of |products|| so that cancellation in a sum does not tighten it:
FP32 2e-6 of that scale, FP16 and BF16 outputs one rounding of the output type
(2^-11, 2^-8) plus the FP32 sum.  FP64 inputs accumulate in FP64 (1e-12)."""
import numpy as np
import pytest
import torch

from oracle import spatial_corr as OS

pytestmark = pytest.mark.gpu

CASES = [
    # kH kW pH pW padH padW dilH dilW dpH dpW dH dW
    (1, 1, 15, 15, 0, 0, 1, 1, 1, 1, 1, 1),   # local attention: kernel 1, 15 x 15 window
    (1, 1, 9, 9, 0, 0, 1, 1, 2, 2, 1, 1),
    (3, 3, 5, 3, 1, 1, 1, 1, 1, 1, 1, 1),
    (3, 2, 3, 5, 2, 0, 2, 1, 2, 1, 2, 1),
    (2, 3, 4, 2, 1, 2, 1, 2, 3, 2, 1, 3),    # even patches, dilation_patch > 1
    (5, 1, 1, 3, 0, 3, 1, 1, 1, 1, 3, 2),
    (1, 1, 5, 7, 2, 3, 1, 1, 1, 3, 1, 1),    # kernel 1 with padding, a 3-tile band
    (1, 1, 3, 13, 0, 0, 1, 1, 2, 4, 1, 1),   # kernel 1, a 4-tile band ((13-1)*4 = 48)
    (1, 1, 2, 4, 1, 0, 2, 1, 3, 2, 1, 1),    # kernel 1, even patches
]
TOL = {torch.float32: 2e-6, torch.float16: 2 ** -11 + 2e-6, torch.bfloat16: 2 ** -8 + 2e-6}


def _inputs(shape, dtype, seed):
    g = torch.Generator().manual_seed(seed)
    a = torch.randn(shape, generator=g).to(dtype)
    b = torch.randn(shape, generator=g).to(dtype)
    return a, b


def _scale(a, b, case):
    """Largest sum of |products| of any output (the oracle on |a|, |b|)."""
    return float(np.abs(OS.forward(np.abs(a), np.abs(b), *case)).max())


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.bfloat16, torch.float64])
def test_forward_matches_oracle(case, dtype):
    from vipe_amd.ext import corr_ext

    a, b = _inputs((2, 19, 13, 70), dtype, seed=sum(case))
    out = corr_ext.forward(a.cuda(), b.cuda(), *case)
    an, bn = a.double().numpy(), b.double().numpy()
    ref = OS.forward(an, bn, *case)
    assert out.dtype == dtype and tuple(out.shape) == ref.shape
    err = np.abs(out.cpu().double().numpy() - ref).max()
    if dtype == torch.float64:
        assert err < 1e-12 * max(1.0, _scale(an, bn, case))
    else:
        assert err <= TOL[dtype] * _scale(an, bn, case), err


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16])
def test_backward_matches_oracle(case, dtype):
    from vipe_amd.ext import corr_ext

    a, b = _inputs((2, 7, 12, 67), dtype, seed=11 + sum(case))
    oH, oW = OS.out_size(12, 67, case[0], case[1], case[4], case[5], case[6], case[7], case[10], case[11])
    go = torch.randn((2, case[2], case[3], oH, oW), generator=torch.Generator().manual_seed(5)).to(dtype)
    g1, g2 = corr_ext.backward(a.cuda(), b.cuda(), go.cuda(), *case)
    an, bn, gn = a.double().numpy(), b.double().numpy(), go.double().numpy()
    r1, r2 = OS.backward(an, bn, gn, *case)
    # scale: the same gathers over |values|
    s1, s2 = OS.backward(np.abs(an), np.abs(bn), np.abs(gn), *case)
    tol = 1e-12 if dtype == torch.float64 else TOL[dtype]
    assert np.abs(g1.cpu().double().numpy() - r1).max() <= tol * max(1.0, s1.max())
    assert np.abs(g2.cpu().double().numpy() - r2).max() <= tol * max(1.0, s2.max())


def test_module_autograd_and_gradcheck():
    from vipe_amd.ext.corr import SpatialCorrelationSampler, spatial_correlation_sample

    g = torch.Generator().manual_seed(2)
    q = torch.randn(2, 6, 9, 11, generator=g, dtype=torch.float64).cuda().requires_grad_()
    k = torch.randn(2, 6, 9, 11, generator=g, dtype=torch.float64).cuda().requires_grad_()
    assert torch.autograd.gradcheck(
        lambda x, y: spatial_correlation_sample(x, y, kernel_size=3, patch_size=5, stride=(1, 2), padding=1,
                                                dilation=1, dilation_patch=2), (q, k))
    m = SpatialCorrelationSampler(kernel_size=1, patch_size=7, stride=1, padding=0, dilation=1, dilation_patch=1)
    out = m(q.float(), k.float())
    assert out.shape == (2, 7, 7, 9, 11)
    out.sum().backward()
    assert q.grad is not None and k.grad is not None


def test_local_attention_shape():
    """Shape of the local-attention caller: [n*heads, d, h, w] with a 15 x 15
    window, viewed as [n, heads, 225, h*w]."""
    from vipe_amd.ext.corr import SpatialCorrelationSampler

    n, heads, d, h, w = 2, 4, 32, 30, 54
    g = torch.Generator().manual_seed(9)
    q = torch.randn(n * heads, d, h, w, generator=g).cuda()
    k = torch.randn(n * heads, d, h, w, generator=g).cuda()
    qk = SpatialCorrelationSampler(kernel_size=1, patch_size=15, stride=1, padding=0, dilation=1,
                                   dilation_patch=1)(q, k).view(n, heads, 225, h * w)
    ref = OS.forward(q.cpu().double().numpy(), k.cpu().double().numpy(), 1, 1, 15, 15, 0, 0, 1, 1, 1, 1, 1, 1)
    scale = float(np.abs(OS.forward(np.abs(q.cpu().double().numpy()), np.abs(k.cpu().double().numpy()),
                                     1, 1, 15, 15, 0, 0, 1, 1, 1, 1, 1, 1)).max())
    err = np.abs(qk.cpu().double().numpy().reshape(ref.shape) - ref).max()
    assert err <= 2e-6 * scale


def test_argument_errors_on_device():
    from vipe_amd.ext import corr_ext

    a = torch.randn(1, 3, 8, 8, device="cuda")
    with pytest.raises(RuntimeError, match="contiguous"):
        corr_ext.forward(a.transpose(2, 3), a, 1, 1, 3, 3, 0, 0, 1, 1, 1, 1, 1, 1)
    with pytest.raises(RuntimeError, match="one shape"):
        corr_ext.forward(a, a[:, :2].contiguous(), 1, 1, 3, 3, 0, 0, 1, 1, 1, 1, 1, 1)
    with pytest.raises(RuntimeError, match="empty output"):
        corr_ext.forward(a, a, 9, 9, 3, 3, 0, 0, 1, 1, 1, 1, 1, 1)
    with pytest.raises(RuntimeError, match="grad_output"):
        corr_ext.backward(a, a, torch.zeros(1, 3, 3, 7, 7, device="cuda"), 1, 1, 3, 3, 0, 0, 1, 1, 1, 1, 1, 1)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16])
def test_matrix_core_path_wide_channels_odd_width(dtype):
    """C = 300 (several channel steps and a padded tail), odd width, padding:
    forward and backward of the kernel-1 path against the oracle."""
    from vipe_amd.ext import corr_ext

    case = (1, 1, 7, 9, 1, 2, 1, 1, 1, 2, 1, 1)
    a, b = _inputs((3, 300, 11, 37), dtype, seed=77)
    out = corr_ext.forward(a.cuda(), b.cuda(), *case)
    an, bn = a.double().numpy(), b.double().numpy()
    ref = OS.forward(an, bn, *case)
    assert np.abs(out.cpu().double().numpy() - ref).max() <= TOL[dtype] * _scale(an, bn, case)
    go = torch.randn(out.shape, generator=torch.Generator().manual_seed(3)).to(dtype)
    g1, g2 = corr_ext.backward(a.cuda(), b.cuda(), go.cuda(), *case)
    r1, r2 = OS.backward(an, bn, go.double().numpy(), *case)
    s1, s2 = OS.backward(np.abs(an), np.abs(bn), np.abs(go.double().numpy()), *case)
    assert np.abs(g1.cpu().double().numpy() - r1).max() <= TOL[dtype] * max(1.0, s1.max())
    assert np.abs(g2.cpu().double().numpy() - r2).max() <= TOL[dtype] * max(1.0, s2.max())
