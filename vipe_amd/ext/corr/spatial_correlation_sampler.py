"""Spatial correlation sampling with the interface of
vipe/ext/corr/spatial_correlation_sampler.py:13-128: every size argument is an
int or an (H, W) pair; the output is [N, patchH, patchW, oH, oW]."""
from __future__ import annotations

from torch import nn
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from .. import corr_ext

_FIELDS = ("kernel_size", "patch_size", "padding", "dilation", "dilation_patch", "stride")


def _pair(v):
    if isinstance(v, (tuple, list)):
        if len(v) != 2:
            raise ValueError(f"expected an int or a pair, got {v!r}")
        return int(v[0]), int(v[1])
    return int(v), int(v)


def _flat(params):
    """(kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW) in the
    argument order of corr_ext.forward / backward."""
    return tuple(x for name in _FIELDS for x in params[name])


class SpatialCorrelationSamplerFunction(Function):
    @staticmethod
    def forward(ctx, input1, input2, kernel_size=1, patch_size=1, stride=1, padding=0, dilation=1,
                dilation_patch=1):
        params = {"kernel_size": _pair(kernel_size), "patch_size": _pair(patch_size), "padding": _pair(padding),
                  "dilation": _pair(dilation), "dilation_patch": _pair(dilation_patch), "stride": _pair(stride)}
        ctx.save_for_backward(input1, input2)
        ctx.params = params
        return corr_ext.forward(input1, input2, *_flat(params))

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output):
        input1, input2 = ctx.saved_tensors
        g1, g2 = corr_ext.backward(input1, input2, grad_output, *_flat(ctx.params))
        return g1, g2, None, None, None, None, None, None


def spatial_correlation_sample(input1, input2, kernel_size=1, patch_size=1, stride=1, padding=0, dilation=1,
                               dilation_patch=1):
    """Correlate input1 with input2 over a patch of shifts (FlowNet-style local
    correlation).  kernel_size: window summed per output; patch_size: grid of
    shifts; stride / padding / dilation: of the window; dilation_patch: step
    between shifts."""
    return SpatialCorrelationSamplerFunction.apply(input1, input2, kernel_size, patch_size, stride, padding,
                                                   dilation, dilation_patch)


class SpatialCorrelationSampler(nn.Module):
    def __init__(self, kernel_size=1, patch_size=1, stride=1, padding=0, dilation=1, dilation_patch=1):
        super().__init__()
        self.kernel_size = kernel_size
        self.patch_size = patch_size
        self.stride = stride
        self.padding = padding
        self.dilation = dilation
        self.dilation_patch = dilation_patch

    def forward(self, input1, input2):
        return spatial_correlation_sample(input1, input2, self.kernel_size, self.patch_size, self.stride,
                                          self.padding, self.dilation, self.dilation_patch)
