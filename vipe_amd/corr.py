"""Correlation volume classes of the DROID update loop on gfx950.

Mirrors vipe/slam/networks/droid_net.py:32-196 (same class names, constructor
arguments, call signatures and output layouts):

* ``CorrSampler`` / ``CorrLayer``: autograd Functions over the
  ``droid_net_ext`` lookup ops (droid_net.py:32-45, 105-118).
* ``CorrBlock``: all-pairs volume + 4-level pyramid, looked up at radius r
  (droid_net.py:48-102).  The volume and its pooled levels are built by one
  MFMA kernel plus pooling kernels (``droid_net_ext.corr_build``) instead of
  ``torch.matmul`` + ``F.avg_pool2d``; the values are the same,
  ``(f1/4).(f2/4)`` summed over channels.
* ``AltCorrBlock``: the memory-light on-the-fly variant (droid_net.py:121-196).
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

from .ext import droid_net_ext


class CorrSampler(torch.autograd.Function):
    @staticmethod
    def forward(ctx, volume, coords, radius):
        ctx.save_for_backward(volume, coords)
        ctx.radius = radius
        (corr,) = droid_net_ext.corr_index_forward(volume, coords, radius)
        return corr

    @staticmethod
    def backward(ctx, grad_output):
        volume, coords = ctx.saved_tensors
        grad_output = grad_output.contiguous()
        (grad_volume,) = droid_net_ext.corr_index_backward(volume, coords, grad_output, ctx.radius)
        return grad_volume, None, None


class CorrPyramidSampler(torch.autograd.Function):
    """All levels of one CorrBlock lookup in one kernel; backward scatters each
    level's output gradient into its volume (corr_index_backward, re-tiled for
    tiled pyramids)."""

    @staticmethod
    def forward(ctx, coords, radius, layout, level_dims, *levels):
        ctx.radius, ctx.layout, ctx.level_dims = radius, layout, level_dims
        ctx.save_for_backward(coords, *levels)
        return droid_net_ext.corr_lookup_pyramid(list(levels), coords, radius, level_dims, layout)

    @staticmethod
    def backward(ctx, grad_output):
        coords, *levels = ctx.saved_tensors
        rd = 2 * ctx.radius + 1
        B, H, W, _ = coords.shape
        cxy = coords.permute(0, 3, 1, 2).contiguous()
        grads = []
        for i, vol in enumerate(levels):
            if not ctx.needs_input_grad[4 + i]:
                grads.append(None)
                continue
            g = grad_output[:, i * rd * rd:(i + 1) * rd * rd].reshape(B, rd, rd, H, W).contiguous()
            planar = vol if ctx.layout != "tiled" else droid_net_ext.untile(vol, *ctx.level_dims[i])
            (gv,) = droid_net_ext.corr_index_backward(planar, (cxy / 2**i).contiguous(), g, ctx.radius)
            grads.append(gv if ctx.layout != "tiled" else droid_net_ext.tile(gv))
        return (None, None, None, None, *grads)


class CorrLookupFeatures(torch.autograd.Function):
    """Fused pyramid lookup whose backward returns the gradients of both
    feature maps (CorrBlock built from ``fmap1``, ``fmap2`` [E, C, H, W]).

    By linearity of the pooling, level l of the lookup equals the on-the-fly
    correlation of fmap1/4 with avgpool^l(fmap2)/4 sampled at coords/2^l
    (droid_net.py:56-102, altcorr_kernel.cu:242-366), so each level's feature
    gradients come from the altcorr backward kernel; the level-l gradient of
    fmap2 is spread back over its 2^l x 2^l block with weight 4^-l (the
    adjoint of the repeated 2x2 mean; floor-dropped rows and columns get 0).
    """

    @staticmethod
    def forward(ctx, coords, radius, layout, level_dims, fmap1, fmap2, *levels):
        ctx.radius = radius
        ctx.num_levels = len(levels)
        ctx.save_for_backward(coords, fmap1, fmap2)
        return droid_net_ext.corr_lookup_pyramid(list(levels), coords, radius, level_dims, layout)

    @staticmethod
    def backward(ctx, grad_output):
        coords, f1, f2 = ctx.saved_tensors
        E, Cc, H, W = f1.shape
        r, L = ctx.radius, ctx.num_levels
        rd2 = (2 * r + 1) ** 2
        g = grad_output.float()
        F1 = droid_net_ext.feat_to_nhwc(f1, 0.25)
        F2 = droid_net_ext.feat_to_nhwc(f2, 0.25)
        g1s, g2s = [], []
        for lvl in range(L):
            if lvl > 0:
                F2 = droid_net_ext.pool2x2_nhwc(F2)
            gl = g[:, lvl * rd2:(lvl + 1) * rd2].reshape(E, 1, rd2, H, W).contiguous()
            cl = (coords / 2 ** lvl).reshape(E, 1, H, W, 2).contiguous()
            g1, g2, _ = droid_net_ext.altcorr_backward(F1, F2, cl, gl, r)
            g1s.append(g1)
            g2s.append(g2)
        dt = f1.dtype if f1.dtype in (torch.float32, torch.float16, torch.bfloat16) else torch.float32
        df1, df2 = droid_net_ext.corr_feature_grad(g1s, g2s, 0.25, dt)
        return (None, None, None, None, df1.to(f1.dtype) if ctx.needs_input_grad[4] else None,
                df2.to(f2.dtype) if ctx.needs_input_grad[5] else None, *([None] * L))


class CorrBlock:
    """All-pairs correlation pyramid of two feature maps ``[B, N, C, H, W]``.

    ``volume_dtype`` selects the storage of the pyramid (float32, or float16
    as the reference produces under autocast); by default it follows the
    feature dtype.  ``contraction_dtype`` selects the MFMA operand type of the
    volume build (default: the feature dtype; float16, bfloat16 or
    float8_e4m3fn for the mixed-precision forms, all accumulating in FP32).
    ``layout`` "tiled" (default) keeps each level's planes as 128-byte tiles
    shared by four neighbouring source pixels, which the build writes and the
    lookup reads in whole cache lines (more than 4 levels, or H*W not a multiple of 4, use
    "planar");
    "planar" keeps the reference's [B*N, H, W, h, w] levels.  Callers only build,
    look up, ``cat`` and index along the edge dimension, which both support.
    """

    def __init__(self, fmap1, fmap2, num_levels=4, radius=3, volume_dtype=None, contraction_dtype=None,
                 layout="tiled"):
        self.num_levels = num_levels
        self.radius = radius
        B, N, C, H, W = fmap1.shape
        f1 = fmap1.reshape(B * N, C, H, W).contiguous()
        f2 = fmap2.reshape(B * N, C, H, W).contiguous()
        if layout == "tiled" and (num_levels > 4 or (H * W) % droid_net_ext.TILE_PLANES):
            layout = "planar"
        levels = droid_net_ext.corr_build(f1.detach(), f2.detach(), num_levels, out_dtype=volume_dtype,
                                          contraction_dtype=contraction_dtype, layout=layout)
        self.layout = layout
        self.level_dims = [(H >> i, W >> i) for i in range(num_levels)]
        self.corr_pyramid = levels  # level i: [B*N, H, W, H/2^i, W/2^i] (planar) or [B*N, H, W, Hp*Wp]
        # kept for gradients to the features (the pyramid itself is built without autograd)
        self._fmaps = (f1, f2) if (f1.requires_grad or f2.requires_grad) else None

    def __call__(self, coords):
        """coords [B, N, H, W, 2] -> [B, N, L*(2r+1)^2, H, W] (droid_net.py:71-84), all
        levels in one fused kernel reading the coordinates in place."""
        batch, num, ht, wd, _ = coords.shape
        c = coords.reshape(batch * num, ht, wd, 2).contiguous().float()
        if self._fmaps is not None and torch.is_grad_enabled():
            out = CorrLookupFeatures.apply(c, self.radius, self.layout, self.level_dims, self._fmaps[0],
                                           self._fmaps[1], *self.corr_pyramid)
        else:
            out = CorrPyramidSampler.apply(c, self.radius, self.layout, self.level_dims, *self.corr_pyramid)
        return out.view(batch, num, -1, ht, wd)

    def lookup_per_level(self, coords):
        """The reference's per-level form: one corr_index_forward per level + concat."""
        out_pyramid = []
        batch, num, ht, wd, _ = coords.shape
        coords = coords.permute(0, 1, 4, 2, 3)
        coords = coords.contiguous().view(batch * num, 2, ht, wd)
        for i in range(self.num_levels):
            vol = self.corr_pyramid[i]
            if self.layout == "tiled":
                vol = droid_net_ext.untile(vol, *self.level_dims[i])
            corr = CorrSampler.apply(vol, coords / 2**i, self.radius)
            out_pyramid.append(corr.view(batch, num, -1, ht, wd))
        return torch.cat(out_pyramid, dim=2)

    def cat(self, other):
        self._fmaps = None
        for i in range(self.num_levels):
            self.corr_pyramid[i] = torch.cat([self.corr_pyramid[i], other.corr_pyramid[i]], 0)
        return self

    def __getitem__(self, index):
        self._fmaps = None
        for i in range(self.num_levels):
            self.corr_pyramid[i] = self.corr_pyramid[i][index]
        return self

    @staticmethod
    def corr(fmap1, fmap2):
        """all-pairs correlation [B, N, H, W, H, W] (level 0 only)."""
        batch, num, dim, ht, wd = fmap1.shape
        (lvl0,) = droid_net_ext.corr_build(fmap1.reshape(batch * num, dim, ht, wd).contiguous(),
                                           fmap2.reshape(batch * num, dim, ht, wd).contiguous(), 1, layout="planar")
        return lvl0.view(batch, num, ht, wd, ht, wd)


class CorrLayer(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fmap1, fmap2, coords, r):
        ctx.r = r
        ctx.save_for_backward(fmap1, fmap2, coords)
        (corr,) = droid_net_ext.altcorr_forward(fmap1, fmap2, coords, ctx.r)
        return corr

    @staticmethod
    def backward(ctx, grad_corr):
        fmap1, fmap2, coords = ctx.saved_tensors
        grad_corr = grad_corr.contiguous()
        fmap1_grad, fmap2_grad, coords_grad = droid_net_ext.altcorr_backward(fmap1, fmap2, coords, grad_corr, ctx.r)
        return fmap1_grad, fmap2_grad, coords_grad, None


class AltCorrBlock:
    """On-the-fly correlation over an NHWC feature pyramid (fmaps/4, avg-pooled)."""

    def __init__(self, fmaps, num_levels=4, radius=3):
        self.num_levels = num_levels
        self.radius = radius
        B, N, C, H, W = fmaps.shape
        # fmaps / 4 channels-last, then 2x2 means per level (droid_net.py:146-160)
        f = fmaps.reshape(B * N, C, H, W).contiguous()
        lvl = droid_net_ext.feat_to_nhwc(f, 0.25, out_dtype=f.dtype)
        self.pyramid = []
        for i in range(self.num_levels):
            if i > 0:
                lvl = droid_net_ext.pool2x2_nhwc(lvl)
            self.pyramid.append(lvl.view(B, N, H // 2**i, W // 2**i, C))

    def corr_fn_reference(self, coords, ii, jj):
        B, N, H, W, S, _ = coords.shape
        coords = coords.permute(0, 1, 4, 2, 3, 5)
        corr_list = []
        for i in range(self.num_levels):
            fmap1_i = self.pyramid[0][:, ii]
            fmap2_i = self.pyramid[i][:, jj]
            coords_i = (coords / 2**i).reshape(B * N, S, H, W, 2).contiguous()
            fmap1_i = fmap1_i.reshape((B * N,) + fmap1_i.shape[2:])
            fmap2_i = fmap2_i.reshape((B * N,) + fmap2_i.shape[2:])
            corr = CorrLayer.apply(fmap1_i.float().contiguous(), fmap2_i.float().contiguous(), coords_i, self.radius)
            corr = corr.view(B, N, S, -1, H, W).permute(0, 1, 3, 4, 5, 2)
            corr_list.append(corr)
        return torch.cat(corr_list, dim=2)

    def corr_fn_indexed(self, coords, ii, jj):
        if coords.shape[-2] != 1:
            return self.corr_fn_reference(coords, ii, jj)
        coords = coords.squeeze(dim=-2).contiguous()
        (corr,) = droid_net_ext.altcorr_index_forward(self.pyramid[0], self.pyramid, coords, ii.contiguous(),
                                                      jj.contiguous(), self.radius)
        return corr.unsqueeze(dim=-1)

    def corr_fn(self, coords, ii, jj):
        return self.corr_fn_indexed(coords, ii, jj)

    def __call__(self, coords, ii, jj):
        squeeze_output = False
        if len(coords.shape) == 5:
            coords = coords.unsqueeze(dim=-2)
            squeeze_output = True
        corr = self.corr_fn(coords, ii, jj)
        if squeeze_output:
            corr = corr.squeeze(dim=-1)
        return corr.contiguous()
