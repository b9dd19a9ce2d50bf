"""FP64 restatement of the correlation volume, pyramid and lookups (oracle; test-only).

Follows:
  CorrBlock.corr      vipe/slam/networks/droid_net.py:94-102  (f1/4 . f2/4 over channels)
  pyramid             vipe/slam/networks/droid_net.py:62-69   (avg_pool2d(2,2) over target dims)
  corr_index fwd/bwd  csrc/droid_net_ext/correlation_kernels.cu:22-113
                      (taps (i,j) in [0,2r+1]^2 at floor(x0)-r+i, floor(y0)-r+j, zero
                      padded, splatted into output channel (i*(2r+1)+j): x-major)
  altcorr fwd/bwd     csrc/droid_net_ext/altcorr_kernel.cu:26-138, 242-366
  altcorr indexed     csrc/droid_net_ext/altcorr_kernel.cu:140-240, 394-424
                      (edge m: fmap1[ii[m]] at level 0, fmap2_level[jj[m]], coords / 2^level)
  AltCorrBlock        vipe/slam/networks/droid_net.py:121-196

Pinned by tests/test_oracle_corr.py through: the volume equals an explicit
per-element dot product; the pyramid equals F.avg_pool2d; the lookup equals
F.grid_sample(align_corners=True, padding_mode='zeros') sampled at the window
offsets; the backward equals autograd of the forward (linearity); altcorr
equals a lookup into the materialised volume of the pooled features.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

F64 = torch.float64


def corr_volume(fmap1: torch.Tensor, fmap2: torch.Tensor, dtype=F64) -> torch.Tensor:
    """fmap [B, N, C, H, W] -> [B, N, H, W, H, W] (droid_net.py:94-102).
    ``dtype`` float32 restates the reference's CPU arithmetic (the CPU baseline)."""
    B, N, C, H, W = fmap1.shape
    f1 = fmap1.reshape(B * N, C, H * W).to(dtype) / 4.0
    f2 = fmap2.reshape(B * N, C, H * W).to(dtype) / 4.0
    corr = torch.bmm(f1.transpose(1, 2).contiguous(), f2)
    return corr.view(B, N, H, W, H, W)


def corr_pyramid(corr: torch.Tensor, num_levels: int = 4):
    """[B, N, H1, W1, H2, W2] -> list of [B*N, H1, W1, H2/2^l, W2/2^l] (droid_net.py:62-69)."""
    B, N, H1, W1, H2, W2 = corr.shape
    c = corr.reshape(B * N * H1 * W1, 1, H2, W2)
    out = []
    for i in range(num_levels):
        out.append(c.reshape(B * N, H1, W1, c.shape[-2], c.shape[-1]))
        if i + 1 < num_levels:
            c = F.avg_pool2d(c, 2, stride=2)
    return out


def _window(coord_x, coord_y, r):
    """Integer tap offsets and bilinear fractions shared by every lookup."""
    x0f = torch.floor(coord_x)
    y0f = torch.floor(coord_y)
    dx = coord_x - x0f
    dy = coord_y - y0f
    return x0f.long(), y0f.long(), dx, dy


def corr_index_forward(volume: torch.Tensor, coords: torch.Tensor, r: int, dtype=F64) -> torch.Tensor:
    """volume [B,H1,W1,H2,W2], coords [B,2,H1,W1] -> [B, 2r+1, 2r+1, H1, W1]
    with out[b,i,j] = bilinear(volume plane, x0-r+i, y0-r+j), zero padding
    (correlation_kernels.cu:22-66)."""
    B, H1, W1, H2, W2 = volume.shape
    rd = 2 * r + 1
    v = volume.to(dtype).reshape(B, H1 * W1, H2 * W2)
    xi, yi, dx, dy = _window(coords[:, 0].to(dtype), coords[:, 1].to(dtype), r)
    xi = xi.reshape(B, H1 * W1)
    yi = yi.reshape(B, H1 * W1)
    dx = dx.reshape(B, H1 * W1)
    dy = dy.reshape(B, H1 * W1)
    out = torch.zeros(B, rd, rd, H1 * W1, dtype=dtype)
    for i in range(rd + 1):
        for j in range(rd + 1):
            x1 = xi - r + i
            y1 = yi - r + j
            ok = (x1 >= 0) & (x1 < W2) & (y1 >= 0) & (y1 < H2)
            idx = (y1.clamp(0, H2 - 1) * W2 + x1.clamp(0, W2 - 1))
            s = torch.gather(v, 2, idx[..., None])[..., 0] * ok
            if i > 0 and j > 0:
                out[:, i - 1, j - 1] += s * dx * dy
            if i > 0 and j < rd:
                out[:, i - 1, j] += s * dx * (1 - dy)
            if i < rd and j > 0:
                out[:, i, j - 1] += s * (1 - dx) * dy
            if i < rd and j < rd:
                out[:, i, j] += s * (1 - dx) * (1 - dy)
    return out.view(B, rd, rd, H1, W1)


def corr_index_backward(volume_shape, coords: torch.Tensor, corr_grad: torch.Tensor, r: int) -> torch.Tensor:
    """Adjoint of corr_index_forward (correlation_kernels.cu:68-113)."""
    B, H1, W1, H2, W2 = volume_shape
    rd = 2 * r + 1
    g = corr_grad.to(F64).reshape(B, rd, rd, H1 * W1)
    xi, yi, dx, dy = _window(coords[:, 0].to(F64), coords[:, 1].to(F64), r)
    xi = xi.reshape(B, H1 * W1)
    yi = yi.reshape(B, H1 * W1)
    dx = dx.reshape(B, H1 * W1)
    dy = dy.reshape(B, H1 * W1)
    vg = torch.zeros(B, H1 * W1, H2 * W2, dtype=F64)
    for i in range(rd + 1):
        for j in range(rd + 1):
            x1 = xi - r + i
            y1 = yi - r + j
            ok = (x1 >= 0) & (x1 < W2) & (y1 >= 0) & (y1 < H2)
            acc = torch.zeros(B, H1 * W1, dtype=F64)
            if i > 0 and j > 0:
                acc += g[:, i - 1, j - 1] * dx * dy
            if i > 0 and j < rd:
                acc += g[:, i - 1, j] * dx * (1 - dy)
            if i < rd and j > 0:
                acc += g[:, i, j - 1] * (1 - dx) * dy
            if i < rd and j < rd:
                acc += g[:, i, j] * (1 - dx) * (1 - dy)
            idx = (y1.clamp(0, H2 - 1) * W2 + x1.clamp(0, W2 - 1))
            vg.scatter_add_(2, idx[..., None], (acc * ok)[..., None])
    return vg.view(B, H1, W1, H2, W2)


def corr_block_lookup(pyramid, coords: torch.Tensor, r: int = 3, dtype=F64) -> torch.Tensor:
    """CorrBlock.__call__: coords [B, N, H, W, 2] -> [B, N, L*(2r+1)^2, H, W] (droid_net.py:71-84)."""
    B, N, H, W, _ = coords.shape
    c = coords.permute(0, 1, 4, 2, 3).reshape(B * N, 2, H, W)
    outs = []
    for i, vol in enumerate(pyramid):
        o = corr_index_forward(vol, c / 2**i, r, dtype)
        outs.append(o.view(B, N, -1, H, W))
    return torch.cat(outs, 2)


def _bilinear_dots(f1, f2, xs, ys, r):
    """Shared core of the altcorr kernels.  f1 [P, C] (one feature per output
    pixel), f2 [H2, W2, C], xs/ys [P] sample coordinates -> [P, (2r+1)^2]
    with channel ix*(2r+1)+iy (altcorr_kernel.cu:100-131)."""
    H2, W2, C = f2.shape
    rd = 2 * r + 1
    xi, yi, dx, dy = _window(xs, ys, r)
    out = torch.zeros(f1.shape[0], rd, rd, dtype=F64)  # [P, ix, iy]
    flat = f2.reshape(H2 * W2, C)
    for iy in range(rd + 1):
        for ix in range(rd + 1):
            h2 = yi - r + iy
            w2 = xi - r + ix
            ok = (h2 >= 0) & (h2 < H2) & (w2 >= 0) & (w2 < W2)
            idx = h2.clamp(0, H2 - 1) * W2 + w2.clamp(0, W2 - 1)
            s = (f1 * flat[idx]).sum(-1) * ok
            if iy > 0 and ix > 0:
                out[:, ix - 1, iy - 1] += s * dy * dx
            if iy > 0 and ix < rd:
                out[:, ix, iy - 1] += s * dy * (1 - dx)
            if iy < rd and ix > 0:
                out[:, ix - 1, iy] += s * (1 - dy) * dx
            if iy < rd and ix < rd:
                out[:, ix, iy] += s * (1 - dy) * (1 - dx)
    return out.reshape(f1.shape[0], rd * rd)


def altcorr_forward(fmap1, fmap2, coords, r):
    """fmap1 [B,H1,W1,C], fmap2 [B,H2,W2,C], coords [B,N,H1,W1,2] -> [B,N,(2r+1)^2,H1,W1]
    (altcorr_kernel.cu:26-138, 368-392)."""
    B, H1, W1, C = fmap1.shape
    N = coords.shape[1]
    rd = 2 * r + 1
    out = torch.zeros(B, N, rd * rd, H1, W1, dtype=F64)
    for b in range(B):
        f1 = fmap1[b].to(F64).reshape(H1 * W1, C)
        f2 = fmap2[b].to(F64)
        for n in range(N):
            xs = coords[b, n, ..., 0].to(F64).reshape(-1)
            ys = coords[b, n, ..., 1].to(F64).reshape(-1)
            out[b, n] = _bilinear_dots(f1, f2, xs, ys, r).T.reshape(rd * rd, H1, W1)
    return out


def altcorr_backward(fmap1, fmap2, coords, corr_grad, r):
    """Gradients of altcorr_forward to fmap1 and fmap2; coords_grad is zeros
    (altcorr_kernel.cu:242-366, 426-455).  Computed by autograd through the
    forward restatement (the op is bilinear in (fmap1, fmap2))."""
    f1 = fmap1.to(F64).detach().requires_grad_(True)
    f2 = fmap2.to(F64).detach().requires_grad_(True)
    out = altcorr_forward(f1, f2, coords, r)
    (out * corr_grad.to(F64)).sum().backward()
    return f1.grad, f2.grad, torch.zeros(coords.shape, dtype=F64)


def altcorr_pyramid(fmaps: torch.Tensor, num_levels: int = 4):
    """AltCorrBlock.__init__: fmaps [B,N,C,H,W] -> list of [B,N,H/2^l,W/2^l,C] of fmaps/4
    avg-pooled (droid_net.py:130-142)."""
    B, N, C, H, W = fmaps.shape
    f = fmaps.reshape(B * N, C, H, W).to(F64) / 4.0
    out = []
    for i in range(num_levels):
        out.append(f.permute(0, 2, 3, 1).reshape(B, N, f.shape[-2], f.shape[-1], C))
        if i + 1 < num_levels:
            f = F.avg_pool2d(f, 2, stride=2)
    return out


def altcorr_index_forward(fmap1, pyramid, coords, ii, jj, r):
    """fmap1 [B,N,H,W,C], pyramid list of [B,N,H_l,W_l,C], coords [B,M,H,W,2],
    ii/jj [M] -> [B, M, L*(2r+1)^2, H, W] (altcorr_kernel.cu:140-240, 394-424)."""
    B, M, H, W, _ = coords.shape
    C = fmap1.shape[-1]
    rd = 2 * r + 1
    L = len(pyramid)
    out = torch.zeros(B, M, L * rd * rd, H, W, dtype=F64)
    for b in range(B):
        for m in range(M):
            f1 = fmap1[b, int(ii[m])].to(F64).reshape(H * W, C)
            for lvl, f2l in enumerate(pyramid):
                xs = coords[b, m, ..., 0].to(F64).reshape(-1) / 2**lvl
                ys = coords[b, m, ..., 1].to(F64).reshape(-1) / 2**lvl
                o = _bilinear_dots(f1, f2l[b, int(jj[m])].to(F64), xs, ys, r)
                out[b, m, lvl * rd * rd:(lvl + 1) * rd * rd] = o.T.reshape(rd * rd, H, W)
    return out
