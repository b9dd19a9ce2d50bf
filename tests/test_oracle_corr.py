"""Pins oracle/corr.py against torch reference operators it does not use:
explicit dot products, F.avg_pool2d, F.grid_sample and autograd."""
import math

import torch
import torch.nn.functional as F

from oracle import corr as OC

F64 = torch.float64


def test_volume_is_scaled_dot_product():
    g = torch.Generator().manual_seed(0)
    f1 = torch.randn(1, 2, 16, 5, 6, generator=g, dtype=F64)
    f2 = torch.randn(1, 2, 16, 5, 6, generator=g, dtype=F64)
    v = OC.corr_volume(f1, f2)
    for (n, y1, x1, y2, x2) in [(0, 0, 0, 0, 0), (1, 4, 5, 2, 3), (0, 2, 1, 4, 0)]:
        ref = sum(f1[0, n, c, y1, x1] * f2[0, n, c, y2, x2] for c in range(16)) / 16.0
        assert abs(v[0, n, y1, x1, y2, x2].item() - ref.item()) < 1e-12


def test_pyramid_is_mean_of_2x2_blocks():
    g = torch.Generator().manual_seed(1)
    v = torch.randn(1, 1, 2, 2, 9, 11, generator=g, dtype=F64)
    pyr = OC.corr_pyramid(v, 3)
    assert pyr[1].shape[-2:] == (4, 5) and pyr[2].shape[-2:] == (2, 2)
    manual = 0.25 * (v[..., 0:8:2, 0:10:2] + v[..., 1:8:2, 0:10:2] + v[..., 0:8:2, 1:10:2] + v[..., 1:8:2, 1:10:2])
    assert (pyr[1] - manual.reshape(pyr[1].shape)).abs().max() < 1e-12


def _grid_sample_lookup(volume, coords, r):
    """Lookup via F.grid_sample(align_corners=True, zeros) at the (2r+1)^2 window."""
    B, H1, W1, H2, W2 = volume.shape
    rd = 2 * r + 1
    planes = volume.reshape(B * H1 * W1, 1, H2, W2)
    x0 = coords[:, 0].reshape(-1)
    y0 = coords[:, 1].reshape(-1)
    out = torch.zeros(B * H1 * W1, rd, rd, dtype=F64)
    for i in range(rd):
        for j in range(rd):
            xs = x0 - r + i
            ys = y0 - r + j
            gx = 2 * xs / (W2 - 1) - 1
            gy = 2 * ys / (H2 - 1) - 1
            grid = torch.stack([gx, gy], -1).view(-1, 1, 1, 2)
            out[:, i, j] = F.grid_sample(planes, grid, mode="bilinear", padding_mode="zeros",
                                         align_corners=True).view(-1)
    return out.view(B, H1, W1, rd, rd).permute(0, 3, 4, 1, 2)


def test_corr_index_equals_grid_sample_and_backward_is_adjoint():
    g = torch.Generator().manual_seed(2)
    B, H1, W1, H2, W2, r = 2, 4, 5, 7, 9, 2
    vol = torch.randn(B, H1, W1, H2, W2, generator=g, dtype=F64)
    coords = torch.stack([torch.rand(B, H1, W1, generator=g, dtype=F64) * (W2 + 4) - 2,
                          torch.rand(B, H1, W1, generator=g, dtype=F64) * (H2 + 4) - 2], 1)
    ours = OC.corr_index_forward(vol, coords, r)
    ref = _grid_sample_lookup(vol, coords, r)
    assert (ours - ref).abs().max() < 1e-12
    # backward = autograd of the grid_sample formulation
    gout = torch.randn(ours.shape, generator=g, dtype=F64)
    v = vol.clone().requires_grad_(True)
    (_grid_sample_lookup(v, coords, r) * gout).sum().backward()
    assert (OC.corr_index_backward(vol.shape, coords, gout, r) - v.grad).abs().max() < 1e-12


def test_altcorr_equals_lookup_of_materialised_volume():
    g = torch.Generator().manual_seed(3)
    B, N, H, W, C, r = 1, 2, 5, 6, 8, 3
    f1 = torch.randn(B, H, W, C, generator=g, dtype=F64)
    f2 = torch.randn(B, H, W, C, generator=g, dtype=F64)
    coords = torch.rand(B, N, H, W, 2, generator=g, dtype=F64) * 6 - 0.5
    out = OC.altcorr_forward(f1, f2, coords, r)
    vol = torch.einsum("bhwc,bxyc->bhwxy", f1, f2)  # unscaled volume [B,H,W,H,W]
    for n in range(N):
        c = coords[:, n].permute(0, 3, 1, 2)
        ref = OC.corr_index_forward(vol, c, r).reshape(B, -1, H, W)
        assert (out[:, n] - ref).abs().max() < 1e-12


def test_indexed_altcorr_equals_corr_block_lookup():
    """Pooling is linear: looking up a pooled volume equals dotting pooled
    features (the reference's CorrBlock vs AltCorrBlock equivalence)."""
    g = torch.Generator().manual_seed(4)
    Nf, C, H, W = 3, 16, 8, 12
    fmaps = torch.randn(1, Nf, C, H, W, generator=g, dtype=F64) / math.sqrt(C)
    ii = torch.tensor([0, 1, 2])
    jj = torch.tensor([1, 2, 0])
    M = ii.shape[0]
    coords = torch.rand(1, M, H, W, 2, generator=g, dtype=F64) * torch.tensor([W, H], dtype=F64)
    pyr = OC.altcorr_pyramid(fmaps, 4)
    alt = OC.altcorr_index_forward(pyr[0], pyr, coords, ii, jj, 3)
    vol = OC.corr_volume(fmaps[:, ii], fmaps[:, jj])
    blk = OC.corr_block_lookup(OC.corr_pyramid(vol, 4), coords, 3)
    assert (alt - blk).abs().max() < 1e-12
