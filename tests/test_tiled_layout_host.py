"""Tiled correlation-plane geometry on the host: tile / untile invert and the
tile padding holds zeros."""
import torch

from vipe_amd.ext import droid_net_ext as D


def test_tile_untile_round_trip():
    g = torch.Generator().manual_seed(0)
    for dtype in (torch.float32, torch.float16):
        for Hl, Wl in ((64, 80), (13, 9), (8, 10), (1, 1)):
            v = torch.randn(1, 3, 2, Hl, Wl, generator=g).to(dtype)
            t = D.tile(v)
            Hp, Wp, tw = D.tiled_plane(Hl, Wl, dtype)
            assert Hp % 4 == 0 and Wp % tw == 0 and Hp * Wp * t.element_size() % 64 == 0
            assert t.shape[-1] == Hp * Wp
            assert torch.equal(D.untile(t, Hl, Wl), v)
            # the padding rows and columns of each plane are zero
            full = t.view(1, 3, 2, Hp // 4, Wp // tw, 4, tw).permute(0, 1, 2, 3, 5, 4, 6).reshape(1, 3, 2, Hp, Wp)
            assert torch.all(full[..., Hl:, :] == 0) and torch.all(full[..., :, Wl:] == 0)
