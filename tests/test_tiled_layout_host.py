"""Tiled correlation-plane geometry on the host: tile / untile invert, the tile
padding holds zeros, and elements sit where the C header's layout puts them
(planes grouped, 128-byte tiles of the group's rows x tw columns)."""
import pytest
import torch

from vipe_amd.ext import droid_net_ext as D


def _offset(p, y, x, Hp, Wp, tw):
    """Element offset of plane p, row y, column x (VH_LAYOUT_TILED)."""
    G, R = D.TILE_PLANES, D.TILE_ROWS
    return ((p // G) * G * Hp * Wp + ((y // R) * (Wp // tw) + x // tw) * R * G * tw + ((y % R) * G + p % G) * tw +
            x % tw)


def test_tile_untile_round_trip_and_element_positions():
    g = torch.Generator().manual_seed(0)
    for dtype in (torch.float32, torch.float16):
        for Hl, Wl in ((64, 80), (13, 9), (8, 10), (1, 1)):
            v = torch.randn(1, 2, 4, Hl, Wl, generator=g).to(dtype)
            t = D.tile(v)
            Hp, Wp, tw = D.tiled_plane(Hl, Wl, dtype)
            assert Hp % D.TILE_ROWS == 0 and Hp - Hl < D.TILE_ROWS and Wp % tw == 0
            assert D.TILE_BYTES == 16 * D.TILE_PLANES * D.TILE_ROWS
            assert D.TILE_PLANES * Hp * Wp * t.element_size() % D.TILE_BYTES == 0
            assert tw * t.element_size() == 16
            assert t.shape[-1] == Hp * Wp
            assert torch.equal(D.untile(t, Hl, Wl), v)
            flat, planes = t.reshape(-1), v.reshape(8, Hl, Wl)
            for p in range(8):
                for y in range(Hp):
                    for x in range(0, Wp, max(1, Wp // 7)):
                        want = planes[p, y, x] if (y < Hl and x < Wl) else torch.zeros((), dtype=dtype)
                        assert flat[_offset(p, y, x, Hp, Wp, tw)] == want, (dtype, Hl, Wl, p, y, x)


def test_tiled_layout_needs_whole_plane_groups_per_edge():
    v = torch.zeros(2, 3, 3, 4, 4)
    with pytest.raises(RuntimeError, match="multiple"):
        D.tile(v)
    with pytest.raises(RuntimeError, match="multiple"):
        D.untile(torch.zeros(2, 3, 3, 16), 4, 4)
