"""scatter_ext host-side logic (no GPU): index broadcast and the stride
collapse that lets a broadcast index reach the kernels without a copy."""
import pytest
import torch

from oracle import scatter as OSc


def test_broadcast_matches_oracle_rule():
    from vipe_amd.ext.scatter_ext import broadcast

    src = torch.zeros(3, 5, 4, 2)
    for idx, dim in ((torch.arange(5), 1), (torch.arange(3), 0), (torch.zeros(3, 5, dtype=torch.long), 1),
                     (torch.arange(2), -1)):
        b = broadcast(idx, src, dim)
        assert tuple(b.shape) == tuple(src.shape)
        assert torch.equal(b, torch.from_numpy(OSc.broadcast(idx.numpy(), tuple(src.shape), dim).copy()))


def test_index_strides_collapse_without_copy():
    from vipe_amd.ext.scatter_ext import _index_strides, broadcast

    src = torch.zeros(2, 3, 7, 4, 5)
    idx = broadcast(torch.arange(7), src, 2)
    same, sB, sE, sK = _index_strides(idx, 2)
    assert same.data_ptr() == idx.data_ptr() and (sB, sE, sK) == (0, 1, 0)
    full = torch.zeros(2, 3, 7, 4, 5, dtype=torch.long)
    _, sB, sE, sK = _index_strides(full, 2)
    assert (sB, sE, sK) == (7 * 20, 20, 1)
    # [B, E] index over trailing K: strides (E, 1, 0)
    part = broadcast(torch.zeros(6, 7, dtype=torch.long), torch.zeros(6, 7, 3), 1)
    _, sB, sE, sK = _index_strides(part, 1)
    assert (sB, sE, sK) == (7, 1, 0)
    # a transposed [E, B] index still collapses: (b, e) at b + 4 e
    t = torch.zeros(5, 4, dtype=torch.long).t()
    got, sB, sE, sK = _index_strides(broadcast(t, torch.zeros(4, 5, 6), 1), 1)
    assert got.data_ptr() == t.data_ptr() and (sB, sE, sK) == (1, 4, 0)
    # leading axes in swapped order cannot share one stride: made contiguous
    p = torch.zeros(3, 2, 5, dtype=torch.long).permute(1, 0, 2)
    got, sB, sE, sK = _index_strides(p, 2)
    assert got.is_contiguous() and (sB, sE, sK) == (5, 1, 0)


def test_host_tensors_raise():
    from vipe_amd.ext import scatter_ext

    with pytest.raises(RuntimeError, match="src must be CUDA tensor"):
        scatter_ext.scatter_max(torch.zeros(4), torch.zeros(4, dtype=torch.long), 0)
