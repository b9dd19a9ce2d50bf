"""Pins oracle/scatter.py (host only) against torch.Tensor.scatter_reduce_,
the reference docstring's shape example, and first-occurrence arg min / max."""
import numpy as np
import pytest
import torch

from oracle import scatter as OSc

TORCH_REDUCE = {"sum": "sum", "mul": "prod", "min": "amin", "max": "amax", "mean": "mean"}


@pytest.mark.parametrize("reduce", ["sum", "mul", "min", "max", "mean"])
@pytest.mark.parametrize("dim", [0, 1, -1])
def test_matches_scatter_reduce(reduce, dim):
    g = torch.Generator().manual_seed(5)
    src = torch.rand(4, 9, 3, generator=g, dtype=torch.float64) + 0.5
    d = dim % 3
    idx = torch.randint(0, 5, (src.shape[d],), generator=g)
    index = OSc.broadcast(idx.numpy(), tuple(src.shape), d)
    out, _ = OSc.scatter(src.numpy(), idx.numpy(), dim, reduce=reduce)
    size = list(src.shape)
    size[d] = int(idx.max()) + 1
    ident = {"sum": 0.0, "mean": 0.0, "mul": 1.0, "min": 0.0, "max": 0.0}[reduce]
    ref = torch.full(size, ident, dtype=torch.float64)
    ref.scatter_reduce_(d, torch.from_numpy(np.ascontiguousarray(index)), src, TORCH_REDUCE[reduce],
                        include_self=False)
    np.testing.assert_allclose(out, ref.numpy(), rtol=1e-13)


def test_docstring_shape_and_given_out():
    src = np.random.default_rng(0).standard_normal((10, 6, 64))
    index = np.array([0, 1, 0, 1, 2, 1])
    out, _ = OSc.scatter(src, index, 1, reduce="sum")
    assert out.shape == (10, 3, 64)
    np.testing.assert_allclose(out[:, 1], src[:, 1] + src[:, 3] + src[:, 5])
    base = np.ones((10, 3, 64))
    out2, _ = OSc.scatter(src, index, 1, out=base, reduce="sum")
    np.testing.assert_allclose(out2, out + 1)


@pytest.mark.parametrize("reduce", ["min", "max"])
def test_arg_is_first_occurrence_and_empty_is_zero(reduce):
    src = np.array([3.0, 1.0, 1.0, 5.0, 5.0, 2.0, np.nan], dtype=np.float32)
    index = np.array([0, 0, 0, 2, 2, 2, 3])
    out, arg = OSc.scatter(src, index, 0, dim_size=5, reduce=reduce)
    if reduce == "min":
        assert out.tolist() == [1.0, 0.0, 2.0, 0.0, 0.0] and arg.tolist() == [1, 7, 5, 7, 7]
    else:
        assert out.tolist() == [3.0, 0.0, 5.0, 0.0, 0.0] and arg.tolist() == [0, 7, 3, 7, 7]
    # the identity itself in a segment also reads as empty (the reference's masked fill)
    big = np.array([np.finfo(np.float32).max], dtype=np.float32)
    o, a = OSc.scatter(big, np.array([0]), 0, reduce="min")
    assert o.tolist() == [0.0] and a.tolist() == [1]


def test_integer_products_wrap_and_mean_floors():
    src = np.array([100, 3, -7, 9], dtype=np.int8)
    out, _ = OSc.scatter(src, np.array([0, 0, 1, 1]), 0, reduce="mul")
    assert out.dtype == np.int8 and out.tolist() == [np.int8(44), np.int8(-63)]   # 300 mod 256 = 44
    m, _ = OSc.scatter(np.array([1, 2, 4, -3], dtype=np.int64), np.array([0, 0, 1, 1]), 0, reduce="mean")
    assert m.tolist() == [1, 0]                                                 # floor(3/2), floor(1/2)
