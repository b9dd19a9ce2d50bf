"""corr_ext host-side behaviour (no GPU): the reference's argument checks and
the output-size rule through the C ABI."""
import pytest
import torch

from oracle import spatial_corr as OS


def test_host_tensors_raise_like_the_reference():
    from vipe_amd.ext import corr_ext

    a = torch.randn(1, 2, 5, 5)
    with pytest.raises(RuntimeError, match="input1 must be a CUDA tensor"):
        corr_ext.forward(a, a, 1, 1, 3, 3, 0, 0, 1, 1, 1, 1, 1, 1)
    with pytest.raises(RuntimeError, match="input1 must be a CUDA tensor"):
        corr_ext.backward(a, a, torch.zeros(1, 3, 3, 5, 5), 1, 1, 3, 3, 0, 0, 1, 1, 1, 1, 1, 1)


@pytest.mark.parametrize("args", [(11, 13, 3, 2, 2, 0, 2, 1, 2, 1), (30, 54, 1, 1, 0, 0, 1, 1, 1, 1),
                                  (7, 9, 5, 5, 0, 0, 2, 2, 3, 1), (4, 4, 9, 1, 0, 0, 1, 1, 1, 1)])
def test_output_size_matches_oracle(args):
    from vipe_amd.ext import corr_ext

    assert corr_ext.output_size(*args) == OS.out_size(*args)


def test_module_pairs_arguments():
    from vipe_amd.ext.corr import spatial_correlation_sampler as S

    assert S._pair(3) == (3, 3) and S._pair((2, 5)) == (2, 5)
    with pytest.raises(ValueError):
        S._pair((1, 2, 3))
