"""Drop-in for ``vipe.ext.scatter`` (vipe/ext/scatter.py:12-205): the
torch_scatter-style helpers.  Every reduction, including ``scatter_sum`` /
``scatter_add`` / ``scatter_mean`` (which the reference runs through
``torch.Tensor.scatter_add_``), runs on the gfx950 ``vh_scatter`` kernels;
``scatter_ext`` documents the semantics and the differences."""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import scatter_ext
from .scatter_ext import broadcast  # noqa: F401  (same helper as the reference's)


def scatter_sum(src: torch.Tensor, index: torch.Tensor, dim: int = -1, out: Optional[torch.Tensor] = None,
                dim_size: Optional[int] = None) -> torch.Tensor:
    return scatter_ext.scatter_sum(src, index, dim, out, dim_size)


def scatter_add(src: torch.Tensor, index: torch.Tensor, dim: int = -1, out: Optional[torch.Tensor] = None,
                dim_size: Optional[int] = None) -> torch.Tensor:
    return scatter_sum(src, index, dim, out, dim_size)


def scatter_mul(src: torch.Tensor, index: torch.Tensor, dim: int = -1, out: Optional[torch.Tensor] = None,
                dim_size: Optional[int] = None) -> torch.Tensor:
    return scatter_ext.scatter_mul(src, index, dim, out, dim_size)


def scatter_mean(src: torch.Tensor, index: torch.Tensor, dim: int = -1, out: Optional[torch.Tensor] = None,
                 dim_size: Optional[int] = None) -> torch.Tensor:
    return scatter_ext.scatter_mean(src, index, dim, out, dim_size)


def scatter_min(src: torch.Tensor, index: torch.Tensor, dim: int = -1, out: Optional[torch.Tensor] = None,
                dim_size: Optional[int] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    return scatter_ext.scatter_min(src, index, dim, out, dim_size)


def scatter_max(src: torch.Tensor, index: torch.Tensor, dim: int = -1, out: Optional[torch.Tensor] = None,
                dim_size: Optional[int] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    return scatter_ext.scatter_max(src, index, dim, out, dim_size)


_REDUCE = {"sum": scatter_sum, "add": scatter_sum, "mul": scatter_mul, "mean": scatter_mean}


def scatter(src: torch.Tensor, index: torch.Tensor, dim: int = -1, out: Optional[torch.Tensor] = None,
            dim_size: Optional[int] = None, reduce: str = "sum") -> torch.Tensor:
    """Reduce src into out along dim at the positions index gives ("sum" / "add",
    "mul", "mean", "min", "max"); any other reduce raises ValueError."""
    if reduce in _REDUCE:
        return _REDUCE[reduce](src, index, dim, out, dim_size)
    if reduce == "min":
        return scatter_min(src, index, dim, out, dim_size)[0]
    if reduce == "max":
        return scatter_max(src, index, dim, out, dim_size)[0]
    raise ValueError(reduce)
