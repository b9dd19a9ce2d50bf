"""Lie groups on gfx950: ``from vipe_amd.lietorch import SE3, Sim3, ...``
(the interface of ``vipe.ext.lietorch``)."""
from .groups import SE3, SO3, LieGroupParameter, RxSO3, Sim3, cat, stack

__all__ = ["SE3", "SO3", "RxSO3", "Sim3", "LieGroupParameter", "cat", "stack"]
