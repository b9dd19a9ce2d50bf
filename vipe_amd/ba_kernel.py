"""Robust kernels of the BA solver with the interface of vipe/slam/ba/kernel.py:21-162.

``RobustKernel.apply(x)`` returns the per-residual IRLS weight through the
gfx950 ``vh_robust_weights`` kernel (float32; other float dtypes are computed
in float32 and cast back).  ``fused_args()`` gives the (type, threshold, mu,
mu_max) that ``slam_ext.ba_extended_robust`` and the fused BA caller pass to
``vh_ba_run_robust``, where the same weight multiplies each flow residual
component's weight inside the frame kernel.
"""
from __future__ import annotations

from abc import ABC, abstractmethod

import torch

from . import _lib

NONE, HUBER, TUKEY, GNC_TLS = 0, 1, 2, 3


def _weights(kind: int, threshold: float, mu: float, mu_max: float, x: torch.Tensor) -> torch.Tensor:
    if x.device.type != "cuda":
        raise RuntimeError(f"robust kernels run on the GPU; got a tensor on {x.device}")
    xf = x.to(torch.float32).contiguous()
    w = torch.empty_like(xf)
    with torch.cuda.device(x.device):
        rc = _lib.load().vh_robust_weights(kind, float(threshold), float(mu), float(mu_max), _lib.ptr(xf),
                                           _lib.ptr(w), xf.numel(), _lib.stream_of(xf))
    _lib.check(rc, "robust_weights")
    return w if x.dtype == torch.float32 else w.to(x.dtype)


class RobustKernel(ABC):
    @abstractmethod
    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """Return a per-residual weight of the same shape as ``x``."""

    @abstractmethod
    def fused_args(self) -> tuple:
        """(type, threshold, mu, mu_max) of vh_ba_run_robust."""

    def is_gnc(self) -> bool:
        return False

    def set_mu(self, mu: float) -> None:
        pass

    def update_mu(self, mu: float) -> float:
        return mu

    @property
    def mu_max(self) -> float:
        return float("inf")


class HuberRobustKernel(RobustKernel):
    """``w(r) = 1`` if ``|r| <= k`` else ``k / |r|``."""

    def __init__(self, threshold: float = 1.0) -> None:
        self.threshold = float(threshold)

    def apply(self, x):
        return _weights(HUBER, self.threshold, 1.0, 1.0, x)

    def fused_args(self):
        return (HUBER, self.threshold, 1.0, 1.0)


class TukeyRobustKernel(RobustKernel):
    """``w(r) = (1 - (r/c)^2)^2`` if ``|r| <= c`` else ``0``."""

    def __init__(self, c: float = 2.0) -> None:
        self.c = float(c)

    def apply(self, x):
        return _weights(TUKEY, self.c, 1.0, 1.0, x)

    def fused_args(self):
        return (TUKEY, self.c, 1.0, 1.0)


class TLSGncRobustKernel(RobustKernel):
    """Graduated non-convexity with truncated least squares (Yang, Antonante,
    Tzoumas, Carlone, RA-L 2020).  At ``mu`` with ``c2 = threshold^2``: weight 1
    below ``mu/(mu+1) c2``, 0 above ``(mu+1)/mu c2``, ``sqrt(c2 mu (mu+1) / r^2) - mu``
    between; ``mu >= mu_max`` (or no ``mu`` set) is the indicator ``r^2 < c2``.
    Schedule ``mu <- min(mu * mu_step, mu_max)``."""

    def __init__(self, threshold: float = 3.0, mu_step: float = 3.0, mu_init: float = 1.0,
                 mu_max: float = 1.0e6) -> None:
        self.c_bar = float(threshold)
        self.c_bar_sq = self.c_bar * self.c_bar
        self.mu_step = float(mu_step)
        self.mu_init = float(mu_init)
        self._mu_max = float(mu_max)
        self._mu: float | None = None

    def is_gnc(self) -> bool:
        return True

    @property
    def mu_max(self) -> float:
        return self._mu_max

    def set_mu(self, mu: float) -> None:
        self._mu = float(mu)

    def update_mu(self, mu: float) -> float:
        return min(mu * self.mu_step, self._mu_max)

    def _mu_now(self) -> float:
        return self._mu_max if self._mu is None else self._mu

    def apply(self, x):
        return _weights(GNC_TLS, self.c_bar, self._mu_now(), self._mu_max, x)

    def fused_args(self):
        return (GNC_TLS, self.c_bar, self._mu_now(), self._mu_max)


def build_robust_kernel(name, threshold: float, *, gnc_mu_init: float = 1.0, gnc_mu_step: float = 3.0,
                        gnc_mu_max: float = 1.0e6):
    """``None`` / ``"none"`` / ``"null"`` / ``"off"`` / ``""`` -> None (plain least squares)."""
    if name is None:
        return None
    n = str(name).lower()
    if n in ("none", "null", "off", ""):
        return None
    if n == "huber":
        return HuberRobustKernel(threshold=threshold)
    if n == "tukey":
        return TukeyRobustKernel(c=threshold)
    if n in ("gnc_tls", "gnc-tls"):
        return TLSGncRobustKernel(threshold=threshold, mu_init=gnc_mu_init, mu_step=gnc_mu_step,
                                  mu_max=gnc_mu_max)
    raise ValueError(f"Unknown robust_kernel: {name!r}; expected 'huber' | 'tukey' | 'gnc_tls' | None")
