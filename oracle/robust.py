"""FP64 restatement of the robust kernels of the generic BA solver (oracle; test-only).

Follows vipe/slam/ba/kernel.py:
  HuberRobustKernel.apply    :41-51   w = 1 if |r| <= k else k / max(|r|, 1e-8)
  TukeyRobustKernel.apply    :54-64   w = (max(1 - (r/c)^2, 0))^2
  TLSGncRobustKernel.apply   :66-131  at mu with c2 = c_bar^2:
                                      r^2 < mu/(mu+1) c2 -> 1; r^2 <= (mu+1)/mu c2 ->
                                      max(sqrt(c2 mu (mu+1) / max(r^2, 1e-12)) - mu, 0);
                                      else 0; mu >= mu_max (or unset) -> 1{r^2 < c2}
  schedule                   :104-105 mu <- min(mu * mu_step, mu_max)
  build_robust_kernel        :134-162 names none/null/off/'' -> None, huber, tukey, gnc_tls
and their use as a factor on each residual component's weight,
vipe/slam/ba/terms.py:73-75.  Pinned by tests/test_oracle_robust.py: each weight
equals rho'(r) / (rho_0'' r) of its loss (Huber's, Tukey's biweight, the GNC
surrogate of truncated least squares of Yang et al. 2020), checked by central
differences of the loss, plus the GNC limits and continuity at the band edges.
"""
from __future__ import annotations

import torch

NONE, HUBER, TUKEY, GNC_TLS = 0, 1, 2, 3


def weight(r: torch.Tensor, kind: int, c: float, mu: float = 1.0, mu_max: float = 1.0e6) -> torch.Tensor:
    r = r.to(torch.float64)
    if kind == NONE:
        return torch.ones_like(r)
    if kind == HUBER:
        a = r.abs()
        return torch.where(a <= c, torch.ones_like(r), c / a.clamp_min(1e-8))
    if kind == TUKEY:
        return (1.0 - (r / c) ** 2).clamp(min=0.0) ** 2
    if kind == GNC_TLS:
        r2 = r * r
        c2 = c * c
        if mu >= mu_max:
            return (r2 < c2).to(torch.float64)
        lo = mu / (mu + 1.0) * c2
        hi = (mu + 1.0) / mu * c2
        mid = (torch.sqrt(c2 * mu * (mu + 1.0) / r2.clamp_min(1e-12)) - mu).clamp(min=0.0)
        return torch.where(r2 < lo, torch.ones_like(r), torch.where(r2 <= hi, mid, torch.zeros_like(r)))
    raise ValueError(kind)


def gnc_schedule(mu_init: float, mu_step: float, mu_max: float, n_steps: int):
    """The mu of each outer GNC step (buffer.py:686-695 with kernel.py:104-105)."""
    mus, mu = [], float(mu_init)
    for _ in range(n_steps):
        mus.append(mu)
        mu = min(mu * mu_step, mu_max)
    return mus
