"""Drop-in for ``vipe.ext`` on MI355X: ``slam_ext``, ``droid_net_ext``,
``lietorch_ext``, ``corr_ext`` and ``scatter_ext`` with the reference's
function names and signatures (vipe/ext/__init__.py:39-45), backed by
libvipe_hip.so; ``corr`` holds the spatial correlation sampler module
(vipe/ext/corr) and ``scatter`` the torch_scatter-style helpers
(vipe/ext/scatter.py)."""
from . import corr_ext, droid_net_ext, lietorch_ext, scatter_ext, slam_ext  # noqa: F401

__all__ = ["slam_ext", "droid_net_ext", "lietorch_ext", "corr_ext", "scatter_ext"]
