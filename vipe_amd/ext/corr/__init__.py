"""Drop-in for ``vipe.ext.corr`` (vipe/ext/corr/__init__.py): the spatial
correlation sampler as an autograd function and an ``nn.Module``."""
from .spatial_correlation_sampler import (  # noqa: F401
    SpatialCorrelationSampler,
    SpatialCorrelationSamplerFunction,
    spatial_correlation_sample,
)

__all__ = ["SpatialCorrelationSampler", "SpatialCorrelationSamplerFunction", "spatial_correlation_sample"]
