"""Seismic norms accepted by the TTI models (agd/Metrics/Seismic)."""
from .tti import TTI  # noqa: F401

__all__ = ["TTI"]
