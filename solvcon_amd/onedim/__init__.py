"""One-dimensional Euler CESE core (drop-in for _solvcon.onedim)."""

from .euler1d import Euler1DCore  # noqa: F401

__all__ = ['Euler1DCore']
