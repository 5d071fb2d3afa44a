"""One-dimensional Euler CESE solver (drop-in for solvcon.onedim)."""

from .euler1d import Euler1DCore, Euler1DSolver, ShockTube  # noqa: F401

__all__ = ['Euler1DCore', 'Euler1DSolver', 'ShockTube']
