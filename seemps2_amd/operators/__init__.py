"""Operators on device MPS (drop-in for ``seemps.operators``)."""

from .mpo import MPO
from .quadratic import QuadraticForm

__all__ = ["MPO", "QuadraticForm"]
