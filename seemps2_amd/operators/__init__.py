"""Operators on device MPS (drop-in for ``seemps.operators``)."""

from .mpo import MPO
from .mposum import MPOSum
from .quadratic import QuadraticForm

__all__ = ["MPO", "MPOSum", "QuadraticForm"]
