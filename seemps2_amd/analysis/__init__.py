"""Function-loading inputs (drop-in for the parts of ``seemps.analysis`` that
build the finite-difference operators of the function-loading workloads)."""

from .mesh import (
    Interval,
    IntervalTuple,
    QuantizedInterval,
    RegularInterval,
    interleaving_permutation,
)
from .finite_differences import finite_differences_mpo, tridiagonal_mpo

__all__ = [
    "Interval",
    "IntervalTuple",
    "QuantizedInterval",
    "RegularInterval",
    "finite_differences_mpo",
    "interleaving_permutation",
    "tridiagonal_mpo",
]
