"""Ground-state search (drop-in for ``seemps.optimization.dmrg``)."""

from .dmrg import OptimizeResults, dmrg

__all__ = ["OptimizeResults", "dmrg"]
