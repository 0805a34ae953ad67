"""Linear systems A x = b with MPO A and MPS b (drop-in for ``seemps.solve.dmrg``)."""

from .dmrg import dmrg_solve

__all__ = ["dmrg_solve"]
