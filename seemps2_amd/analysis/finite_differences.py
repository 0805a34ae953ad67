"""Finite-difference operators as MPOs (reference:
src/seemps/analysis/derivatives/finite_differences.py:9-150).

Both are built on the host (a handful of small tensors) and uploaded.
`finite_differences_mpo` is a weighted sum of register shifts; its stencils
are Holoborodko's smooth noise-robust differentiators
(http://www.holoborodko.com/pavel/numerical-methods/numerical-derivative/
smooth-low-noise-differentiators) for the first derivative and the
corresponding smoothed second-difference stencils, with the same coefficient
tables as the reference (the tables fix the operator, so they are kept
identical).
"""

from __future__ import annotations

import numpy as np

from ..core import DEFAULT_STRATEGY, Strategy
from ..operators.mpo import MPO
from ..register.transforms import mpo_weighted_shifts
from .mesh import QuantizedInterval

# (order, stencil size) -> (weights, shifts): the operator maps f to
# sum_k weights[k] f(x + shifts[k] dx) / dx^order.
_STENCILS = {
    (1, 3): ([-1 / 2, 1 / 2], [1, -1]),
    (1, 5): ([-1 / 8, -2 / 8, 2 / 8, 1 / 8], [2, 1, -1, -2]),
    (1, 7): ([-1 / 32, -4 / 32, -5 / 32, 5 / 32, 4 / 32, 1 / 32], [3, 2, 1, -1, -2, -3]),
    (1, 9): ([-1 / 128, -6 / 128, -14 / 128, -14 / 128, 14 / 128, 14 / 128, 6 / 128,
              1 / 128], [4, 3, 2, 1, -1, -2, -3, -4]),
    (1, 11): ([-1 / 512, -8 / 512, -27 / 512, -48 / 512, -42 / 512, 42 / 512, 48 / 512,
               27 / 512, 8 / 512, 1 / 512], [4, 3, 2, 1, -1, -2, -3, -4]),
    (2, 3): ([1, -2, 1], [-1, 0, 1]),
    (2, 5): ([1 / 4, -2 / 4, 1 / 4], [-2, 0, 2]),
    (2, 7): ([1 / 16, 2 / 16, -1 / 16, -4 / 16, -1 / 16, 2 / 16, 1 / 16],
             [-3, -2, -1, 0, 1, 2, 3]),
    (2, 9): ([1 / 64, 4 / 64, 4 / 64, -4 / 64, -10 / 64, -4 / 64, 4 / 64, 4 / 64, 1 / 64],
             [-4, -3, -2, -1, 0, 1, 2, 3, 4]),
}


def tridiagonal_mpo(n: int, a: float, b: float, c: float, closed: bool = True,
                    strategy: Strategy = DEFAULT_STRATEGY) -> MPO:
    """MPO of the tridiagonal matrix a I + b S^+ + c S^- on a register of n
    qubits (S^+ adds one, S^- subtracts one), with the wrap-around terms when
    `closed` (finite_differences.py:9-40).  Bond channels: 0 = carry settled,
    1 = +1 pending, 2 = -1 pending."""
    A = np.zeros((3, 2, 2, 3))
    A[0, 0, 0, 0] = A[0, 1, 1, 0] = 1.0      # identity, nothing pending
    A[0, 1, 0, 1] = A[1, 0, 1, 1] = 1.0      # +1: 0 -> 1 settles, 1 -> 0 carries
    A[2, 1, 0, 2] = A[0, 0, 1, 2] = 1.0      # -1: 1 -> 0 settles, 0 -> 1 borrows
    last = a * A[:, :, :, [0]] + b * A[:, :, :, [1]] + c * A[:, :, :, [2]]
    first = A.sum(axis=0, keepdims=True) if closed else A[[0]]
    return MPO([first] + [A] * (n - 2) + [last], strategy)


def finite_differences_mpo(order: int, interval, filter: int = 3, periodic: bool = False,
                           base: int = 2, tol: float = 1e-4) -> MPO:
    """Derivative of the given order on the register grid of `interval`
    (finite_differences.py:91-146); steps below `tol` are doubled (and the
    shifts scaled to match) to keep the weights finite."""
    if isinstance(interval, tuple):
        interval = QuantizedInterval(*interval)
    stencil = _STENCILS.get((order, filter))
    if stencil is None:
        raise ValueError(f"Unknown finite difference derivative of order {order} "
                         f"with noise filter of size {filter}")
    weights, shifts = stencil
    dx = interval.step
    stride = 1
    while dx < tol:
        dx *= 2
        stride *= 2
    return mpo_weighted_shifts(interval.qubits, np.asarray(weights) / dx ** order,
                               [stride * s for s in shifts], periodic=periodic, base=base)


__all__ = ["finite_differences_mpo", "tridiagonal_mpo"]
