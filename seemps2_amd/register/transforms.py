"""Shift operators on integers encoded in base-`base` registers (reference:
src/seemps/register/transforms.py:10-118).

The register holds s = sum_i s_i base^(L-1-i), site 0 the most significant.
Adding a shift r is done digit by digit from the least significant site: the
digit s_i plus the incoming carry c gives the new digit (s_i + c) mod base and
the outgoing carry (s_i + c) div base.  The MPO bond between two sites is the
carry; its values are those reachable from the requested shifts, so the bond
dimension stays tiny (2-3 for the usual finite-difference stencils).  The
last site's right bond indexes the shifts themselves (that is what
`mpo_weighted_shifts` contracts with the weights); the first site's left
bond keeps only carry 0 (open boundary) or sums all carries (periodic).
Tensors are built on the host (setup, a few hundred entries) and uploaded.
"""

from __future__ import annotations

import numpy as np

from ..operators.mpo import MPO


def _carry_tensors(L: int, shifts: np.ndarray, base: int):
    """Site tensors W[c_out, s_new, s_old, c_in] (site order, last site first
    built) and the carry values leaving the most significant site."""
    carries = np.asarray(shifts, dtype=np.int64)  # values entering the current site
    tensors = []
    digits = np.arange(base)
    for _ in range(L):
        total = digits[:, None] + carries[None, :]          # [s_old, carry_in]
        new_digit = np.mod(total, base)
        carry_out = np.floor_divide(total, base)
        out_values = np.unique(carry_out)                     # sorted
        out_index = np.searchsorted(out_values, carry_out)
        W = np.zeros((out_values.size, base, base, carries.size))
        for s_old in range(base):
            for j in range(carries.size):
                W[out_index[s_old, j], new_digit[s_old, j], s_old, j] = 1.0
        tensors.append(W)
        carries = out_values
    return tensors[::-1], carries


def mpo_shifts(L: int, shifts, periodic: bool = False, base: int = 2) -> MPO:
    """MPO mapping |s> to |s + r_j> on its last bond index j, for the shifts
    r_j (a list, or a (start, stop) range tuple) (transforms.py:51-118)."""
    if isinstance(shifts, tuple):
        r = np.arange(shifts[0], shifts[1], dtype=np.int64)
    else:
        r = np.asarray(shifts, dtype=np.int64)
    tensors, carries = _carry_tensors(L, r, base)
    first = tensors[0]
    if periodic:
        first = first.sum(axis=0, keepdims=True)    # wrap-around: any carry out
    else:
        zero = np.nonzero(carries == 0)[0]
        first = first[zero[:1]] if zero.size else 0.0 * first[:1]
    tensors[0] = first
    return MPO(tensors)


def mpo_weighted_shifts(L: int, weights, shifts, periodic: bool = False,
                        base: int = 2) -> MPO:
    """MPO for |s> -> sum_j w_j |s + r_j> (transforms.py:10-48)."""
    O = mpo_shifts(L, shifts, periodic, base)
    last = O[L - 1]
    a, i, j, m = last.shape
    w = np.asarray(weights, dtype=np.complex128).reshape(m, 1)
    from .. import kernels as K

    O[L - 1] = K.gemm(last.reshape(a * i * j, m), K.as_dev(w)).view(a, i, j, 1)
    return O


__all__ = ["mpo_shifts", "mpo_weighted_shifts"]
