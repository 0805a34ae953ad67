"""Tensor products of MPS (reference: src/seemps/state/compose.py:11-115).

Order "A" concatenates the states' sites; order "B" interleaves them, taking
site k of every state before site k + 1 of any.  For "B" each state is first
padded to the joint sites with identity tensors that carry its current bond
(a bond-D identity times an all-ones physical vector), and the padded states
are multiplied site by site (``MPS * MPS``, the device site-product kernel),
optionally simplified.
"""

from __future__ import annotations

import numpy as np

from .. import kernels as K
from .mps import MPS
from .simplification import simplify


def _joint_sites(mps_list, order: str) -> list[tuple[int, int]]:
    """(state, site within the state) for every joint site."""
    if order == "A":
        return [(s, i) for s, mps in enumerate(mps_list) for i in range(mps.size)]
    if order == "B":
        longest = max(mps.size for mps in mps_list)
        return [(s, i) for i in range(longest) for s, mps in enumerate(mps_list)
                if i < mps.size]
    raise ValueError(f"Invalid mps order {order}")


def _padded_terms(mps_list, order: str) -> list[MPS]:
    sites = _joint_sites(mps_list, order)
    terms = []
    for s, mps in enumerate(mps_list):
        data, D = [], 1
        for owner, i in sites:
            if owner == s:
                A = mps[i]
                D = A.shape[-1]
            else:
                d = mps_list[owner][i].shape[1]
                A = np.broadcast_to(np.eye(D)[:, None, :], (D, d, D)).astype(np.complex128)
            data.append(A)
        terms.append(MPS(data))
    return terms


def mps_tensor_product(mps_list, mps_order: str = "A", strategy=None,
                       simplify_steps: bool = False) -> MPS:
    """Tensor product of the states in `mps_order` ("A" or "B"), simplified with
    `strategy` at the end, or after every product when `simplify_steps`
    (compose.py:73-115)."""
    if mps_order == "A":
        result = MPS([A for mps in mps_list for A in mps])
    elif mps_order == "B":
        terms = _padded_terms(mps_list, mps_order)
        result = terms[0]
        for term in terms[1:]:
            result = result * term
            if strategy and simplify_steps:
                result = simplify(result, strategy=strategy)
    else:
        raise ValueError(f"Invalid mps order {mps_order}")
    if strategy and not simplify_steps:
        result = simplify(result, strategy=strategy)
    return result


__all__ = ["mps_tensor_product"]
