"""TEBD: chains of two-site unitaries applied with truncation (reference:
src/seemps/evolution/trotter.py:10-260).

Each gate exp(-i dt h_{j,j+1}) is a small dense matrix built on the host
(scipy's expm of the interaction term, once per step size) and uploaded.  A
sweep applies them in order along the chain: the two-site tensor is
U . (A_j A_{j+1}) (``_contract_nrjl_ijk_klm``: one GEMM and one local
contraction on the device) and is split back with the Strategy's truncation
(``update_2site_right``/``update_2site_left``), moving the centre with it.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg

from ..core import DEFAULT_STRATEGY, Strategy, _contract_nrjl_ijk_klm
from ..state import CanonicalMPS, MPS


def _sweep_gates(state, gates, strategy: Strategy, direction: int | None = None):
    """Apply gates[j] on sites (j, j+1), left to right when the centre is in
    the left half (or direction > 0), right to left otherwise
    (trotter.py:55-91)."""
    if not isinstance(state, CanonicalMPS):
        state = CanonicalMPS(state, center=0, strategy=strategy)
    L = state.size
    if direction is None:
        direction = +1 if state.center < L // 2 else -1
    if direction > 0:
        if state.center > 1:
            state.recenter(1)
        for j in range(L - 1):
            state.update_2site_right(_contract_nrjl_ijk_klm(gates[j], state[j], state[j + 1]), j,
                                     strategy)
    else:
        if state.center < L - 2:
            state.recenter(L - 2)
        for j in range(L - 2, -1, -1):
            state.update_2site_left(_contract_nrjl_ijk_klm(gates[j], state[j], state[j + 1]), j,
                                    strategy)
    return state


class PairwiseUnitaries:
    """U_j = exp(-i dt h_{j,j+1}) for the interactions of an NNHamiltonian
    (trotter.py:10-91)."""

    def __init__(self, H, dt: float, strategy: Strategy):
        self.U = [np.asarray(scipy.linalg.expm(
            (-1j * dt) * np.asarray(H.interaction_term(k).todense()
                                    if hasattr(H.interaction_term(k), "todense")
                                    else H.interaction_term(k))), dtype=np.complex128)
                  for k in range(H.size - 1)]
        self.strategy = strategy

    def apply(self, state: MPS) -> CanonicalMPS:
        return self.apply_inplace(state.copy() if isinstance(state, CanonicalMPS) else state)

    def apply_inplace(self, state: MPS) -> CanonicalMPS:
        return _sweep_gates(state, self.U, self.strategy)

    def __matmul__(self, state: MPS) -> CanonicalMPS:
        return self.apply(state)


class Trotter2ndOrder:
    """Two sweeps of exp(-i dt/2 h_{j,j+1}), forth and back (trotter.py:121-179)."""

    def __init__(self, H, dt: float, strategy: Strategy = DEFAULT_STRATEGY):
        self.U = PairwiseUnitaries(H, 0.5 * dt, strategy)

    def apply(self, state: MPS) -> CanonicalMPS:
        return self.U.apply_inplace(self.U.apply(state))

    def apply_inplace(self, state: MPS) -> CanonicalMPS:
        return self.U.apply_inplace(self.U.apply_inplace(state))

    def __matmul__(self, state: MPS) -> CanonicalMPS:
        return self.apply(state)


class Trotter3rdOrder:
    """Sweeps of exp(-i dt/4 h), exp(-i dt/2 h), exp(-i dt/4 h) (trotter.py:200-260)."""

    def __init__(self, H, dt: float, strategy: Strategy = DEFAULT_STRATEGY):
        self.Umid = PairwiseUnitaries(H, 0.5 * dt, strategy)
        self.U = PairwiseUnitaries(H, 0.25 * dt, strategy)

    def apply(self, state: MPS) -> CanonicalMPS:
        state = self.U.apply(state)
        state = self.Umid.apply_inplace(state)
        return self.U.apply_inplace(state)

    def apply_inplace(self, state: MPS) -> CanonicalMPS:
        state = self.U.apply_inplace(state)
        state = self.Umid.apply_inplace(state)
        return self.U.apply_inplace(state)

    def __matmul__(self, state: MPS) -> CanonicalMPS:
        return self.apply(state)


__all__ = ["PairwiseUnitaries", "Trotter2ndOrder", "Trotter3rdOrder", "_sweep_gates"]
