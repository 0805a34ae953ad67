"""Layers of two-qubit gates (reference: src/seemps/register/circuit.py:16-53,
223-289).  A layer applies the same 4 x 4 gate to every neighbouring pair in
one sweep, with the truncation of the layer's Strategy, through the same
device path as TEBD (``seemps2_amd.evolution.trotter``)."""

from __future__ import annotations

import numpy as np

from ..core import DEFAULT_STRATEGY, Strategy
from ..state import CanonicalMPS, MPS

_sx = np.array([[0.0, 1.0], [1.0, 0.0]])
_sz = np.array([[1.0, 0.0], [0.0, -1.0]])
_sy = -1j * _sz @ _sx
_cnot = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0], [0, 0, 1.0, 0]])

#: Named operators (circuit.py:27-42; the same names and matrices).
known_operators = {
    "sx": _sx / 2.0, "sy": _sy / 2.0, "sz": _sz / 2.0,
    "σx": _sx, "σy": _sy, "σz": _sz,
    "cnot": _cnot, "cx": _cnot, "cz": np.diag([1.0, 1.0, 1.0, -1.0]),
    "h": np.asarray([[1, 1], [1, -1]]) / np.sqrt(2),
}


def interpret_operator(op) -> np.ndarray:
    """A named operator or a square matrix (circuit.py:45-53)."""
    if isinstance(op, str):
        return known_operators[op.lower()]
    if not isinstance(op, np.ndarray) or op.ndim != 2 or op.shape[0] != op.shape[1]:
        raise TypeError(f"Invalid qubit operator of type '{type(op)}")
    return op


class TwoQubitGatesLayer:
    """The gate on qubits (0, 1), (1, 2), ..., in one sweep whose direction is
    given or follows the centre of the state (circuit.py:223-289)."""

    def __init__(self, register_size: int, operator, direction: int | None = None,
                 strategy: Strategy = DEFAULT_STRATEGY):
        O = interpret_operator(operator)
        if O.shape != (4, 4):
            raise ValueError("Not a valid two-qubit operator")
        self.register_size = register_size
        self.operator = np.asarray(O, dtype=np.complex128)
        self.direction = direction
        self.strategy = strategy

    def apply(self, state: MPS, parameters=None) -> CanonicalMPS:
        return self.apply_inplace(state.copy() if isinstance(state, CanonicalMPS) else state,
                                  parameters)

    def apply_inplace(self, state: MPS, parameters=None) -> CanonicalMPS:
        from ..evolution.trotter import _sweep_gates

        if self.register_size != state.size:
            raise ValueError("Circuit register size does not match the state size")
        if parameters is not None and len(parameters) > 0:
            raise ValueError(f"{type(self).__name__} does not accept parameters")
        gates = [self.operator] * (self.register_size - 1)
        return _sweep_gates(state, gates, self.strategy, self.direction)

    def __matmul__(self, state: MPS) -> CanonicalMPS:
        return self.apply(state)


__all__ = ["TwoQubitGatesLayer", "interpret_operator", "known_operators"]
