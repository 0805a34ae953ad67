"""Quantum-register transforms (drop-in for ``seemps.register``, the parts the
function-loading inputs need)."""

from .circuit import TwoQubitGatesLayer, interpret_operator, known_operators
from .transforms import mpo_shifts, mpo_weighted_shifts

__all__ = ["TwoQubitGatesLayer", "interpret_operator", "known_operators", "mpo_shifts",
           "mpo_weighted_shifts"]
