"""Quantum-register transforms (drop-in for ``seemps.register``, the parts the
function-loading inputs need)."""

from .transforms import mpo_shifts, mpo_weighted_shifts

__all__ = ["mpo_shifts", "mpo_weighted_shifts"]
