"""Nearest-neighbour Hamiltonians and their MPOs (drop-in for ``seemps.hamiltonians``)."""

from .nearest_neighbor import (
    ConstantNNHamiltonian,
    ConstantTIHamiltonian,
    HeisenbergHamiltonian,
    NNHamiltonian,
    XXZHamiltonian,
)

__all__ = [
    "ConstantNNHamiltonian",
    "ConstantTIHamiltonian",
    "HeisenbergHamiltonian",
    "NNHamiltonian",
    "XXZHamiltonian",
]
