"""Uniform grids on quantized registers (reference: src/seemps/analysis/mesh.py:11-177,
410-426).  Host-side bookkeeping: no tensors live here."""

from __future__ import annotations

from typing import TypeAlias

import numpy as np


class Interval:
    """`size` points between `start` and `stop`, indexable and iterable
    (mesh.py:11-75)."""

    def __init__(self, start: float, stop: float, size: int):
        self.start = start
        self.stop = stop
        self.size = size

    def __len__(self) -> int:
        return self.size

    def _validate_index(self, idx):
        if isinstance(idx, (int, np.integer)):
            idx = int(idx)
            if idx < 0:
                idx += self.size
            if not 0 <= idx < self.size:
                raise IndexError("Index out of range")
            return idx
        if isinstance(idx, np.ndarray):
            idx = np.where(idx < 0, idx + self.size, idx)
            if not np.all((idx >= 0) & (idx < self.size)):
                raise IndexError("Index out of range")
            return idx
        raise TypeError("Index must be an integer or a NumPy array")

    def __getitem__(self, idx):
        raise NotImplementedError

    def __iter__(self):
        return (self[i] for i in range(self.size))

    def to_vector(self) -> np.ndarray:
        return np.array([self[i] for i in range(self.size)])

    def map_to(self, start: float, stop: float) -> "Interval":
        return type(self)(start, stop, self.size)


class RegularInterval(Interval):
    """Equispaced points; [start, stop) by default, either end open or closed
    (mesh.py:100-150)."""

    def __init__(self, start: float, stop: float, size: int, endpoint_left: bool = True,
                 endpoint_right: bool = False):
        super().__init__(start, stop, size)
        self.endpoint_left = endpoint_left
        self.endpoint_right = endpoint_right
        # number of steps between the two (possibly excluded) endpoints
        self.num_steps = size - 1 + (0 if endpoint_left else 1) + (0 if endpoint_right else 1)
        self.step = (stop - start) / self.num_steps
        self._first = start if endpoint_left else start + self.step

    def __getitem__(self, idx):
        return self._first + self._validate_index(idx) * self.step

    def map_to(self, start: float, stop: float) -> "RegularInterval":
        return RegularInterval(start, stop, self.size, self.endpoint_left, self.endpoint_right)


class QuantizedInterval(RegularInterval):
    """RegularInterval with 2**qubits points (mesh.py:155-177)."""

    def __init__(self, start: float, stop: float, qubits: int, endpoint_left: bool = True,
                 endpoint_right: bool = False):
        if not (isinstance(qubits, (int, np.integer)) and qubits > 0):
            raise ValueError("qubits must be a positive integer")
        super().__init__(start, stop, 2 ** int(qubits), endpoint_left, endpoint_right)
        self.qubits = int(qubits)

    def map_to(self, start: float, stop: float) -> "QuantizedInterval":
        return QuantizedInterval(start, stop, self.qubits, self.endpoint_left,
                                 self.endpoint_right)


#: (a, b, n): the semi-open QuantizedInterval [a, b) with n qubits.
IntervalTuple: TypeAlias = tuple[float, float, int]


def interleaving_permutation(sites_per_dimension: list[int]) -> np.ndarray:
    """Site order that takes the k-th site of every dimension before the
    (k+1)-th of any ("B" order), dimensions in their given order
    (mesh.py:410-426)."""
    starts = np.concatenate([[0], np.cumsum(sites_per_dimension)[:-1]]).astype(int)
    order = [starts[dim] + k
             for k in range(max(sites_per_dimension))
             for dim, n in enumerate(sites_per_dimension) if k < n]
    return np.array(order, dtype=int)


__all__ = ["Interval", "IntervalTuple", "QuantizedInterval", "RegularInterval",
           "interleaving_permutation"]
