"""Seeded state builders (reference: src/seemps/state/factories.py:12-238).

Draws happen on the host with the caller's ``np.random.Generator`` in the
reference's order (``normal(size=(DL, d, DR))`` then the imaginary part), and
the tensors are uploaded once.
"""

from __future__ import annotations

from collections.abc import Iterable

import numpy as np

from .mps import MPS

SEED = 0x874665212
DEFAULT_RNG = np.random.default_rng(SEED)


def product_state(vectors, length: int | None = None) -> MPS:
    """Product state MPS (factories.py:12-42)."""

    def to_tensor(v):
        v = np.asarray(v)
        if v.ndim != 1:
            raise ValueError("Each site vector must be one-dimensional")
        return v.reshape(1, v.size, 1)

    if length is not None:
        return MPS([to_tensor(vectors)] * length)
    if isinstance(vectors, Iterable):
        return MPS([to_tensor(v) for v in list(vectors)])
    raise TypeError("Invalid argument to product_state")


def random_mps(dimensions: list[int], D: int = 1, truncate: bool = True, complex: bool = False,
               rng: np.random.Generator = DEFAULT_RNG) -> MPS:
    """Random MPS with bond dimension up to D (factories.py:176-221)."""
    N = len(dimensions)
    tensors = []
    DR = 1
    if N > 60:
        truncate = False
    for i, d in enumerate(dimensions):
        DL = DR
        if not truncate and i != N - 1:
            DR = D
        else:
            DR = min(DR * d, D, d ** (N - i - 1))
        T = rng.normal(size=(DL, d, DR))
        if complex:
            T = T + 1j * rng.normal(size=T.shape)
        tensors.append(T)
    return MPS(tensors)


def random_uniform_mps(d: int, N: int, D: int = 1, truncate: bool = True, complex: bool = False,
                       rng: np.random.Generator = DEFAULT_RNG) -> MPS:
    """``random_mps([d] * N, ...)`` (factories.py:147-173)."""
    return random_mps([d] * N, D, truncate, complex, rng)


def mps_ones(dimensions: list[int], normalize: bool = False) -> MPS:
    return MPS([np.ones((1, d, 1)) / (np.sqrt(d) if normalize else 1) for d in dimensions])
