"""Oracle: seeded input builders.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``random_mps``     -- ``state/factories.py:176-221`` (same draw order on an
                        explicit ``np.random.Generator``).
* ``product_state``  -- ``state/factories.py:12-42``.
* ``nn_mpo``         -- ``hamiltonians/nearest_neighbor.py:105-165`` (MPO of a
                        sum of two-site terms, bond 2 + rank(h)).
* ``xxz_interaction`` / ``heisenberg_interaction`` -- two-site terms of
                        ``nearest_neighbor.py:318-334`` and the XXZ chain.

Pinning: ``nn_mpo`` is pinned by dense equality with the Kronecker-sum
Hamiltonian (tests/test_oracle_algorithms.py), as the reference pins
``to_mpo`` (tests/test_operators/test_hamiltonians.py:92-101).  MPO tensor
entries themselves depend on the SVD gauge of degenerate singular values:
parity unpinned for raw MPO entries (only the operator is compared).
``random_mps`` is pinned by the reference's shape rules
(tests/test_states/test_random_mps.py:14-27,64-75) and by its draw order
against the generator's stream (tests/test_oracle_algorithms.py); the
reference package cannot be imported here (its ``__init__`` imports h5py) to
write golden tensors.
"""

from __future__ import annotations

import numpy as np

from . import core
from .algorithms import OMPS

SX = np.array([[0.0, 1.0], [1.0, 0.0]])
SZ = np.array([[1.0, 0.0], [0.0, -1.0]])
SY = -1j * SZ @ SX


def random_mps(dimensions, D=1, truncate=True, complex=False, rng=None) -> OMPS:
    """``factories.py:176-221``: bond rule and draw order of the reference."""
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
    return OMPS(tensors)


def product_state(vectors, length=None) -> OMPS:
    if length is not None:
        v = np.asarray(vectors)
        return OMPS([v.reshape(1, v.size, 1).copy() for _ in range(length)])
    return OMPS([np.asarray(v).reshape(1, -1, 1).copy() for v in vectors])


def heisenberg_interaction():
    """0.25 (sx sx + sy sy + sz sz), real (``nearest_neighbor.py:329``)."""
    return (0.25 * (np.kron(SX, SX) + np.kron(SY, SY) + np.kron(SZ, SZ))).real


def xxz_interaction(Jxy=1.0, Jz=0.5):
    """Jxy (Sx Sx + Sy Sy) + Jz Sz Sz with S = sigma / 2."""
    return (
        Jxy * 0.25 * (np.kron(SX, SX) + np.kron(SY, SY)) + Jz * 0.25 * np.kron(SZ, SZ)
    ).real


def nn_mpo(interactions: list, dims: list, strategy=core.DEFAULT_STRATEGY) -> list:
    """``nearest_neighbor.py:105-165``: MPO of sum_i h_{i,i+1}.

    Bond layout: index 0 = "nothing yet", 1 = "done", 2.. = the Schmidt
    channels of h_{i,i+1}.
    """
    size = len(dims)
    tensors = [np.zeros((2, d, d, 2)) for d in dims]
    for i in range(size - 1):
        di, dj = dims[i], dims[i + 1]
        h = np.asarray(interactions[i]).reshape(di, dj, di, dj).transpose(0, 2, 1, 3)
        U, s, Vh = core.svd(h.reshape(di * di, dj * dj))
        ds, _ = core.truncate(s, strategy)
        s = np.sqrt(s[:ds])
        A = tensors[i]
        a, j, k, b = A.shape
        B = np.zeros((a, j, k, b + ds), dtype=np.result_type(U, A))
        B[:, :, :, :b] = A
        B[0, :, :, 2:] = (U[:, :ds] * s).reshape(di, di, ds)
        B[1, :, :, 1] = np.eye(j)
        B[0, :, :, 0] = np.eye(j)
        tensors[i] = B
        A = tensors[i + 1]
        a, j, k, b = A.shape
        B = np.zeros((a + ds, j, k, b), dtype=np.result_type(Vh, A))
        B[:a] = A
        B[2:, :, :, 1] = (s[:, None] * Vh[:ds, :]).reshape(ds, dj, dj)
        B[1, :, :, 1] = np.eye(j)
        B[0, :, :, 0] = np.eye(j)
        tensors[i + 1] = B
    tensors[0] = tensors[0][[0]]
    tensors[-1] = tensors[-1][:, :, :, [1]]
    return tensors


def xxz_mpo(N, Jxy=1.0, Jz=0.5):
    return nn_mpo([xxz_interaction(Jxy, Jz)] * (N - 1), [2] * N)


def heisenberg_mpo(N):
    return nn_mpo([heisenberg_interaction()] * (N - 1), [2] * N)
