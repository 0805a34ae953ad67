"""Closed forms that pin the oracle (TEST INFRASTRUCTURE ONLY).

* ``mpo_to_matrix``        -- dense operator of an MPO (``mpo.py:192-207``).
* ``nn_hamiltonian_dense`` -- Kronecker sum of two-site terms
                              (``nearest_neighbor.py:77-103``).
* ``ground_energy_dense``  -- exact diagonalisation.
* ``xx_free_fermion_energy`` -- ground energy of the open XX chain,
  H = Jxy sum (Sx Sx + Sy Sy) = (Jxy/2) sum (c+_i c_i+1 + h.c.) after a
  Jordan-Wigner transform: the sum of the negative eigenvalues
  Jxy cos(k pi / (N+1)), k = 1..N.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg


def mpo_to_matrix(mpo: list) -> np.ndarray:
    out = np.ones((1, 1, 1))
    Di = Dj = 1
    for W in mpo:
        _, i, j, b = W.shape
        Di *= i
        Dj *= j
        out = np.einsum("xya,aijb->xiyjb", out, W).reshape(Di, Dj, b)
    return out[:, :, 0]


def nn_hamiltonian_sparse(interactions: list, dims: list):
    H = sp.csr_matrix((dims[0], dims[0]))
    dleft = 1
    for i in range(len(dims) - 1):
        H = sp.kron(H, sp.eye(dims[i + 1]))
        H = H + sp.kron(sp.eye(dleft), sp.csr_matrix(interactions[i]))
        dleft *= dims[i]
    return H.tocsr()


def ground_energy_sparse(H) -> float:
    if H.shape[0] <= 64:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(scipy.sparse.linalg.eigsh(H, k=1, which="SA", tol=1e-14)[0][0])


def xx_free_fermion_energy(N: int, Jxy: float = 1.0) -> float:
    k = np.arange(1, N + 1)
    e = Jxy * np.cos(k * np.pi / (N + 1))
    return float(np.sum(e[e < 0]))
