"""Nearest-neighbour Hamiltonians (reference: src/seemps/hamiltonians/nearest_neighbor.py).

``to_mpo`` follows nearest_neighbor.py:105-165: each two-site term is split by
an SVD (on the device, truncated with the Strategy rule) into the Schmidt
channels of an MPO with bond dimension 2 + rank(h).  The small MPO tensors
are assembled on the host from the device SVD factors and uploaded once.
"""

from __future__ import annotations

import warnings
from abc import ABC, abstractmethod
from math import sqrt

import numpy as np
import scipy.sparse as sp

from .. import kernels as K
from ..core import DEFAULT_STRATEGY, Strategy
from ..operators.mpo import MPO

SX = np.array([[0.0, 1.0], [1.0, 0.0]])
SZ = np.array([[1.0, 0.0], [0.0, -1.0]])
SY = -1j * SZ @ SX


class NNHamiltonian(ABC):
    """H = sum_i h_{i,i+1} on a chain (nearest_neighbor.py:16-66)."""

    def __init__(self, size: int, constant: bool = False):
        self.size = size
        self.constant = constant

    @abstractmethod
    def dimension(self, i: int) -> int:
        pass

    def interaction_term(self, i: int, t: float = 0.0):
        d1 = self.dimension(i)
        d2 = self.dimension(i + 1)
        return sp.csr_matrix((d1 * d2, d1 * d2))

    def tomatrix(self, t: float = 0.0):
        warnings.warn("Method Hamiltonian.tomatrix() has been renamed to_matrix()",
                      DeprecationWarning, stacklevel=2)
        return self.to_matrix(t)

    def to_matrix(self, t: float = 0.0):
        """Sparse matrix (nearest_neighbor.py:77-103)."""
        dleft = 1
        H = sp.bsr_matrix((self.dimension(0), self.dimension(0)))
        for i in range(self.size - 1):
            H = sp.kron(H, sp.eye(self.dimension(i + 1)))
            H = H + sp.kron(sp.eye(dleft if dleft else 1), self.interaction_term(i, t))
            dleft *= self.dimension(i)
        return H.tocsr()

    def to_mpo(self, t: float = 0.0, strategy: Strategy = DEFAULT_STRATEGY) -> MPO:
        """MPO with bond 2 + rank(h_{i,i+1}) (nearest_neighbor.py:105-165)."""
        tensors = [np.zeros((2, di, di, 2)) for i in range(self.size)
                   for di in [self.dimension(i)]]
        for i in range(self.size - 1):
            Hij = np.asarray(sp.csr_matrix(self.interaction_term(i, t)).toarray())
            di = self.dimension(i)
            dj = self.dimension(i + 1)
            Hij = Hij.reshape(di, dj, di, dj).transpose(0, 2, 1, 3).reshape(di * di, dj * dj)
            U, s, V = K.svd(K.as_dev(Hij))
            ds, _ = K.truncate(s, strategy.method, strategy.tolerance,
                               strategy.max_bond_dimension)
            U, s, V = K.to_numpy(U)[:, :ds], s.cpu().numpy()[:ds], K.to_numpy(V)[:ds, :]
            s = np.sqrt(s)
            A = tensors[i]
            a, j, k, b = A.shape
            B = np.zeros((a, j, k, b + ds), dtype=np.result_type(U, A))
            B[:, :, :, :b] = A
            B[0, :, :, 2:] = (U * s).reshape(di, di, ds)
            B[1, :, :, 1] = np.eye(j)
            B[0, :, :, 0] = np.eye(j)
            tensors[i] = B
            A = tensors[i + 1]
            a, j, k, b = A.shape
            B = np.zeros((a + ds, j, k, b), dtype=np.result_type(V, A))
            B[:a, :, :, :] = A
            B[2:, :, :, 1] = (s.reshape(ds, 1) * V).reshape(ds, dj, dj)
            B[1, :, :, 1] = np.eye(j)
            B[0, :, :, 0] = np.eye(j)
            tensors[i + 1] = B
        tensors[0] = tensors[0][[0], :, :, :]
        tensors[-1] = tensors[-1][:, :, :, [1]]
        return MPO(tensors)


class ConstantNNHamiltonian(NNHamiltonian):
    """Nearest-neighbour Hamiltonian with constant terms (nearest_neighbor.py:168-280)."""

    def __init__(self, size: int, dimension):
        super().__init__(size, True)
        if isinstance(dimension, list):
            if len(dimension) != size:
                raise ValueError("len(dimension) must equal the Hamiltonian size")
            self.dimensions = dimension
        else:
            self.dimensions = [dimension] * size
        self.interactions = [np.zeros((si * sj, si * sj))
                             for si, sj in zip(self.dimensions[:-1], self.dimensions[1:])]

    def add_local_term(self, site: int, operator) -> "ConstantNNHamiltonian":
        if site < 0 or site >= self.size:
            raise IndexError("Site {site} out of bounds in add_local_term()")
        if site == 0:
            self.add_interaction_term(site, operator, np.eye(self.dimensions[1]))
        elif site == self.size - 1:
            self.add_interaction_term(site - 1, np.eye(self.dimensions[site - 1]), operator)
        else:
            half = 0.5 * sp.csr_matrix(operator)
            self.add_interaction_term(site - 1, np.eye(self.dimensions[site - 1]), half)
            self.add_interaction_term(site, half, np.eye(self.dimensions[site + 1]))
        return self

    def add_interaction_term(self, i: int, op1, op2=None) -> "ConstantNNHamiltonian":
        if i < 0 or i >= self.size - 1:
            raise IndexError("Site {site} out of bounds in add_interaction_term()")
        H12 = op1 if op2 is None else sp.kron(op1, op2)
        if (H12.ndim != 2 or H12.shape[0] != H12.shape[1]
                or H12.shape[1] != self.dimension(i) * self.dimension(i + 1)):
            raise ValueError("Invalid operators supplied to add_interaction_term()")
        self.interactions[i] = self.interactions[i] + H12
        return self

    def dimension(self, i: int) -> int:
        return self.dimensions[i]

    def interaction_term(self, i: int, t: float = 0.0):
        return self.interactions[i]


class ConstantTIHamiltonian(ConstantNNHamiltonian):
    """Translationally invariant chain (nearest_neighbor.py:283-315)."""

    def __init__(self, size: int, interaction=None, local_term=None):
        if local_term is not None:
            dimension = local_term.shape[0]
        elif interaction is not None:
            dimension = round(sqrt(interaction.shape[0]))
        else:
            raise ValueError("Either interactions or local term must be supplied")
        super().__init__(size, dimension)
        for site in range(size - 1):
            if interaction is not None:
                self.add_interaction_term(site, interaction)
            if local_term is not None:
                self.add_local_term(site, local_term)


class HeisenbergHamiltonian(ConstantTIHamiltonian):
    """0.25 sigma.sigma between neighbouring spins 1/2 (nearest_neighbor.py:318-334)."""

    def __init__(self, size: int, field=None):
        Hint = 0.25 * (sp.kron(SX, SX) + sp.kron(SY, SY) + sp.kron(SZ, SZ)).real
        Hlocal = None if field is None else field[0] * SX + field[1] * SY + field[2] * SZ
        super().__init__(size, interaction=Hint, local_term=Hlocal)


class XXZHamiltonian(ConstantTIHamiltonian):
    """Jxy (Sx Sx + Sy Sy) + Jz Sz Sz with S = sigma / 2 (open XXZ chain)."""

    def __init__(self, size: int, Jxy: float = 1.0, Jz: float = 0.5):
        Hint = (Jxy * 0.25 * (np.kron(SX, SX) + np.kron(SY, SY))
                + Jz * 0.25 * np.kron(SZ, SZ)).real
        super().__init__(size, interaction=Hint)
