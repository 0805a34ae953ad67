"""Nearest-neighbour Hamiltonians H = sum_i h_{i,i+1} and their MPOs.

Drop-in for ``seemps.hamiltonians`` (reference: src/seemps/hamiltonians/
nearest_neighbor.py:16-334): same classes, constructor arguments and methods.

MPO channel layout (the operator a user of the reference gets, kept as an
interface).  The bond between sites i and i+1 carries 2 + r_i channels, r_i
being the rank of h_{i,i+1} seen as a matrix between the operator spaces of
the two sites, h = sum_k A_k (x) B_k:

* channel 0: no term has been placed to the left of the bond yet;
* channel 1: one term has been placed (and completed) to the left;
* channel 2 + k: term k of the bond has been opened on site i by A_k and is
  closed on site i+1 by B_k.

A site tensor W[a, out, in, b] therefore holds the identity on (0 -> 0) and
(1 -> 1), A_k on (0 -> 2 + k) towards its right bond and B_k on (2 + k -> 1)
from its left bond.  The first site keeps row channel 0 only, the last site
column channel 1 only.  The split h = sum_k A_k (x) B_k is the Schmidt
decomposition of h (device SVD, truncated with the Strategy rule), with
sqrt(s_k) on each side.
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
SY = np.array([[0.0, -1j], [1j, 0.0]])


def _schmidt_terms(h, d1: int, d2: int, strategy: Strategy):
    """Operator Schmidt decomposition of the two-site term h (d1 d2 x d1 d2).

    Returns (A, B) with A[k] (d1 x d1) and B[k] (d2 x d2) such that
    h = sum_k kron(A[k], B[k]), the weights split as sqrt(s_k) on each side."""
    dense = h.toarray() if sp.issparse(h) else np.asarray(h)
    # h[(o1, o2), (i1, i2)] -> M[(o1, i1), (o2, i2)]
    M = dense.reshape(d1, d2, d1, d2).transpose(0, 2, 1, 3).reshape(d1 * d1, d2 * d2)
    U, s, VH = K.svd(K.as_dev(np.ascontiguousarray(M)))
    rank, _ = K.truncate(s, strategy.method, strategy.tolerance, strategy.max_bond_dimension)
    w = np.sqrt(s.cpu().numpy()[:rank])
    U = K.to_numpy(U)[:, :rank] * w
    VH = w[:, None] * K.to_numpy(VH)[:rank, :]
    return U.T.reshape(rank, d1, d1), VH.reshape(rank, d2, d2)


def _site_tensor(d: int, closing, opening, first: bool, last: bool) -> np.ndarray:
    """W[a, out, in, b] of one site in the channel layout of the module.

    ``closing`` (r_left x d x d or None) are the B_k of the bond on the left,
    ``opening`` (r_right x d x d or None) the A_k of the bond on the right."""
    r_left = 0 if closing is None else closing.shape[0]
    r_right = 0 if opening is None else opening.shape[0]
    dtype = np.result_type(np.float64, *(x.dtype for x in (closing, opening) if x is not None))
    W = np.zeros((2 + r_left, d, d, 2 + r_right), dtype=dtype)
    identity = np.eye(d)
    W[0, :, :, 0] = identity
    W[1, :, :, 1] = identity
    for k in range(r_right):
        W[0, :, :, 2 + k] = opening[k]
    for k in range(r_left):
        W[2 + k, :, :, 1] = closing[k]
    if first:
        W = W[:1]
    if last:
        W = W[:, :, :, 1:2]
    return np.ascontiguousarray(W)


class NNHamiltonian(ABC):
    """H = sum_i h_{i,i+1} on a chain of ``size`` subsystems
    (nearest_neighbor.py:16-66).  Subclasses give :meth:`dimension` and
    :meth:`interaction_term`."""

    def __init__(self, size: int, constant: bool = False):
        self.size = size
        self.constant = constant

    @abstractmethod
    def dimension(self, i: int) -> int:
        """Physical dimension of subsystem ``i``."""

    def interaction_term(self, i: int, t: float = 0.0):
        """h_{i,i+1} at time ``t``; zero unless a subclass says otherwise."""
        n = self.dimension(i) * self.dimension(i + 1)
        return sp.csr_matrix((n, n))

    def tomatrix(self, t: float = 0.0):
        warnings.warn("Method Hamiltonian.tomatrix() has been renamed to_matrix()",
                      DeprecationWarning, stacklevel=2)
        return self.to_matrix(t)

    def to_matrix(self, t: float = 0.0):
        """Sparse matrix sum_i 1_{<i} (x) h_{i,i+1} (x) 1_{>i+1}."""
        dims = [self.dimension(i) for i in range(self.size)]
        total = int(np.prod(dims))
        H = sp.csr_matrix((total, total))
        for i in range(self.size - 1):
            left = int(np.prod(dims[:i]))
            right = int(np.prod(dims[i + 2:]))
            term = sp.csr_matrix(self.interaction_term(i, t))
            H = H + sp.kron(sp.kron(sp.identity(left), term), sp.identity(right))
        return sp.csr_matrix(H)

    def to_mpo(self, t: float = 0.0, strategy: Strategy = DEFAULT_STRATEGY) -> MPO:
        """MPO in the channel layout of the module, bond 2 + rank(h_{i,i+1})
        (nearest_neighbor.py:105-165)."""
        N = self.size
        terms = [_schmidt_terms(self.interaction_term(i, t), self.dimension(i),
                                self.dimension(i + 1), strategy) for i in range(N - 1)]
        tensors = []
        for i in range(N):
            closing = terms[i - 1][1] if i > 0 else None
            opening = terms[i][0] if i < N - 1 else None
            tensors.append(_site_tensor(self.dimension(i), closing, opening, i == 0,
                                        i == N - 1))
        return MPO(tensors)


class ConstantNNHamiltonian(NNHamiltonian):
    """Nearest-neighbour Hamiltonian with time-independent terms
    (nearest_neighbor.py:168-280).  ``dimension`` is one int for every site or
    a list of ``size`` ints."""

    def __init__(self, size: int, dimension):
        super().__init__(size, True)
        if isinstance(dimension, list):
            if len(dimension) != size:
                raise ValueError("len(dimension) must equal the Hamiltonian size")
            self.dimensions = list(dimension)
        else:
            self.dimensions = [dimension] * size
        self.interactions = [np.zeros((self.dimensions[i] * self.dimensions[i + 1],) * 2)
                             for i in range(size - 1)]

    def add_local_term(self, site: int, operator) -> "ConstantNNHamiltonian":
        """Add ``operator`` acting on ``site``.  It is shared evenly between the
        bonds that touch the site (one bond at either end of the chain, two in
        the bulk)."""
        bonds = [b for b in (site - 1, site) if 0 <= b < self.size - 1]
        if not (0 <= site < self.size) or not bonds:
            raise IndexError(f"Site {site} out of bounds in add_local_term()")
        weight = 1.0 / len(bonds)
        part = weight * sp.csr_matrix(operator) if weight != 1.0 else operator
        for b in bonds:
            if b == site:   # the site is the left member of the bond
                self.add_interaction_term(b, part, np.eye(self.dimensions[b + 1]))
            else:           # the right member
                self.add_interaction_term(b, np.eye(self.dimensions[b]), part)
        return self

    def add_interaction_term(self, i: int, op1, op2=None) -> "ConstantNNHamiltonian":
        """Add ``op1`` (a two-site operator) or ``op1 (x) op2`` to h_{i,i+1}."""
        if not (0 <= i < self.size - 1):
            raise IndexError(f"Site {i} out of bounds in add_interaction_term()")
        H12 = op1 if op2 is None else sp.kron(op1, op2)
        n = self.dimension(i) * self.dimension(i + 1)
        if H12.ndim != 2 or H12.shape != (n, n):
            raise ValueError("Invalid operators supplied to add_interaction_term()")
        self.interactions[i] = self.interactions[i] + H12
        return self

    def dimension(self, i: int) -> int:
        return self.dimensions[i]

    def interaction_term(self, i: int, t: float = 0.0):
        return self.interactions[i]


class ConstantTIHamiltonian(ConstantNNHamiltonian):
    """Translationally invariant chain (nearest_neighbor.py:283-315): the same
    ``interaction`` on every bond and ``local_term`` on sites 0 .. size-2 (the
    reference's loop over bonds leaves the last site without it)."""

    def __init__(self, size: int, interaction=None, local_term=None):
        if local_term is not None:
            dimension = local_term.shape[0]
        elif interaction is not None:
            dimension = round(sqrt(interaction.shape[0]))
        else:
            raise ValueError("Either interactions or local term must be supplied")
        super().__init__(size, dimension)
        for bond in range(size - 1):
            if interaction is not None:
                self.add_interaction_term(bond, interaction)
            if local_term is not None:
                self.add_local_term(bond, local_term)


class HeisenbergHamiltonian(ConstantTIHamiltonian):
    """(1/4) sigma . sigma between neighbouring spins 1/2, plus an optional
    field (hx, hy, hz) . sigma (nearest_neighbor.py:318-334)."""

    def __init__(self, size: int, field=None):
        exchange = 0.25 * sum(sp.kron(P, P) for P in (SX, SY, SZ)).real
        local = None if field is None else sum(f * P for f, P in zip(field, (SX, SY, SZ)))
        super().__init__(size, interaction=exchange, local_term=local)


class XXZHamiltonian(ConstantTIHamiltonian):
    """Jxy (Sx Sx + Sy Sy) + Jz Sz Sz with S = sigma / 2 (open XXZ chain)."""

    def __init__(self, size: int, Jxy: float = 1.0, Jz: float = 0.5):
        coupling = (0.25 * Jxy * (np.kron(SX, SX) + np.kron(SY, SY))
                    + 0.25 * Jz * np.kron(SZ, SZ)).real
        super().__init__(size, interaction=coupling)
