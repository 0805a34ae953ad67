"""Drop-in counterpart of the ``seemps.cython`` kernel facade.

Mirrors the configuration types (``Strategy``, ``Truncation``,
``Simplification``, ``GemmOrder``, constants) of
``src/seemps/cython/truncation.pxi:1-220`` and the 17 kernels re-exported by
``src/seemps/cython/__init__.py:26-54``.  Tensors are torch complex128 device
tensors; every numerical step runs in ``libseemps_hip.so``.  NumPy arrays
passed in are uploaded first.  Also importable as ``seemps2_amd.cython``.
"""

from __future__ import annotations

import ctypes
import math
from typing import Sequence

import numpy as np
import torch

from . import _lib
from . import kernels as K
from ._lib import OP_C, OP_H, OP_N, OP_T

MAX_BOND_DIMENSION = 0x7FFFFFFF
"""Maximum bond dimension for any MPS."""

DEFAULT_TOLERANCE = float(np.finfo(np.float64).eps)
"""Relative or absolute tolerance in various algorithms."""


class Truncation:
    """SVD truncation algorithm when splitting tensors (truncation.pxi:17-37)."""

    DO_NOT_TRUNCATE = 0
    RELATIVE_SINGULAR_VALUE = 1
    RELATIVE_NORM_SQUARED_ERROR = 2
    ABSOLUTE_SINGULAR_VALUE = 3


class Simplification:
    """Tensor network simplification algorithms (truncation.pxi:39-55)."""

    DO_NOT_SIMPLIFY = 0
    CANONICAL_FORM = 1
    VARIATIONAL = 2
    VARIATIONAL_EXACT_GUESS = 3


class GemmOrder:
    """Matrix transpose/adjoint flags for ``_gemm`` (gemm.pxi)."""

    NORMAL = 0
    TRANSPOSE = 1
    ADJOINT = 2


class Strategy:
    """MPS and MPO simplification strategies (truncation.pxi:66-196).

    Same arguments, defaults, validation (``AssertionError``) and the rule that a
    zero tolerance with a truncating method selects ``ABSOLUTE_SINGULAR_VALUE``.
    """

    __slots__ = (
        "method",
        "tolerance",
        "simplification_tolerance",
        "max_bond_dimension",
        "normalize",
        "simplify",
        "max_sweeps",
    )

    def __init__(
        self,
        method: int = Truncation.RELATIVE_NORM_SQUARED_ERROR,
        tolerance: float = DEFAULT_TOLERANCE,
        simplification_tolerance: float = DEFAULT_TOLERANCE,
        max_bond_dimension: int = MAX_BOND_DIMENSION,
        normalize: bool = False,
        simplify: int = Simplification.VARIATIONAL,
        max_sweeps: int = 16,
    ):
        if tolerance < 0 or tolerance >= 1.0:
            raise AssertionError("Invalid tolerance argument passed to Strategy")
        if tolerance == 0 and method != Truncation.DO_NOT_TRUNCATE:
            method = Truncation.ABSOLUTE_SINGULAR_VALUE
        if max_bond_dimension <= 0 or max_bond_dimension > MAX_BOND_DIMENSION:
            raise AssertionError("Invalid bond dimension in Strategy")
        if simplify < 0 or simplify > Simplification.VARIATIONAL_EXACT_GUESS:
            raise AssertionError("Invalid simplify argument passed to Strategy")
        if max_sweeps < 0:
            raise AssertionError("Negative or zero number of sweeps in Strategy")
        if method not in (0, 1, 2, 3):
            raise AssertionError("Invalid method argument passed to Strategy")
        self.method = int(method)
        self.tolerance = float(tolerance)
        self.simplification_tolerance = float(simplification_tolerance)
        self.max_bond_dimension = int(max_bond_dimension)
        self.normalize = bool(normalize)
        self.simplify = int(simplify)
        self.max_sweeps = int(max_sweeps)

    def replace(self, method=None, tolerance=None, simplification_tolerance=None,
                max_bond_dimension=None, normalize=None, simplify=None, max_sweeps=None):
        return Strategy(
            method=self.method if method is None else method,
            tolerance=self.tolerance if tolerance is None else tolerance,
            simplification_tolerance=(self.simplification_tolerance
                                      if simplification_tolerance is None
                                      else simplification_tolerance),
            max_bond_dimension=(self.max_bond_dimension if max_bond_dimension is None
                                else max_bond_dimension),
            normalize=self.normalize if normalize is None else normalize,
            simplify=self.simplify if simplify is None else simplify,
            max_sweeps=self.max_sweeps if max_sweeps is None else max_sweeps,
        )

    def set_normalization(self, normalize: bool) -> "Strategy":
        return self.replace(normalize=normalize)

    def get_method(self) -> int:
        return self.method

    def get_simplification_method(self) -> int:
        return self.simplify

    def get_tolerance(self) -> float:
        return self.tolerance

    def get_simplification_tolerance(self) -> float:
        return self.simplification_tolerance

    def get_max_bond_dimension(self) -> int:
        return self.max_bond_dimension

    def get_normalize_flag(self) -> bool:
        return self.normalize

    def get_max_sweeps(self) -> int:
        return self.max_sweeps

    def get_simplify_flag(self) -> bool:
        return self.simplify != Simplification.DO_NOT_SIMPLIFY

    def __str__(self) -> str:
        m = {0: "None", 1: "RelativeSVD", 2: "RelativeNorm", 3: "AbsoluteSVD"}[self.method]
        s = {0: "None", 1: "CanonicalForm", 2: "Variational",
             3: "Variational (exact guess)"}[self.simplify]
        return (f"Strategy(method={m}, tolerance={self.tolerance:5g}, "
                f"max_bond_dimension={self.max_bond_dimension}, normalize={self.normalize}, "
                f"simplify={s}, simplification_tolerance={self.simplification_tolerance:5g}, "
                f"max_sweeps={self.max_sweeps})")

    __repr__ = __str__


DEFAULT_STRATEGY = Strategy(
    method=Truncation.RELATIVE_NORM_SQUARED_ERROR,
    simplify=Simplification.VARIATIONAL,
    tolerance=DEFAULT_TOLERANCE,
    simplification_tolerance=DEFAULT_TOLERANCE,
    max_bond_dimension=MAX_BOND_DIMENSION,
    normalize=False,
)

NO_TRUNCATION = DEFAULT_STRATEGY.replace(
    method=Truncation.DO_NOT_TRUNCATE, simplify=Simplification.DO_NOT_SIMPLIFY
)

# --------------------------------------------------------------- contractions


def _contract_last_and_first(A, B) -> torch.Tensor:
    """tensordot(A, B, 1) (contractions.pxi:6-38)."""
    if not isinstance(A, torch.Tensor) or not isinstance(B, torch.Tensor):
        if not hasattr(A, "shape") or not hasattr(B, "shape"):
            raise ValueError("_contract_last_and_first expects tensors")
        A, B = K.as_dev(A), K.as_dev(B)
    k = A.shape[-1]
    if B.shape[0] != k:
        raise ValueError("_contract_last_and_first: dimension mismatch")
    out = K.gemm(A.reshape(-1, k), B.reshape(k, -1))
    return out.view(*A.shape[:-1], *B.shape[1:])


def _contract_nrjl_ijk_klm(U, A, B) -> torch.Tensor:
    """einsum('ijk,klm,nrjl->inrm', A, B, U) with U as [n*r, j*l] (contractions.pxi:102-132)."""
    A, B, U = K.as_dev(A), K.as_dev(B), K.as_dev(U)
    if A.dim() != 3 or B.dim() != 3 or U.dim() != 2:
        raise ValueError("Invalid arguments to _contract_nrjl_ijk_klm")
    a, d, b = A.shape
    _, e, c = B.shape
    AB = K.gemm(A.reshape(a * d, b), B.reshape(b, e * c))
    Ut = K.permute(U, (1, 0))  # Mop[(j,l), (n,r)]
    out = K.local(AB, Ut, a, d * e, c, d * e, 1)
    return out.view(a, d, e, c)


# --------------------------------------------------------------- environments


def _begin_environment(D: int = 1) -> torch.Tensor:
    """Identity environment (environments.pxi:1-8); a constant uploaded once."""
    return K.as_dev(np.eye(D))


def _update_left_environment(B, A, rho) -> torch.Tensor:
    """rho'[n,k] = sum B*[l,j,n] rho[l,i] A[i,j,k] (environments.pxi:10-32)."""
    B, A, rho = K.as_dev(B), K.as_dev(A), K.as_dev(rho)
    if B.dim() != 3 or A.dim() != 3 or rho.dim() != 2:
        raise ValueError("Invalid arguments to _update_left_environment")
    l, j, n = B.shape
    i, _, k = A.shape
    T = K.gemm(rho, A.reshape(i, j * k))                       # [l, (j,k)]
    return K.gemm(B.reshape(l * j, n), T.view(l * j, k), opA=OP_H)  # [n, k]


def _update_right_environment(B, A, rho) -> torch.Tensor:
    """rho'[i,l] = sum A[i,j,k] rho[k,n] B*[l,j,n] (environments.pxi:39-62)."""
    B, A, rho = K.as_dev(B), K.as_dev(A), K.as_dev(rho)
    if B.dim() != 3 or A.dim() != 3 or rho.dim() != 2:
        raise ValueError("Invalid arguments to _update_right_environment")
    i, j, k = A.shape
    l, _, n = B.shape
    T = K.gemm(A.reshape(i * j, k), rho)                       # [(i,j), n]
    return K.gemm(T.view(i, j * n), B.reshape(l, j * n), opB=OP_H)  # [i, l]


def _end_environment(rho) -> complex:
    """Scalar product from the last environment (environments.pxi:64-66)."""
    if isinstance(rho, torch.Tensor):
        return complex(rho.reshape(-1)[0].item())
    return complex(np.asarray(rho).reshape(-1)[0])


def _join_environments(rhoL, rhoR) -> complex:
    """sum rhoL[a,b] rhoR[b,a] (environments.pxi:79-85)."""
    rhoL, rhoR = K.as_dev(rhoL), K.as_dev(rhoR)
    if rhoL.dim() != 2 or rhoR.dim() != 2 or rhoL.shape[::-1] != rhoR.shape:
        raise ValueError("Invalid arguments to _join_environments")
    if rhoL.shape[0] == 1:
        return _end_environment(rhoL) * _end_environment(rhoR)
    return K.dotc(K.conj(rhoL), K.permute(rhoR, (1, 0)))


def scprod(bra, ket) -> complex:
    """<bra|ket> by folding left environments (environments.pxi:87-100)."""
    A = bra._data
    B = ket._data
    if len(A) != len(B):
        raise ValueError("Invalid arguments to scprod")
    rho = None
    for Ai, Bi in zip(A, B):
        if rho is None:
            # first site: rho = 1
            l, j, n = Ai.shape
            i, _, k = Bi.shape
            rho = K.gemm(Ai.reshape(l * j, n), Bi.reshape(i * j, k), opA=OP_H)
        else:
            rho = _update_left_environment(Ai, Bi, rho)
    return _end_environment(rho)


def vdot(bra, ket) -> complex:
    """Alias of :func:`scprod`."""
    return scprod(bra, ket)


def _gemm(B, BT: int, A, AT: int) -> torch.Tensor:
    """op(B, BT) @ op(A, AT) (gemm.pxi:16-122)."""
    B, A = K.as_dev(B), K.as_dev(A)
    if A.dim() != 2 or B.dim() != 2:
        raise ValueError()
    if BT not in (0, 1, 2) or AT not in (0, 1, 2):
        raise ValueError(f"Invalid GEMM order flag {BT}, {AT}")
    return K.gemm(B, A, opA=BT, opB=AT)


# --------------------------------------------------------------------- SVD

_svd_driver = "gesdd"


def _select_svd_driver(name: str) -> None:
    """Accepts the reference's driver names (svd.pxi:25-32); the device SVD is
    always the one-sided Jacobi of libseemps_hip.so."""
    global _svd_driver
    if name not in ("gesvd", "gesdd"):
        raise ValueError(f"Invalid LAPACK SVD driver name: {name}")
    _svd_driver = name


def _destructive_svd(A):
    """Economy SVD (U, s, Vh) in NumPy order (svd.pxi:17-83)."""
    A = K.as_dev(A)
    if A.dim() != 2:
        raise ValueError("Invalid argument to SVD")
    return K.svd(A)


def _shrink_ndarray(s: np.ndarray, D: int) -> None:
    """Shorten the 1-D array ``s`` to its first ``D`` entries in place.

    The reference sets the array's first dimension in place
    (truncation.pxi:284-286), which also works for views that do not own their
    data.  An owning array is resized by NumPy; for a view the dimension field
    of the array object is rewritten the same way, after checking that the
    object layout is the one expected (data pointer, ndim and length agree)."""
    if s.flags.owndata and s.base is None:
        s.resize(D, refcheck=False)
        return
    addr = id(s)
    data = ctypes.c_void_p.from_address(addr + 16).value
    nd = ctypes.c_int.from_address(addr + 24).value
    dims = ctypes.c_void_p.from_address(addr + 32).value
    if data != s.ctypes.data or nd != 1 or not dims or \
            ctypes.c_ssize_t.from_address(dims).value != s.shape[0]:
        raise TypeError("destructively_truncate_vector: cannot shrink this array in place")
    ctypes.c_ssize_t.from_address(dims).value = D


def destructively_truncate_vector(s, strategy: Strategy) -> float:
    """Truncate ``s`` in place, return the squared error (truncation.pxi:329-334)."""
    if isinstance(s, np.ndarray):
        if s.ndim != 1:
            raise TypeError("destructively_truncate_vector requires a 1D float64 ndarray")
        if s.size == 0:
            return 0.0
        D, err = K.truncate(torch.as_tensor(s, dtype=torch.float64).to(_lib.device()),
                            strategy.method, strategy.tolerance, strategy.max_bond_dimension)
        if D != s.shape[0]:
            _shrink_ndarray(s, D)
        return err
    if s.dim() != 1:
        raise TypeError("destructively_truncate_vector requires a 1D float64 tensor")
    D, err = K.truncate(s.to(torch.float64).contiguous(), strategy.method, strategy.tolerance,
                        strategy.max_bond_dimension)
    if D != s.shape[0]:
        s.resize_(D)
    return err


def _split(AA, strategy: Strategy, mode: int, basis=None):
    from .learned import two_site_split

    AA = K.as_dev(AA)
    a, d1, d2, b = AA.shape
    A2, B2, D, err2 = two_site_split(AA, a * d1, d2 * b, mode, strategy, basis)
    return A2.view(a, d1, D), B2.view(D, d2, b), math.sqrt(err2)


def _left_orth_2site(AA, strategy: Strategy, basis=None):
    """AA[a,i,j,b] -> (U[a,i,D] isometry, sV[D,j,b], sqrt(err)) (schmidt.pxi:123-147).

    ``basis`` (optional): the current right neighbour B[k,j,b] when it is
    right-canonical; only the learned-rank split uses it, as its sketch."""
    return _split(AA, strategy, 0, basis)


def _right_orth_2site(AA, strategy: Strategy, basis=None):
    """AA[a,i,j,b] -> (Us[a,i,D], V[D,j,b] isometry, sqrt(err)) (schmidt.pxi:150-174).

    ``basis`` (optional): the current left neighbour A[a,i,k] when it is
    left-canonical; only the learned-rank split uses it, as its sketch."""
    return _split(AA, strategy, 1, basis)


def _move_right(state: list, A, site: int, strategy: Strategy) -> float:
    """__update_in_canonical_form_right (schmidt.pxi:18-42)."""
    A = K.as_dev(A)
    a, i, b = A.shape
    U, sV, D, err2, _ = K.svd_split(A, a * i, b, 0, strategy.method, strategy.tolerance,
                                    strategy.max_bond_dimension)
    state[site] = U.view(a, i, D)
    nxt = state[site + 1]
    state[site + 1] = K.gemm(sV, nxt.reshape(nxt.shape[0], -1)).view(D, *nxt.shape[1:])
    return math.sqrt(err2)


def _move_left(state: list, A, site: int, strategy: Strategy) -> float:
    """__update_in_canonical_form_left (schmidt.pxi:59-82)."""
    A = K.as_dev(A)
    a, i, b = A.shape
    Us, V, D, err2, _ = K.svd_split(A, a, i * b, 1, strategy.method, strategy.tolerance,
                                    strategy.max_bond_dimension)
    state[site] = V.view(D, i, b)
    prv = state[site - 1]
    state[site - 1] = K.gemm(prv.reshape(-1, prv.shape[-1]), Us).view(*prv.shape[:-1], D)
    return math.sqrt(err2)


def _update_in_canonical_form_right(state: list, A, site: int, truncation: Strategy):
    """schmidt.pxi:45-56."""
    if site + 1 == len(state):
        state[site] = K.as_dev(A)
        return site, 0.0
    return site + 1, _move_right(state, A, site, truncation)


def _update_in_canonical_form_left(state: list, A, site: int, truncation: Strategy):
    """schmidt.pxi:84-93."""
    if site == 0:
        state[0] = K.as_dev(A)
        return 0, 0.0
    return site - 1, _move_left(state, A, site, truncation)


def _recanonicalize(state: list, oldcenter: int, newcenter: int, truncation: Strategy) -> float:
    """schmidt.pxi:96-106."""
    err = 0.0
    while oldcenter > newcenter:
        err += _move_left(state, state[oldcenter], oldcenter, truncation)
        oldcenter -= 1
    while oldcenter < newcenter:
        err += _move_right(state, state[oldcenter], oldcenter, truncation)
        oldcenter += 1
    return err


def _canonicalize(state: list, center: int, truncation: Strategy) -> float:
    """schmidt.pxi:109-120."""
    err = 0.0
    for i in range(0, center):
        err += _move_right(state, state[i], i, truncation)
    for i in range(len(state) - 1, center, -1):
        err += _move_left(state, state[i], i, truncation)
    return err


__all__ = [
    "DEFAULT_STRATEGY",
    "DEFAULT_TOLERANCE",
    "GemmOrder",
    "MAX_BOND_DIMENSION",
    "NO_TRUNCATION",
    "Simplification",
    "Strategy",
    "Truncation",
    "_begin_environment",
    "_canonicalize",
    "_contract_last_and_first",
    "_contract_nrjl_ijk_klm",
    "_destructive_svd",
    "_end_environment",
    "_gemm",
    "_join_environments",
    "_left_orth_2site",
    "_right_orth_2site",
    "_recanonicalize",
    "_select_svd_driver",
    "_update_in_canonical_form_right",
    "_update_in_canonical_form_left",
    "_update_left_environment",
    "_update_right_environment",
    "destructively_truncate_vector",
    "scprod",
    "vdot",
]
