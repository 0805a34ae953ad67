"""Oracle: Strategy, truncation rules, SVD split, canonical form, environments.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/seemps/cython/truncation.pxi``, ``svd.pxi``, ``schmidt.pxi``,
``contractions.pxi`` and ``environments.pxi`` (mirrored in
``src/seemps/cython/_python_core.py``) in NumPy.

Pinning (tests/test_oracle_core.py):
* ``truncate``            -- the reference's own known-answer tests
                             (tests/golden/truncation_known_answers.json) plus a
                             brute-force restatement of the rule on random spectra.
* ``svd``                 -- reconstruction, unitarity and sigma vs LAPACK.
* ``left/right_orth_2site``, ``canonicalize`` -- isometry + state preservation
                             against the dense vector.
* environments / ``scprod`` -- dense ``<bra|ket>`` of the state vectors.
* ``contract_nrjl_ijk_klm`` -- dense two-site gate applied to the vector.
Gauge-dependent raw tensors are not pinned (the reference's own two backends
disagree on singular-vector phases); comparisons are gauge invariant.
"""

from __future__ import annotations

import numpy as np

MAX_BOND_DIMENSION = 0x7FFFFFFF
DEFAULT_TOLERANCE = float(np.finfo(np.float64).eps)


class Truncation:
    """Codes of ``truncation.pxi:1-8``."""

    DO_NOT_TRUNCATE = 0
    RELATIVE_SINGULAR_VALUE = 1
    RELATIVE_NORM_SQUARED_ERROR = 2
    ABSOLUTE_SINGULAR_VALUE = 3


class Simplification:
    """Codes of ``truncation.pxi:10-15``."""

    DO_NOT_SIMPLIFY = 0
    CANONICAL_FORM = 1
    VARIATIONAL = 2
    VARIATIONAL_EXACT_GUESS = 3


class Strategy:
    """Plain record of the reference ``Strategy`` (``truncation.pxi:66-147``)."""

    def __init__(
        self,
        method=Truncation.RELATIVE_NORM_SQUARED_ERROR,
        tolerance=DEFAULT_TOLERANCE,
        simplification_tolerance=DEFAULT_TOLERANCE,
        max_bond_dimension=MAX_BOND_DIMENSION,
        normalize=False,
        simplify=Simplification.VARIATIONAL,
        max_sweeps=16,
    ):
        # truncation.pxi:103-131
        if tolerance < 0 or tolerance >= 1.0:
            raise AssertionError("Invalid tolerance")
        if tolerance == 0 and method != Truncation.DO_NOT_TRUNCATE:
            method = Truncation.ABSOLUTE_SINGULAR_VALUE
        if max_bond_dimension <= 0 or max_bond_dimension > MAX_BOND_DIMENSION:
            raise AssertionError("Invalid bond dimension")
        if simplify < 0 or simplify > 3:
            raise AssertionError("Invalid simplify")
        if max_sweeps < 0:
            raise AssertionError("Negative sweeps")
        if method not in (0, 1, 2, 3):
            raise AssertionError("Invalid method")
        self.method = int(method)
        self.tolerance = float(tolerance)
        self.simplification_tolerance = float(simplification_tolerance)
        self.max_bond_dimension = int(max_bond_dimension)
        self.normalize = bool(normalize)
        self.simplify = int(simplify)
        self.max_sweeps = int(max_sweeps)

    def replace(self, **kw):
        args = dict(
            method=self.method,
            tolerance=self.tolerance,
            simplification_tolerance=self.simplification_tolerance,
            max_bond_dimension=self.max_bond_dimension,
            normalize=self.normalize,
            simplify=self.simplify,
            max_sweeps=self.max_sweeps,
        )
        args.update({k: v for k, v in kw.items() if v is not None})
        return Strategy(**args)


DEFAULT_STRATEGY = Strategy()
NO_TRUNCATION = DEFAULT_STRATEGY.replace(
    method=Truncation.DO_NOT_TRUNCATE, simplify=Simplification.DO_NOT_SIMPLIFY
)


def truncate(s: np.ndarray, strategy: Strategy) -> tuple[int, float]:
    """Kept size ``D`` and squared error for descending ``s``.

    Follows the C loops of ``truncation.pxi:251-327`` literally, including the
    summation order of the error table (smallest singular value first) so that
    threshold comparisons round exactly like the reference.
    """
    N = s.shape[0]
    method = strategy.method
    if method == Truncation.DO_NOT_TRUNCATE:  # truncation.pxi:222-223
        return N, 0.0
    max_bond = strategy.max_bond_dimension
    if method == Truncation.RELATIVE_NORM_SQUARED_ERROR:  # :251-287
        errors = np.empty(N + 1)
        total = 0.0
        for i in range(N):
            errors[i] = total
            x = float(s[N - 1 - i])
            total += x * x
        errors[N] = total
        max_error = total * strategy.tolerance
        i_stop = N - 1
        for i in range(N):
            if errors[i] > max_error:
                i_stop = i - 1
                break
        D = min(N - i_stop, max_bond)
        return D, float(errors[N - D])
    if method == Truncation.RELATIVE_SINGULAR_VALUE:  # :289-307
        threshold = strategy.tolerance * float(s[0])
    elif method == Truncation.ABSOLUTE_SINGULAR_VALUE:  # :309-327
        threshold = strategy.tolerance
    else:
        raise AssertionError("Invalid truncation method")
    D = min(N, max_bond)
    for i in range(1, D):
        if s[i] <= threshold:
            D = i
            break
    err = 0.0
    for i in range(D, N):
        err += float(s[i]) * float(s[i])
    return D, err


def svd(A: np.ndarray):
    """Economy SVD with the reference's dtype promotion (``svd.pxi:74-77``)."""
    if A.dtype == np.complex64:
        A = A.astype(np.complex128)
    elif A.dtype not in (np.float64, np.complex128):
        A = A.astype(np.float64)
    return np.linalg.svd(A, full_matrices=False)


def left_orth_2site(AA: np.ndarray, strategy: Strategy):
    """``schmidt.pxi:123-147``: AA[a,i,j,b] -> (U[a,i,D], sV[D,j,b], sqrt(err))."""
    a, d1, d2, b = AA.shape
    U, s, Vh = svd(AA.reshape(a * d1, d2 * b))
    D, err = truncate(s, strategy)
    return (
        U[:, :D].reshape(a, d1, D),
        (s[:D, None] * Vh[:D, :]).reshape(D, d2, b),
        np.sqrt(err),
    )


def right_orth_2site(AA: np.ndarray, strategy: Strategy):
    """``schmidt.pxi:150-174``: AA[a,i,j,b] -> (Us[a,i,D], V[D,j,b], sqrt(err))."""
    a, d1, d2, b = AA.shape
    U, s, Vh = svd(AA.reshape(a * d1, d2 * b))
    D, err = truncate(s, strategy)
    return (
        (U[:, :D] * s[:D]).reshape(a, d1, D),
        Vh[:D, :].reshape(D, d2, b),
        np.sqrt(err),
    )


def contract_last_and_first(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """``contractions.pxi:6-38``: tensordot over A's last and B's first index."""
    return np.tensordot(A, B, axes=1)


def contract_nrjl_ijk_klm(U: np.ndarray, A: np.ndarray, B: np.ndarray):
    """``contractions.pxi:102-132``: ``einsum('ijk,klm,nrjl->inrm')``, U as [nr, jl]."""
    a, d, b = A.shape
    _, e, c = B.shape
    U4 = U.reshape(d, e, d, e)
    return np.einsum("ijk,klm,nrjl->inrm", A, B, U4, optimize=True)


def update_in_canonical_form_right(state: list, A: np.ndarray, site: int, strategy):
    """``schmidt.pxi:18-56``: move the centre right with a truncated SVD."""
    if site + 1 == len(state):
        state[site] = A
        return site, 0.0
    a, i, b = A.shape
    U, s, Vh = svd(A.reshape(a * i, b).copy())
    D, err = truncate(s, strategy)
    state[site] = U[:, :D].reshape(a, i, D)
    state[site + 1] = contract_last_and_first(s[:D, None] * Vh[:D, :], state[site + 1])
    return site + 1, float(np.sqrt(err))


def update_in_canonical_form_left(state: list, A: np.ndarray, site: int, strategy):
    """``schmidt.pxi:59-93``: move the centre left with a truncated SVD."""
    if site == 0:
        state[0] = A
        return 0, 0.0
    a, i, b = A.shape
    U, s, Vh = svd(A.reshape(a, i * b).copy())
    D, err = truncate(s, strategy)
    state[site] = Vh[:D, :].reshape(D, i, b)
    state[site - 1] = contract_last_and_first(state[site - 1], U[:, :D] * s[:D])
    return site - 1, float(np.sqrt(err))


def canonicalize(state: list, center: int, strategy) -> float:
    """``schmidt.pxi:109-120``: right-orthogonalise [0,center), then left (center,L)."""
    err = 0.0
    for i in range(center):
        err += update_in_canonical_form_right(state, state[i], i, strategy)[1]
    for i in range(len(state) - 1, center, -1):
        err += update_in_canonical_form_left(state, state[i], i, strategy)[1]
    return err


def recanonicalize(state: list, oldcenter: int, newcenter: int, strategy) -> float:
    """``schmidt.pxi:96-106``."""
    err = 0.0
    while oldcenter > newcenter:
        err += update_in_canonical_form_left(state, state[oldcenter], oldcenter, strategy)[1]
        oldcenter -= 1
    while oldcenter < newcenter:
        err += update_in_canonical_form_right(state, state[oldcenter], oldcenter, strategy)[1]
        oldcenter += 1
    return err


def update_left_environment(B: np.ndarray, A: np.ndarray, rho: np.ndarray):
    """``environments.pxi:10-32``: rho'[n,k] = sum B*[l,j,n] rho[l,i] A[i,j,k]."""
    return np.einsum("ljn,li,ijk->nk", B.conj(), rho, A, optimize=True)


def update_right_environment(B: np.ndarray, A: np.ndarray, rho: np.ndarray):
    """``environments.pxi:39-62``: rho'[i,l] = sum A[i,j,k] rho[k,n] B*[l,j,n]."""
    return np.einsum("ijk,kn,ljn->il", A, rho, B.conj(), optimize=True)


def join_environments(rhoL: np.ndarray, rhoR: np.ndarray):
    """``environments.pxi:79-85``: sum rhoL[a,b] rhoR[b,a]."""
    if rhoL.shape[0] == 1:
        return rhoL.flat[0] * rhoR.flat[0]
    return np.dot(rhoL.ravel(), rhoR.T.ravel())


def scprod(bra: list, ket: list):
    """``environments.pxi:87-100``: fold left environments over all sites."""
    if len(bra) != len(ket):
        raise ValueError("Invalid arguments to scprod")
    rho = np.eye(1)
    for B, A in zip(bra, ket):
        rho = update_left_environment(B, A, rho)
    return rho.flat[0]
