"""A x = b by two-site DMRG sweeps (reference: src/seemps/solve/dmrg.py:28-188).

Each bond solves the local system H_eff(i, i+1) x = b_local, with b_local the
two-site tensor of the AntilinearForm <x|b> and the current state as the
start, by a Krylov solver on device vectors (``krylov.py``), then splits with
the Strategy.  Convergence: the full residual ||A x - b|| after each sweep, or
the relative change of x between sweeps.
"""

from __future__ import annotations

from math import sqrt

import numpy as np

from .. import kernels as K
from ..core import DEFAULT_STRATEGY, Strategy, _contract_last_and_first, scprod
from ..operators import MPO
from ..operators.quadratic import QuadraticForm
from ..state import CanonicalMPS, MPS
from ..state.antilinear import AntilinearForm
from .krylov import SOLVERS


def _relative_change(a: MPS, b: MPS) -> float:
    a2, b2 = a.norm_squared(), b.norm_squared()
    if a2 == 0:
        return 0.0 if b2 == 0 else np.inf
    d2 = a2 - 2.0 * complex(scprod(a, b)).real + b2
    return sqrt(max(d2, 0.0) / a2)


def _solve_local(QF: QuadraticForm, i: int, b_tensor, atol: float, rtol: float, solver):
    """The two-site system at bond i (dmrg.py:38-56), early exit when the
    current state already solves it."""
    v = _contract_last_and_first(QF.state[i], QF.state[i + 1])
    b_tensor = K.as_dev(b_tensor)

    def apply(x):
        return QF.two_site_operator_matvec(i, x)

    r = K.copy(b_tensor)
    K.axpby(-1.0, apply(v), 1.0, r)
    nb = sqrt(max(K.dotc(b_tensor, b_tensor).real, 0.0))
    if sqrt(max(K.dotc(r, r).real, 0.0)) <= max(atol, rtol * nb):
        return v
    x, _ = solver(apply, b_tensor, v, atol=atol, rtol=rtol)
    return x.view(v.shape)


def _sweep(QF, LF, direction, atol, rtol, solver, strategy):
    size = QF.state.size
    sites = range(size - 1) if direction > 0 else range(size - 2, -1, -1)
    for i in sites:
        AB = _solve_local(QF, i, LF.tensor2site(direction), atol, rtol, solver)
        if direction > 0:
            QF.update_2site_right(AB, i, strategy)
            LF.update_right()
        else:
            QF.update_2site_left(AB, i, strategy)
            LF.update_left()


def dmrg_solve(A: MPO, b: MPS, guess: MPS | None = None, maxiter: int = 20, atol: float = 0,
               rtol: float = 1e-5, strategy: Strategy = DEFAULT_STRATEGY,
               method: str = "bicgstab", compute_residuals: bool = True):
    """Solve A x = b (dmrg.py:91-188).  Returns (x, residual or None).
    Methods: 'cg', 'bicgstab', 'gmres'; 'lgmres' runs restarted GMRES."""
    if maxiter < 1:
        raise ValueError("maxiter must be positive")
    if guess is None:
        guess = b.copy()
    solver = SOLVERS.get(method)
    if solver is None:
        raise ValueError(f'Unknown solver "{method}"')
    b_norm = b.norm()
    tol = max(atol, rtol * b_norm)
    strat = strategy.replace(normalize=False)
    if not isinstance(guess, CanonicalMPS) or guess.center not in (0, A.size - 1):
        guess = CanonicalMPS(guess, center=0, strategy=strat)
    if guess.center == 0:
        direction = +1
        QF = QuadraticForm(A, guess, start=0)
        LF = AntilinearForm(guess, b, center=0)
    else:
        direction = -1
        QF = QuadraticForm(A, guess, start=A.size - 2)
        LF = AntilinearForm(guess, b, center=A.size - 1)
    residual = np.inf
    change_tol = max(rtol, atol / b_norm) if b_norm > 0 else rtol
    if compute_residuals:
        residual = (A @ QF.state - b).norm()
        if residual <= tol:
            return QF.state, residual
    for _ in range(maxiter):
        psi_old = None if compute_residuals else QF.state.copy()
        _sweep(QF, LF, direction, atol, rtol, solver, strat)
        direction = -direction
        if compute_residuals:
            residual = (A @ QF.state - b).norm()
            if residual <= tol:
                break
        elif _relative_change(QF.state, psi_old) <= change_tol:
            break
    return QF.state, (residual if compute_residuals else None)


__all__ = ["dmrg_solve"]
