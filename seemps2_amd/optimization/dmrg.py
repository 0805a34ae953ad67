"""Two-site DMRG on the device (reference: src/seemps/optimization/dmrg.py:15-199).

Same driver as the reference: normalised canonical guess, QuadraticForm
environments, alternating two-site sweeps, energy / variance after each sweep
(energy NOT divided by <psi|psi>, variance on the unsimplified H psi), the best
state kept, and the ``_convergence_reason`` stopping rule.  The local
eigenproblem (scipy eigsh -> ARPACK in the reference) is the device Lanczos of
libseemps_hip.so with Krylov dimension ``ncv`` (ARPACK's default 20).
"""

from __future__ import annotations

import dataclasses
from typing import Callable

import numpy as np

from .. import kernels as K
from ..core import DEFAULT_STRATEGY, Strategy, _contract_last_and_first
from ..hamiltonians import NNHamiltonian
from ..operators.mpo import MPO
from ..operators.quadratic import QuadraticForm
from ..state import MPS, CanonicalMPS, random_mps
from ..state.environments import begin_mpo2_environment, update_left_mpo2_environment

NCV = 20
# Arithmetic of the local eigensolver's contractions: "fp64" (the reference's),
# or "mixed" (FP32 products until the residual reaches 1e-5 of the eigenvalue,
# FP64 from there to tol_eigs; same stopping test as "fp64").
PRECISIONS = ("fp64", "mixed")


@dataclasses.dataclass
class OptimizeResults:
    """Results from ground state search (optimization/descent.py:22-49)."""

    state: MPS
    energy: float
    converged: bool
    message: str
    trajectory: list = dataclasses.field(default_factory=list)
    variances: list = dataclasses.field(default_factory=list)


class SweepStats:
    """Counters of the last sweeps (bond updates, matvecs), for benchmarks."""

    def __init__(self):
        self.bond_updates = 0
        self.matvecs = 0
        self.matvecs_fp32 = 0
        self.restarts = 0

    def reset(self):
        self.__init__()


STATS = SweepStats()


def _diagonalize_two_site(QF: QuadraticForm, i: int, tol_eigs: float, ncv: int = NCV,
                          precision: str = "fp64"):
    """Lowest eigenpair of the two-site effective Hamiltonian (dmrg.py:15-23)."""
    L, Hm, R = QF.two_site_block(i)
    v = _contract_last_and_first(QF.state[i], QF.state[i + 1])
    d1 = QF.O[i].shape[1]
    d2 = QF.O[i + 1].shape[1]
    eval_, evec, stats = K.eigensolve(L, Hm, R, v, d1, d2, ncv=ncv, tol=tol_eigs,
                                      precision=precision)
    STATS.matvecs += stats["matvecs"]
    STATS.matvecs_fp32 += stats["matvecs_fp32"]
    STATS.restarts += stats["restarts"]
    STATS.bond_updates += 1
    return eval_, evec


def _sweep(QF: QuadraticForm, direction: int, tol_eigs: float, strategy: Strategy,
           precision: str = "fp64") -> float:
    """One two-site sweep updating QF in place (dmrg.py:26-42)."""
    size = QF.state.size
    sites = range(size - 1) if direction > 0 else range(size - 2, -1, -1)
    E = np.inf
    for i in sites:
        E, AB = _diagonalize_two_site(QF, i, tol_eigs, precision=precision)
        if direction > 0:
            QF.update_2site_right(AB, i, strategy)
        else:
            QF.update_2site_left(AB, i, strategy)
    return E


def _convergence_reason(energy_change: float, energy_scale: float, tol: float,
                        tol_up: float) -> str | None:
    """dmrg.py:45-52."""
    if energy_change > tol_up * energy_scale:
        return f"Energy fluctuation above tolerance {tol_up}"
    if -tol * energy_scale <= energy_change <= 0:
        return f"Energy decrease slower than tolerance {tol}"
    return None


def _applied_norm_squared(H: MPO, state: MPS) -> float:
    """|H psi|^2 = <psi|H^H H|psi> by environments that carry both MPO layers
    (bond w^2 chi^2), the same number as the norm of the unsimplified H psi
    (bond w chi) that the reference builds, at about a quarter of the
    contraction work for w = 5."""
    E = begin_mpo2_environment()
    for A, O in zip(state, H):
        E = update_left_mpo2_environment(E, A, O)
    return abs(complex(E.reshape(-1)[0].item()))


def _energy_and_variance(H: MPO, state: MPS) -> tuple[float, float]:
    """dmrg.py:55-59: energy <psi|H|psi> (not divided by <psi|psi>) and the
    variance | |H psi|^2 - energy^2 |."""
    energy = H.expectation(state).real
    variance = abs(_applied_norm_squared(H, state) - energy * energy)
    return energy, variance


def _state_deepcopy(state: MPS) -> MPS:
    data = [K.copy(A) for A in state]
    if isinstance(state, CanonicalMPS):
        return CanonicalMPS(data, center=state.center, normalize=False,
                            strategy=state.strategy, is_canonical=True, error=state.error())
    return MPS(data, error=state.error())


def dmrg(H, guess: MPS | None = None, maxiter: int = 20, tol: float = 1e-10,
         tol_up: float | None = None, tol_eigs: float | None = None,
         strategy: Strategy = DEFAULT_STRATEGY, callback: Callable | None = None,
         precision: str = "fp64") -> OptimizeResults:
    """Ground state of an MPO (or NNHamiltonian) by two-site DMRG (dmrg.py:76-199).

    ``precision`` (not in the reference) selects the local eigensolver's
    arithmetic: "fp64", or "mixed" (see ``PRECISIONS``)."""
    if maxiter < 1:
        raise ValueError("maxiter must be positive")
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    if isinstance(H, NNHamiltonian):
        H = H.to_mpo()
    if H.size < 2:
        raise ValueError("DMRG requires at least two sites")
    if guess is None:
        guess = random_mps(H.physical_dimensions(), D=2)
    if tol < 0:
        raise ValueError("tol must be non-negative")
    if tol_up is None:
        tol_up = tol
    elif tol_up < 0:
        raise ValueError("tol_up must be non-negative")
    if tol_eigs is None:
        tol_eigs = tol
    strat = strategy.replace(normalize=True)
    if not isinstance(guess, CanonicalMPS) or guess.center not in (0, H.size - 1):
        guess = CanonicalMPS(guess, center=0, strategy=strat)
    guess.normalize_inplace()
    if guess.center == 0:
        direction = +1
        QF = QuadraticForm(H, guess, start=0)
    else:
        direction = -1
        QF = QuadraticForm(H, guess, start=H.size - 2)
    energy, variance = _energy_and_variance(H, QF.state)
    results = OptimizeResults(state=_state_deepcopy(QF.state), energy=energy, converged=False,
                              message=f"Exceeded maximum number of steps {maxiter}",
                              trajectory=[energy], variances=[variance])
    if callback is not None:
        callback(QF.state, results)
    last_energy = energy
    for _sweep_no in range(1, maxiter + 1):
        _sweep(QF, direction, tol_eigs, strat, precision)
        direction = -direction
        energy, variance = _energy_and_variance(H, QF.state)
        results.trajectory.append(energy)
        results.variances.append(variance)
        if energy < results.energy:
            results.energy, results.state = energy, _state_deepcopy(QF.state)
        if callback is not None:
            callback(QF.state, results)
        reason = _convergence_reason(energy - last_energy, max(1.0, abs(energy)), tol, tol_up)
        if reason is not None:
            results.converged = True
            results.message = reason
            break
        last_energy = energy
    return results
