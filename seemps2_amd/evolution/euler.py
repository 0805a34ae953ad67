"""Explicit Euler for d|psi>/dt = L|psi> (reference: src/seemps/evolution/euler.py:10-70,
common.py:30-73).  Each step is psi + dt (L @ psi), simplified with the
integrator's strategy: one MPO application (device local contraction +
simplify with L's own strategy) and one simplify of a two-term MPSSum."""

from __future__ import annotations

from collections.abc import Callable

import numpy as np

from ..core import DEFAULT_STRATEGY, Strategy
from ..state import MPS, simplify


def ode_solver(evolve_for_dt: Callable, time, state: MPS, steps: int = 1000,
               strategy: Strategy = DEFAULT_STRATEGY, callback: Callable | None = None):
    """Step `evolve_for_dt` over the time grid: a total time T (from 0), a
    (t0, t1) pair split into `steps` steps, or an explicit sequence of times
    (common.py:30-73).  Returns the final state, or the callback's values."""
    if isinstance(time, (int, float, np.number)):
        time = (0.0, float(time))
    if isinstance(time, tuple):
        grid = np.linspace(float(time[0]), float(time[1]), steps + 1, dtype=np.float64)
    else:
        grid = np.asarray(time, dtype=np.float64)
    output = []
    last = grid[0]
    for t in grid:
        if t != last:
            state = evolve_for_dt(t, state, float(t - last), strategy)
        if callback:
            output.append(callback(t, state))
        last = t
    return output if callback else state


def euler(L, time, state: MPS, steps: int = 1000, strategy: Strategy = DEFAULT_STRATEGY,
          callback: Callable | None = None):
    """psi(t + dt) = psi(t) + dt L psi(t) (euler.py:10-70)."""

    def step(t: float, psi: MPS, dt: float, strategy: Strategy) -> MPS:
        return simplify(psi + dt * (L @ psi), strategy=strategy)

    return ode_solver(step, time, state, steps, strategy, callback)


__all__ = ["euler", "ode_solver"]
