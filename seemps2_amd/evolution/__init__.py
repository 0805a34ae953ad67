"""Time integrators that drive MPO.apply + simplify (drop-in for the Euler
integrator of ``seemps.evolution``)."""

from .euler import euler, ode_solver
from .trotter import PairwiseUnitaries, Trotter2ndOrder, Trotter3rdOrder

__all__ = ["PairwiseUnitaries", "Trotter2ndOrder", "Trotter3rdOrder", "euler", "ode_solver"]
