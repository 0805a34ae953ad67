"""Time integrators that drive MPO.apply + simplify (drop-in for the Euler
integrator of ``seemps.evolution``)."""

from .euler import euler, ode_solver
from .trotter import PairwiseUnitaries, Trotter2ndOrder, Trotter3rdOrder
from .tdvp import tdvp, tdvp_step

__all__ = ["PairwiseUnitaries", "Trotter2ndOrder", "Trotter3rdOrder", "euler", "ode_solver",
           "tdvp", "tdvp_step"]
