"""Time integrators that drive MPO.apply + simplify (drop-in for the Euler
integrator of ``seemps.evolution``)."""

from .euler import euler, ode_solver

__all__ = ["euler", "ode_solver"]
