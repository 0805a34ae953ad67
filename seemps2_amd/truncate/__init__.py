"""Alias module kept for API parity with ``seemps.truncate`` (truncate/__init__.py:1-6)."""

from ..state.antilinear import AntilinearForm
from ..state.simplification import simplify

__all__ = ["AntilinearForm", "simplify"]
