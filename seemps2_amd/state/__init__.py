"""Device-resident MPS objects (drop-in for ``seemps.state``)."""

from ..core import (
    DEFAULT_STRATEGY,
    DEFAULT_TOLERANCE,
    MAX_BOND_DIMENSION,
    NO_TRUNCATION,
    Simplification,
    Strategy,
    Truncation,
    scprod,
    vdot,
)
from .mps import MPS, TensorArray
from .mpssum import MPSSum, to_mps
from .canonical_mps import CanonicalMPS
from .factories import mps_ones, product_state, random_mps, random_uniform_mps
from .antilinear import AntilinearForm
from .simplification import SIMPLIFICATION_STRATEGY, combine, simplify, simplify_mps
from .compose import mps_tensor_product

__all__ = [
    "AntilinearForm",
    "CanonicalMPS",
    "DEFAULT_STRATEGY",
    "DEFAULT_TOLERANCE",
    "MAX_BOND_DIMENSION",
    "MPS",
    "MPSSum",
    "NO_TRUNCATION",
    "SIMPLIFICATION_STRATEGY",
    "Simplification",
    "Strategy",
    "TensorArray",
    "Truncation",
    "combine",
    "mps_ones",
    "mps_tensor_product",
    "product_state",
    "random_mps",
    "random_uniform_mps",
    "scprod",
    "simplify",
    "simplify_mps",
    "to_mps",
    "vdot",
]
