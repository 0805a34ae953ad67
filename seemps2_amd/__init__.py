"""seemps2_amd -- MI355X (gfx950) implementation of the SeeMPS tensor-train hot path.

Drop-in entry points: ``MPS``, ``CanonicalMPS``, ``MPSSum``, ``MPO.apply``,
``simplify``, ``Strategy`` and ``dmrg``, with the names, arguments and
defaults of the reference package.  All numerical work runs in the hand-written
HIP kernels of ``libseemps_hip.so`` (FP64 complex MFMA contractions, one-sided
Jacobi SVD, device Lanczos); torch provides device buffers and streams only.
"""

import sys as _sys

from . import core
from .core import (
    DEFAULT_STRATEGY,
    DEFAULT_TOLERANCE,
    MAX_BOND_DIMENSION,
    NO_TRUNCATION,
    GemmOrder,
    Simplification,
    Strategy,
    Truncation,
    scprod,
    vdot,
)
from .state import (
    MPS,
    MPSSum,
    CanonicalMPS,
    AntilinearForm,
    SIMPLIFICATION_STRATEGY,
    simplify,
    simplify_mps,
    product_state,
    random_mps,
    random_uniform_mps,
)
from .operators import MPO, MPOSum, QuadraticForm
from .state import mps_tensor_product
from .analysis import (
    QuantizedInterval,
    RegularInterval,
    finite_differences_mpo,
    interleaving_permutation,
    tridiagonal_mpo,
)
from .register import mpo_shifts, mpo_weighted_shifts
from .evolution import euler
from .solve import dmrg_solve
from .hamiltonians import (
    ConstantNNHamiltonian,
    ConstantTIHamiltonian,
    HeisenbergHamiltonian,
    NNHamiltonian,
    XXZHamiltonian,
)
from .optimization import OptimizeResults, dmrg

# ``seemps.cython`` is the reference's kernel facade; expose ours under that name.
cython = core
_sys.modules[__name__ + ".cython"] = core

__all__ = [
    "dmrg_solve",
    "MPOSum",
    "QuantizedInterval",
    "RegularInterval",
    "euler",
    "finite_differences_mpo",
    "interleaving_permutation",
    "mpo_shifts",
    "mpo_weighted_shifts",
    "mps_tensor_product",
    "tridiagonal_mpo",
    "AntilinearForm",
    "CanonicalMPS",
    "ConstantNNHamiltonian",
    "ConstantTIHamiltonian",
    "DEFAULT_STRATEGY",
    "DEFAULT_TOLERANCE",
    "GemmOrder",
    "HeisenbergHamiltonian",
    "MAX_BOND_DIMENSION",
    "MPO",
    "MPS",
    "MPSSum",
    "NNHamiltonian",
    "NO_TRUNCATION",
    "OptimizeResults",
    "QuadraticForm",
    "SIMPLIFICATION_STRATEGY",
    "Simplification",
    "Strategy",
    "Truncation",
    "XXZHamiltonian",
    "core",
    "dmrg",
    "product_state",
    "random_mps",
    "random_uniform_mps",
    "scprod",
    "simplify",
    "simplify_mps",
    "vdot",
]
