"""CPU oracle for the SeeMPS tensor-train hot path (TEST INFRASTRUCTURE ONLY).

This package is a plain NumPy/SciPy restatement of the reference algorithm, used
as the checker for the HIP product path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to check or time the CPU restatement.  Nothing under
``seemps2_amd/`` imports it; the product path never routes through it.

Modules
-------
core        Strategy / truncation rules, SVD split, canonical form, MPS
            environments, contractions (reference ``src/seemps/cython``).
algorithms  MPS / MPSSum / MPO containers as tensor lists, ``simplify``,
            ``MPO.apply``, MPO environments, two-site effective operator and the
            two-site ``dmrg`` driver (reference ``src/seemps/state``,
            ``operators``, ``optimization``).
builders    Seeded input builders: ``random_mps``, ``product_state``,
            nearest-neighbour Hamiltonian MPOs (reference ``state/factories.py``,
            ``hamiltonians/nearest_neighbor.py``).
exact       Closed forms used to pin the oracle: dense state vectors and
            operators, exact diagonalisation, free-fermion energies.

Pinning status of each function is listed in the module headers.
"""

from . import core, algorithms, builders, exact  # noqa: F401
