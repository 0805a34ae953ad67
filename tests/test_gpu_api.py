"""Parity of the drop-in API (device path) with the oracle on seeded inputs.

Gauge-invariant comparisons: bond dimensions exactly, state vectors / overlaps /
energies / errors at FP64 tolerances (the reference's singular-vector phases
are not defined, see the oracle headers)."""

import numpy as np
import pytest

from oracle import core as oc, algorithms as oa, builders as ob, exact as ex

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def S():
    import seemps2_amd

    return seemps2_amd


def dev_mps(S, omps):
    return S.MPS([A for A in omps.tensors], error=omps.error)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def fidelity(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def test_library_is_native(S):
    from seemps2_amd import _lib

    assert _lib.LIB_PATH.exists() and _lib.ABI_VERSION == 1


def test_scprod_norm_vector(S, rng):
    a = ob.random_mps([2, 3, 2, 2, 2], D=5, complex=True, rng=rng)
    b = ob.random_mps([2, 3, 2, 2, 2], D=3, complex=True, rng=rng)
    A, B = dev_mps(S, a), dev_mps(S, b)
    assert rel(A.to_vector(), a.to_vector()) < 1e-14
    assert abs(S.scprod(A, B) - oc.scprod(a.tensors, b.tensors)) < 1e-12 * abs(
        oc.scprod(a.tensors, b.tensors))
    assert abs(A.norm_squared() - a.norm_squared()) < 1e-12 * a.norm_squared()


def test_canonical_mps(S, rng):
    a = ob.random_mps([2] * 9, D=6, complex=True, rng=rng)
    for center in (0, 4, 8):
        ref = oa.canonical(a, center, oc.DEFAULT_STRATEGY)
        out = S.CanonicalMPS(dev_mps(S, a), center=center)
        assert out.bond_dimensions() == ref.bond_dimensions()
        assert rel(out.to_vector(), ref.to_vector()) < 1e-12
        assert abs(out.norm_squared() - ref.norm_squared()) < 1e-12 * ref.norm_squared()
        for i, T in enumerate(out):
            M = T.cpu().numpy()
            if i < center:
                M = M.reshape(-1, M.shape[-1])
                assert np.abs(M.conj().T @ M - np.eye(M.shape[1])).max() < 1e-12
            elif i > center:
                M = M.reshape(M.shape[0], -1)
                assert np.abs(M @ M.conj().T - np.eye(M.shape[0])).max() < 1e-12
    # truncating canonical form: same bond dims and error as the oracle
    st = S.Strategy(max_bond_dimension=3)
    ost = oc.Strategy(max_bond_dimension=3)
    out = S.CanonicalMPS(dev_mps(S, a), center=0, strategy=st)
    ref = oa.canonical(a, 0, ost)
    assert out.bond_dimensions() == ref.bond_dimensions()
    assert abs(out.error() - ref.error) < 1e-12 * max(1.0, ref.error)
    assert fidelity(out.to_vector(), ref.to_vector()) > 1 - 1e-12


def test_mpo_apply_expectation(S, rng):
    N = 8
    H = ob.xxz_mpo(N)
    Hd = S.MPO(H)
    a = ob.random_mps([2] * N, D=4, complex=True, rng=rng)
    A = dev_mps(S, a)
    out = Hd.apply(A, simplify=False)
    ref = oa.mpo_apply(H, a, oc.NO_TRUNCATION, simplify_flag=False)
    assert out.bond_dimensions() == ref.bond_dimensions()
    assert rel(out.to_vector(), ref.to_vector()) < 1e-13
    e_ref = oa.mpo_expectation(H, a)
    assert abs(Hd.expectation(A) - e_ref) < 1e-12 * abs(e_ref)
    c = S.CanonicalMPS(A, center=3)
    w = c.to_vector()
    dense = ex.mpo_to_matrix(H)
    assert abs(Hd.expectation(c) - np.vdot(w, dense @ w)) < 1e-11 * abs(np.vdot(w, dense @ w))


def test_hamiltonian_to_mpo(S):
    for N in (2, 5):
        H = S.XXZHamiltonian(N, 1.0, 0.5)
        dense = H.to_matrix().toarray()
        assert np.abs(H.to_mpo().to_matrix() - dense).max() < 1e-13
        Hh = S.HeisenbergHamiltonian(N)
        assert np.abs(Hh.to_mpo().to_matrix() - Hh.to_matrix().toarray()).max() < 1e-13


@pytest.mark.parametrize("seed", [1, 2])
def test_mpo_apply_simplify_parity(S, seed):
    """Config-1 shape at reduced size: MPO.apply + simplify with max bond and 2 sweeps."""
    rng = np.random.default_rng(seed)
    N, D = 12, 8
    a = ob.random_mps([2] * N, D=D, truncate=False, complex=True, rng=rng)
    nrm = np.sqrt(a.norm_squared())
    a.tensors[0] = a.tensors[0] / nrm
    H = ob.heisenberg_mpo(N)
    ost = oc.Strategy(tolerance=1e-10, simplification_tolerance=1e-10, max_bond_dimension=16,
                      max_sweeps=2)
    st = S.Strategy(tolerance=1e-10, simplification_tolerance=1e-10, max_bond_dimension=16,
                    max_sweeps=2)
    ref = oa.mpo_apply(H, a, ost, simplify_flag=True)
    out = S.MPO(H).apply(dev_mps(S, a), strategy=st, simplify=True)
    assert out.bond_dimensions() == ref.bond_dimensions()
    target = ex.mpo_to_matrix(H) @ a.to_vector()
    f_ref = fidelity(ref.to_vector(), target)
    f_out = fidelity(out.to_vector(), target)
    assert abs(f_out - f_ref) < 1e-12
    assert abs(out.error() - ref.error) < 1e-12 * max(1.0, ref.error)
    assert rel(out.to_vector(), ref.to_vector()) < 1e-12


def test_simplify_sum_parity(S, rng):
    N = 10
    a = ob.random_mps([2] * N, D=5, complex=True, rng=rng)
    b = ob.random_mps([2] * N, D=6, complex=True, rng=rng)
    st = S.Strategy(tolerance=1e-8, simplification_tolerance=1e-8)
    ost = oc.Strategy(tolerance=1e-8, simplification_tolerance=1e-8)
    ref = oa.simplify(oa.OMPSSum([1.0, 0.5j], [a, b]), ost)
    out = S.simplify(S.MPSSum([1.0, 0.5j], [dev_mps(S, a), dev_mps(S, b)]), st)
    assert out.bond_dimensions() == ref.bond_dimensions()
    assert rel(out.to_vector(), ref.to_vector()) < 1e-12


def test_simplify_right_to_left_parity(S, rng):
    N = 10
    a = ob.random_mps([2] * N, D=7, complex=True, rng=rng)
    b = ob.random_mps([2] * N, D=5, complex=True, rng=rng)
    st = S.Strategy(tolerance=1e-9, simplification_tolerance=1e-9, max_bond_dimension=9)
    ost = oc.Strategy(tolerance=1e-9, simplification_tolerance=1e-9, max_bond_dimension=9)
    ref = oa.simplify(oa.OMPSSum([1.0, -0.3j], [a, b]), ost, direction=-1)
    out = S.simplify(S.MPSSum([1.0, -0.3j], [dev_mps(S, a), dev_mps(S, b)]), st, direction=-1)
    assert out.bond_dimensions() == ref.bond_dimensions()
    assert rel(out.to_vector(), ref.to_vector()) < 1e-12


def test_simplify_uneven_sums_parity(S):
    """Many-state sums of uneven length and bond dimension, complex64 and
    complex128 inputs (promoted to complex128 as the reference promotes inside
    its SVD/GEMM): simplify(psi_k + phi_k) matches the oracle (same bond
    dimensions, states within 1e-12) and the exact sum (fidelity)."""
    rng = np.random.default_rng(404)
    for k in range(6):
        N = int(rng.integers(10, 15))
        Da, Db = int(rng.integers(4, 25)), int(rng.integers(4, 25))
        a = ob.random_mps([2] * N, D=Da, complex=True, rng=rng)
        b = ob.random_mps([2] * N, D=Db, complex=True, rng=rng)
        if k % 2:
            a = oa.OMPS([A.astype(np.complex64).astype(np.complex128) for A in a.tensors], 0.0)
            da = S.MPS([A.astype(np.complex64) for A in a.tensors])
        else:
            da = dev_mps(S, a)
        st = S.Strategy(tolerance=1e-8, simplification_tolerance=1e-8)
        ost = oc.Strategy(tolerance=1e-8, simplification_tolerance=1e-8)
        ref = oa.simplify(oa.OMPSSum([1.0, 1.0], [a, b]), ost)
        out = S.simplify(S.MPSSum([1.0, 1.0], [da, dev_mps(S, b)]), st)
        assert out.bond_dimensions() == ref.bond_dimensions(), (k, N, Da, Db)
        assert rel(out.to_vector(), ref.to_vector()) < 1e-12, k
        exact = a.to_vector() + b.to_vector()
        v = out.to_vector()
        fid = abs(np.vdot(v, exact)) / (np.linalg.norm(v) * np.linalg.norm(exact))
        assert 1 - fid < 1e-7


def test_dmrg_parity_small(S):
    N = 10
    rng = np.random.default_rng(3)
    g = ob.random_mps([2] * N, D=4, complex=True, rng=rng)
    H = ob.xxz_mpo(N)
    ref = oa.dmrg(H, g.copy(), maxiter=10, strategy=oc.Strategy(max_bond_dimension=32))
    out = S.dmrg(S.MPO(H), dev_mps(S, g), maxiter=10, strategy=S.Strategy(max_bond_dimension=32))
    E0 = ex.ground_energy_sparse(ex.nn_hamiltonian_sparse([ob.xxz_interaction()] * (N - 1),
                                                          [2] * N))
    assert abs(out.energy - E0) < 1e-9
    assert abs(out.energy - ref.energy) < 1e-12 * abs(ref.energy)
    assert out.converged
    assert out.state.bond_dimensions() == ref.state.bond_dimensions()


def test_dmrg_free_fermions_gpu(S):
    N = 16
    rng = np.random.default_rng(5)
    g = ob.random_mps([2] * N, D=8, complex=True, rng=rng)
    out = S.dmrg(S.XXZHamiltonian(N, 1.0, 0.0), dev_mps(S, g), maxiter=12,
                 strategy=S.Strategy(max_bond_dimension=128))
    assert abs(out.energy - ex.xx_free_fermion_energy(N)) < 1e-9


def test_dmrg_mixed_precision_matches_fp64(S):
    """FP32 contractions in the early Lanczos restarts leave the DMRG energy
    within 1e-12 (relative) of the FP64 run and within 1e-9 of the exact
    free-fermion energy (the sweeps' convergence)."""
    import importlib

    # the package re-exports the dmrg function under the module's name
    D = importlib.import_module("seemps2_amd.optimization.dmrg")

    N = 16
    rng = np.random.default_rng(11)
    g = ob.random_mps([2] * N, D=8, complex=True, rng=rng)
    st = S.Strategy(max_bond_dimension=64)
    H = S.XXZHamiltonian(N, 1.0, 0.0)
    E64 = S.dmrg(H, dev_mps(S, g), maxiter=10, strategy=st).energy
    D.STATS.reset()
    out = S.dmrg(H, dev_mps(S, g), maxiter=10, strategy=st, precision="mixed")
    assert D.STATS.matvecs_fp32 > 0
    assert abs(out.energy - E64) <= 1e-12 * abs(E64)
    assert abs(out.energy - ex.xx_free_fermion_energy(N)) < 1e-9
    with pytest.raises(ValueError):
        S.dmrg(H, dev_mps(S, g), maxiter=1, strategy=st, precision="fp16")


def test_energy_and_variance_two_layer_environments(S):
    """The DMRG driver's variance | |H psi|^2 - E^2 | (dmrg.py:55-59) from
    environments that carry both MPO layers equals the norm of the explicit
    H psi (the reference's route) and the oracle's value."""
    import importlib

    D = importlib.import_module("seemps2_amd.optimization.dmrg")
    rng = np.random.default_rng(77)
    N = 9
    H = ob.xxz_mpo(N)
    a = ob.random_mps([2] * N, D=7, complex=True, rng=rng)
    psi = S.CanonicalMPS(dev_mps(S, a), center=0)
    E, var = D._energy_and_variance(S.MPO(H), psi)
    Hpsi = S.MPO(H).apply(psi, simplify=False)
    n_explicit = Hpsi.norm_squared()
    n_two_layer = D._applied_norm_squared(S.MPO(H), psi)
    assert abs(n_two_layer - n_explicit) < 1e-12 * n_explicit
    o = oa.canonical(a, 0, oc.DEFAULT_STRATEGY)
    E_ref, var_ref = oa.energy_and_variance(H, o)
    assert abs(E - E_ref) < 1e-12 * abs(E_ref)
    assert abs(var - var_ref) < 1e-11 * max(abs(var_ref), 1.0)
    # and against the dense operator
    v = psi.to_vector()
    Hd = ex.mpo_to_matrix(H)
    assert abs(n_two_layer - np.vdot(Hd @ v, Hd @ v).real) < 1e-11 * n_two_layer
