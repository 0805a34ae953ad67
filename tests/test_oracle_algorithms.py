"""Pin the oracle's algorithms against dense closed forms."""

import numpy as np
import pytest

from oracle import core, algorithms as al, builders as bl, exact as ex


def test_random_mps_bond_rule():
    # shapes asserted by tests/test_states/test_random_mps.py:14-27,64-75
    rng = np.random.default_rng(1)
    for D, shapes in [(1, [(1, 2, 1)] * 3), (2, [(1, 2, 2), (2, 2, 2), (2, 2, 1)]),
                      (3, [(1, 2, 3), (3, 2, 3), (3, 2, 1)])]:
        assert [A.shape for A in bl.random_mps([2] * 3, D=D, truncate=False, rng=rng).tensors] == shapes
    assert [A.shape for A in bl.random_mps([2, 3, 2], D=3, truncate=False, rng=rng).tensors] == \
        [(1, 2, 3), (3, 3, 3), (3, 2, 1)]
    r1 = bl.random_mps([2] * 10, D=4, rng=np.random.default_rng(0x1231))
    r2 = bl.random_mps([2] * 10, D=4, rng=np.random.default_rng(0x1231))
    assert all(np.array_equal(a, b) for a, b in zip(r1.tensors, r2.tensors))
    # N > 60 forces uniform bonds (factories.py:208-209)
    big = bl.random_mps([2] * 64, D=16, complex=True, rng=np.random.default_rng(2))
    assert big.bond_dimensions() == [1] + [16] * 63 + [1]


def test_random_mps_draw_order():
    """factories.py:209-218 draws, site by site, the real part of each tensor
    (shape (DL, d, DR), C order) and then its imaginary part from the one
    generator.  So the oracle's tensors, read in that order, must be exactly the
    generator's stream.  The reference package itself cannot be imported here
    (seemps/__init__.py imports hdf5.py, which imports h5py, absent), so this
    and the shape rules above are what pin the draw order."""
    for dims, D, cplx in [([2] * 7, 5, True), ([3, 2, 4, 2], 6, True), ([2] * 5, 3, False)]:
        psi = bl.random_mps(dims, D=D, complex=cplx, rng=np.random.default_rng(77))
        parts = []
        for A in psi.tensors:
            parts.append(np.ascontiguousarray(A.real).ravel())
            if cplx:
                parts.append(np.ascontiguousarray(A.imag).ravel())
        drawn = np.concatenate(parts)
        stream = np.random.default_rng(77).normal(size=drawn.size)
        assert np.array_equal(drawn, stream)
        # a planted mistake (imaginary part drawn first) breaks it
        if cplx:
            swapped = []
            for A in psi.tensors:
                swapped += [A.imag.ravel(), A.real.ravel()]
            assert not np.array_equal(np.concatenate(swapped), stream)


@pytest.mark.parametrize("N", [2, 3, 6])
def test_nn_mpo_matches_dense(N):
    for h in (bl.heisenberg_interaction(), bl.xxz_interaction(1.0, 0.5)):
        H = bl.nn_mpo([h] * (N - 1), [2] * N)
        dense = ex.nn_hamiltonian_sparse([h] * (N - 1), [2] * N).toarray()
        assert np.abs(ex.mpo_to_matrix(H) - dense).max() < 1e-14
        assert max(W.shape[0] for W in H) == (5 if N > 2 else 1) or N == 2


def test_mpo_apply_expectation_dense(rng):
    N = 7
    H = bl.xxz_mpo(N)
    Hd = ex.mpo_to_matrix(H)
    psi = bl.random_mps([2] * N, D=4, complex=True, rng=rng)
    v = psi.to_vector()
    Hpsi = al.mpo_apply(H, psi, core.NO_TRUNCATION, simplify_flag=False)
    assert np.linalg.norm(Hpsi.to_vector() - Hd @ v) < 1e-12 * np.linalg.norm(Hd @ v)
    assert abs(al.mpo_expectation(H, psi) - np.vdot(v, Hd @ v)) < 1e-11 * abs(np.vdot(v, Hd @ v))
    can = al.canonical(psi, 3, core.DEFAULT_STRATEGY)
    w = can.to_vector()
    assert abs(al.mpo_expectation(H, can) - np.vdot(w, Hd @ w)) < 1e-11 * abs(np.vdot(w, Hd @ w))


def test_two_site_matvec_is_projected_hamiltonian(rng):
    """H_eff = P^dagger H P with P the isometry built from the other sites."""
    N = 6
    H = bl.xxz_mpo(N)
    Hd = ex.mpo_to_matrix(H)
    psi = al.canonical(bl.random_mps([2] * N, D=4, complex=True, rng=rng), 2, core.DEFAULT_STRATEGY)
    QF = al.QuadraticForm(H, psi, start=2)
    L, H12, R = QF.two_site_block(2)
    T = psi.tensors
    theta_shape = (T[2].shape[0], 2, 2, T[3].shape[2])
    n = int(np.prod(theta_shape))
    P = np.zeros((2 ** N, n), complex)
    for k in range(n):
        e = np.zeros(n, complex)
        e[k] = 1
        psi_k = np.einsum("...a,abcd->...bcd", _left_block(T[:2]), e.reshape(theta_shape))
        psi_k = np.tensordot(psi_k, _right_block(T[4:]), axes=([-1], [0]))
        P[:, k] = psi_k.reshape(-1)
    Heff = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros(n, complex)
        e[k] = 1
        Heff[:, k] = al.two_site_matvec(L, H12, R, e.reshape(theta_shape)).reshape(-1)
    assert np.abs(Heff - P.conj().T @ Hd @ P).max() < 1e-12


def _left_block(Ts):
    x = np.ones((1, 1))
    for A in Ts:
        a, d, b = A.shape
        x = (x @ A.reshape(a, d * b)).reshape(-1, b)
    return x


def _right_block(Ts):
    x = np.ones((1, 1))
    for A in reversed(Ts):
        a, d, b = A.shape
        x = (A.reshape(a * d, b) @ x).reshape(a, -1)
    return x


def test_simplify_exact_and_truncated(rng):
    N = 10
    H = bl.heisenberg_mpo(N)
    psi = bl.random_mps([2] * N, D=6, complex=True, rng=rng)
    target = ex.mpo_to_matrix(H) @ psi.to_vector()
    Hpsi = al.mpo_apply(H, psi, core.NO_TRUNCATION, simplify_flag=False)
    exact = al.simplify(Hpsi, core.Strategy(tolerance=1e-14, simplification_tolerance=1e-14))
    w = exact.to_vector()
    assert np.linalg.norm(w - target) < 1e-10 * np.linalg.norm(target)
    capped = al.simplify(Hpsi, core.Strategy(max_bond_dimension=8, simplification_tolerance=1e-12))
    assert max(capped.bond_dimensions()) <= 8
    w = capped.to_vector()
    fid = abs(np.vdot(w, target)) ** 2 / np.vdot(w, w).real / np.vdot(target, target).real
    assert 0.5 < fid < 1.0
    # normalised output when the strategy asks for it
    nrm = al.simplify(Hpsi, core.Strategy(normalize=True, max_bond_dimension=8))
    assert abs(np.linalg.norm(nrm.to_vector()) - 1) < 1e-12


def test_simplify_sum_dense(rng):
    N = 8
    a = bl.random_mps([2] * N, D=3, complex=True, rng=rng)
    b = bl.random_mps([2] * N, D=4, complex=True, rng=rng)
    s = al.OMPSSum([0.5, -1.2j], [a, b])
    ref = s.to_vector()
    out = al.simplify(s, core.Strategy(tolerance=1e-14, simplification_tolerance=1e-14))
    assert np.linalg.norm(out.to_vector() - ref) < 1e-10 * np.linalg.norm(ref)
    assert max(out.bond_dimensions()) <= 7


def test_dmrg_exact_diagonalisation(rng):
    for N, h in [(6, bl.heisenberg_interaction()), (8, bl.xxz_interaction(1.0, 0.5))]:
        H = bl.nn_mpo([h] * (N - 1), [2] * N)
        E0 = ex.ground_energy_sparse(ex.nn_hamiltonian_sparse([h] * (N - 1), [2] * N))
        r = al.dmrg(H, bl.random_mps([2] * N, D=2, complex=True, rng=rng))
        assert abs(r.energy - E0) < 1e-9
        assert r.converged
        assert len(r.trajectory) == len(r.variances)
        assert r.variances[-1] < 1e-7


def test_dmrg_free_fermions():
    N = 10
    rng = np.random.default_rng(7)
    r = al.dmrg(bl.xxz_mpo(N, Jxy=1.0, Jz=0.0), bl.random_mps([2] * N, D=4, complex=True, rng=rng),
                strategy=core.Strategy(max_bond_dimension=32))
    assert abs(r.energy - ex.xx_free_fermion_energy(N)) < 1e-8


def test_convergence_reason():
    # tests/test_optimization/test_dmrg.py:149-176
    assert al.convergence_reason(1.0, 1.0, 1e-10, 1e-10) is not None
    assert al.convergence_reason(1e-4, 1.0, 1e-10, 1e-3) is None
    assert al.convergence_reason(-1e-11, 1.0, 1e-10, 1e-10) is not None
    assert al.convergence_reason(-1.0, 1.0, 1e-10, 1e-10) is None
    assert al.convergence_reason(-5.0, 100.0, 1e-1, 1e-1) is not None
    assert al.convergence_reason(-5.0, 1.0, 1e-1, 1e-1) is None
