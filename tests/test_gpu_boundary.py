"""Boundary behaviour of the device path: accepted input types, rejected ones,
memory held by cached operators, and sizes beyond the single-tile limits
(d = 4 chains, spectra longer than the truncation kernel's LDS stage)."""

import gc

import numpy as np
import pytest

from oracle import core as oc, algorithms as oa, builders as ob, exact as ex

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def S():
    import seemps2_amd

    return seemps2_amd


def crand(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def test_environment_facade_accepts_ndarrays(S, rng):
    """environments.pxi:10-85 accept NumPy arrays of any float/complex dtype."""
    from seemps2_amd import cython as C

    A = crand(rng, 5, 2, 7)
    B = crand(rng, 4, 2, 6)
    rhoL = crand(rng, 4, 5)
    rhoR = crand(rng, 7, 6)
    left = C._update_left_environment(B, A, rhoL)
    assert np.abs(left.cpu().numpy() - oc.update_left_environment(B, A, rhoL)).max() < 1e-12
    right = C._update_right_environment(B, A, rhoR)
    assert np.abs(right.cpu().numpy() - oc.update_right_environment(B, A, rhoR)).max() < 1e-12
    # real float64 inputs are promoted like the reference's mixed-dtype GEMMs
    Ar, Br, rr = A.real.copy(), B.real.copy(), rhoL.real.copy()
    out = C._update_left_environment(Br, Ar, rr)
    assert np.abs(out.cpu().numpy() - oc.update_left_environment(Br, Ar, rr)).max() < 1e-12
    rL, rR = crand(rng, 6, 7), crand(rng, 7, 6)
    assert abs(C._join_environments(rL, rR) - oc.join_environments(rL, rR)) < 1e-12
    assert abs(C._end_environment(np.array([[2.5 + 1j]])) - (2.5 + 1j)) == 0


def test_kernels_reject_bad_tensors(S):
    """Host-side guards: nothing that is not a contiguous complex128 device
    tensor reaches a launch."""
    from seemps2_amd import kernels as K

    good = K.as_dev(np.eye(4))
    cases = [
        torch.eye(4, dtype=torch.complex128),                      # host tensor
        torch.eye(4, dtype=torch.complex64, device=good.device),   # wrong dtype
        torch.eye(4, dtype=torch.float64, device=good.device),     # real
        K.as_dev(np.ones((4, 8)))[:, ::2],                         # strided view
    ]
    for bad in cases:
        with pytest.raises(ValueError):
            K.gemm(good, bad)
        with pytest.raises(ValueError):
            K.gemm(bad, good)
        with pytest.raises(ValueError):
            K.permute(bad, (1, 0))
    with pytest.raises(ValueError):
        K.axpby(1.0, cases[0], 0.0, good)
    with pytest.raises(ValueError):
        K.local(good, good, 2, 2, 2, 4, 1)  # sizes disagree with the dimensions


def test_mpo_operator_cache_does_not_grow(S, rng):
    """Permuted MPO operators live on the MPO tensors: dropping the MPOs frees them."""
    psi = S.MPS([crand(rng, 1, 2, 4), crand(rng, 4, 2, 4), crand(rng, 4, 2, 1)])
    torch.cuda.synchronize()

    def churn(n):
        for k in range(n):
            H = S.XXZHamiltonian(3, 1.0, 0.1 * k).to_mpo()
            H.expectation(psi)
            (H @ psi)
        gc.collect()
        torch.cuda.synchronize()
        return torch.cuda.memory_allocated()

    base = churn(20)
    after = churn(200)
    assert after <= base + (64 << 10), (base, after)


def spin32_heisenberg():
    m = np.array([1.5, 0.5, -0.5, -1.5])
    Sz = np.diag(m)
    Sp = np.zeros((4, 4))
    for k in range(1, 4):
        Sp[k - 1, k] = np.sqrt(1.5 * 2.5 - m[k] * (m[k] + 1))
    Sm = Sp.T
    return 0.5 * (np.kron(Sp, Sm) + np.kron(Sm, Sp)) + np.kron(Sz, Sz)


def test_two_site_matvec_d4_w5(S, rng):
    """w * d1 * d2 = 80 contracted channels (beyond one LDS tile) against the oracle."""
    from seemps2_amd import kernels as K

    h = spin32_heisenberg()
    mpo = ob.nn_mpo([h] * 3, [4] * 4)
    assert mpo[1].shape == (5, 4, 4, 5)
    L = crand(rng, 6, 5, 6)
    R = crand(rng, 7, 5, 7)
    v = crand(rng, 6, 4, 4, 7)
    ref = oa.two_site_matvec(L, np.einsum("aijb,bklc->aijklc", mpo[1], mpo[2]), R, v)
    Hm = K.two_site_operator(K.as_dev(mpo[1]), K.as_dev(mpo[2]))
    out = K.two_site_matvec(K.as_dev(L), Hm, K.as_dev(R), K.as_dev(v), 4, 4)
    assert np.abs(out.cpu().numpy() - ref).max() < 1e-13 * np.abs(ref).max()


def test_dmrg_spin32_heisenberg_d4(S, rng):
    """Two-site DMRG with d = 4, MPO bond 5 against exact diagonalisation."""
    N = 6
    h = spin32_heisenberg()
    H = S.ConstantTIHamiltonian(N, interaction=h).to_mpo()
    assert max(W.shape[0] for W in H) == 5
    E0 = ex.ground_energy_sparse(ex.nn_hamiltonian_sparse([h] * (N - 1), [4] * N))
    guess = ob.random_mps([4] * N, D=8, complex=True, rng=rng)
    out = S.dmrg(H, S.MPS(guess.tensors), maxiter=12,
                 strategy=S.Strategy(max_bond_dimension=64))
    assert abs(out.energy - E0) < 1e-10 * abs(E0), (out.energy, E0)


def test_dmrg_spin32_heisenberg_d4_learned_splits(S, rng):
    """The same d = 4 chain (N = 8, bond up to 256) with the learned-rank
    basis-sketched splits: the energy still matches exact diagonalisation."""
    from seemps2_amd import learned

    N = 8
    h = spin32_heisenberg()
    H = S.ConstantTIHamiltonian(N, interaction=h).to_mpo()
    E0 = ex.ground_energy_sparse(ex.nn_hamiltonian_sparse([h] * (N - 1), [4] * N))
    guess = ob.random_mps([4] * N, D=8, complex=True, rng=rng)
    lt = learned.LearnedTruncation(learned.RankPredictor(seed=3), min_size=32)
    with learned.learned_truncation(lt):
        out = S.dmrg(H, S.MPS(guess.tensors), maxiter=14,
                     strategy=S.Strategy(max_bond_dimension=256))
    assert lt.calls > 0 and lt.calls > lt.full_splits  # sketched splits ran
    assert abs(out.energy - E0) < 1e-10 * abs(E0), (out.energy, E0, lt)


def test_truncate_long_spectrum(S, rng):
    """Spectra longer than the LDS-staged truncation (4096 values) use the
    unstaged kernel, with the same decisions as the oracle's loops."""
    from seemps2_amd import kernels as K

    for n in (4096, 4097, 6000):
        s = np.sort(np.exp(-rng.uniform(0, 30, size=n)))[::-1].copy()
        for st in (oc.Strategy(tolerance=1e-10, max_bond_dimension=5000),
                   oc.Strategy(tolerance=1e-6, method=oc.Truncation.RELATIVE_SINGULAR_VALUE),
                   oc.Strategy(tolerance=1e-9, method=oc.Truncation.ABSOLUTE_SINGULAR_VALUE)):
            D_ref, e_ref = oc.truncate(s.copy(), st)
            D, e = K.truncate(torch.as_tensor(s, device="cuda"), st.method, st.tolerance,
                              st.max_bond_dimension)
            assert D == D_ref and e == e_ref, (n, st.method, D, D_ref, e, e_ref)


def test_destructively_truncate_vector_on_a_view(S, rng):
    """A non-owning view is shrunk in place, as the reference's dimension write does."""
    from seemps2_amd import cython as C

    base = np.sort(np.exp(-rng.uniform(0, 20, size=40)))[::-1].copy()
    view = base[:30]
    assert not view.flags.owndata
    st = S.Strategy(tolerance=1e-6)
    D_ref, e_ref = oc.truncate(view.copy(), oc.Strategy(tolerance=1e-6))
    err = C.destructively_truncate_vector(view, st)
    assert view.shape == (D_ref,) and err == e_ref
    assert np.array_equal(view, base[:D_ref])
