"""Parity of the raw C-ABI kernels against NumPy / the oracle on the device."""

import numpy as np
import pytest

from oracle import core as ocore

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def K():
    from seemps2_amd import kernels

    return kernels


def crandn(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


OPS = {0: lambda x: x, 1: lambda x: x.T, 2: lambda x: x.conj().T, 3: lambda x: x.conj()}


@pytest.mark.parametrize("opA", [0, 1, 2, 3])
@pytest.mark.parametrize("opB", [0, 1, 2, 3])
@pytest.mark.parametrize("M,N,K_", [(37, 53, 29), (130, 70, 200), (1, 1, 1), (64, 64, 16)])
def test_zgemm_ops(K, rng, opA, opB, M, N, K_):
    A = crandn(rng, *((K_, M) if opA in (1, 2) else (M, K_)))
    B = crandn(rng, *((N, K_) if opB in (1, 2) else (K_, N)))
    C = crandn(rng, M, N)
    alpha, beta = 0.7 - 0.2j, -0.3 + 1.1j
    ref = alpha * OPS[opA](A) @ OPS[opB](B) + beta * C
    Cd = K.as_dev(C)
    out = K.gemm(K.as_dev(A), K.as_dev(B), opA, opB, alpha, beta, Cd)
    assert rel(K.to_numpy(out), ref) < 1e-14


@pytest.mark.parametrize("opA,opB", [(0, 0), (2, 3), (1, 2), (3, 1)])
def test_zgemm_large_tile_path(K, rng, opA, opB):
    """Outputs with at least four 64 x 64 tiles per CU take the 64 x 64 kernel
    (smaller ones the 32 x 32 kernel): ragged M, N and K on that path."""
    M, N, K_ = 2048 + 17, 2048 + 40, 37
    assert (M // 64 + 1) * (N // 64 + 1) >= 4 * torch.cuda.get_device_properties(0).multi_processor_count
    A = crandn(rng, *((K_, M) if opA in (1, 2) else (M, K_)))
    B = crandn(rng, *((N, K_) if opB in (1, 2) else (K_, N)))
    C = crandn(rng, M, N)
    alpha, beta = -0.4 + 0.9j, 0.25 - 0.5j
    ref = alpha * OPS[opA](A) @ OPS[opB](B) + beta * C
    out = K.gemm(K.as_dev(A), K.as_dev(B), opA, opB, alpha, beta, K.as_dev(C))
    assert rel(K.to_numpy(out), ref) < 1e-14


def _int_operands(rng, opA, opB, M, N, K_):
    def ints(*shape):
        return (rng.integers(-3, 4, size=shape) + 1j * rng.integers(-3, 4, size=shape)).astype(
            np.complex128)

    A = ints(*((K_, M) if opA in (1, 2) else (M, K_)))
    B = ints(*((N, K_) if opB in (1, 2) else (K_, N)))
    return A, B


@pytest.mark.parametrize("opA", [0, 1, 2, 3])
@pytest.mark.parametrize("opB", [0, 1, 2, 3])
@pytest.mark.parametrize("M,N,K_", [(37, 53, 29), (130, 70, 200), (2065, 2088, 24)])
def test_zgemm_exact_on_small_integers(K, rng, opA, opB, M, N, K_):
    """Gaussian-integer operands: every partial sum of either product scheme
    (3M on the 32 x 32 tile, 4M on the 64 x 64 one, the last shape) is an
    exact integer, so the result must equal the exact product bit for bit.
    A wrong sign or a dropped term of the 3M recombination fails here."""
    A, B = _int_operands(rng, opA, opB, M, N, K_)
    ref = OPS[opA](A) @ OPS[opB](B)
    out = K.to_numpy(K.gemm(K.as_dev(A), K.as_dev(B), opA, opB))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("opA,opB", [(0, 0), (2, 1), (3, 3)])
def test_zgemm_real_operands_give_real_products(K, rng, opA, opB):
    """Real operands give an imaginary part of exactly zero: the 3M sum
    product (ar + 0)(br + 0) repeats the ar br product operation for
    operation, so t3 - t1 - t2 cancels exactly."""
    M, N, K_ = 300, 200, 257
    A = rng.normal(size=(K_, M) if opA in (1, 2) else (M, K_)).astype(np.complex128)
    B = rng.normal(size=(N, K_) if opB in (1, 2) else (K_, N)).astype(np.complex128)
    out = K.to_numpy(K.gemm(K.as_dev(A), K.as_dev(B), opA, opB))
    assert np.all(out.imag == 0.0)
    assert rel(out, OPS[opA](A) @ OPS[opB](B)) < 1e-14


def test_zgemm_4m_scheme_knob(tmp_path):
    """SEEMPS_GEMM_SCHEME=4m (read once per process) switches the 32 x 32
    tile back to four real products: same results, in a process of its own."""
    import os
    import subprocess
    import sys
    from pathlib import Path

    script = tmp_path / "gemm4m.py"
    script.write_text(
        "import numpy as np\n"
        "from seemps2_amd import kernels as K\n"
        "rng = np.random.default_rng(5)\n"
        "OPS = {0: lambda x: x, 1: lambda x: x.T, 2: lambda x: x.conj().T, 3: lambda x: x.conj()}\n"
        "for opA, opB in [(0, 0), (2, 0), (1, 3), (3, 2)]:\n"
        "    M, N, Kd = 130, 70, 200\n"
        "    A = rng.integers(-3, 4, size=(Kd, M) if opA in (1, 2) else (M, Kd)) + 1j * "
        "rng.integers(-3, 4, size=(Kd, M) if opA in (1, 2) else (M, Kd))\n"
        "    B = rng.integers(-3, 4, size=(N, Kd) if opB in (1, 2) else (Kd, N)) + 1j * "
        "rng.integers(-3, 4, size=(N, Kd) if opB in (1, 2) else (Kd, N))\n"
        "    out = K.to_numpy(K.gemm(K.as_dev(A), K.as_dev(B), opA, opB))\n"
        "    assert np.array_equal(out, OPS[opA](A) @ OPS[opB](B)), (opA, opB)\n"
        "print('4m ok')\n"
    )
    root = Path(__file__).resolve().parents[1]
    env = dict(os.environ, SEEMPS_GEMM_SCHEME="4m", SEEMPS_GEMM_TILE="32x32",
               PYTHONPATH=str(root))
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "4m ok" in r.stdout


@pytest.mark.parametrize("opA,opB", [(0, 0), (2, 0), (0, 1), (3, 2), (1, 3)])
@pytest.mark.parametrize("M,N,K_", [(64, 48, 2048), (512, 512, 1280), (5, 1, 9000)])
def test_zgemm_splitk_through_python(K, rng, opA, opB, M, N, K_):
    """Shapes with few output tiles and long K take the split-K path of the
    Python-level gemm (alpha, beta applied in the slice reduction)."""
    from seemps2_amd._lib import lib

    assert lib.seemps_zgemm_workspace(M, N, K_) > 0
    A = crandn(rng, *((K_, M) if opA in (1, 2) else (M, K_)))
    B = crandn(rng, *((N, K_) if opB in (1, 2) else (K_, N)))
    C = crandn(rng, M, N)
    alpha, beta = 0.3 + 0.4j, -1.1 + 0.2j
    ref = alpha * OPS[opA](A) @ OPS[opB](B) + beta * C
    out = K.gemm(K.as_dev(A), K.as_dev(B), opA, opB, alpha, beta, K.as_dev(C))
    assert rel(K.to_numpy(out), ref) < 1e-13


@pytest.mark.parametrize("opA", [0, 1, 2, 3])
@pytest.mark.parametrize("opB", [0, 3, 1])
@pytest.mark.parametrize("M,N,K_", [(37, 53, 29), (130, 70, 200)])
def test_zgemm_fp32_exact_on_small_integers(K, rng, opA, opB, M, N, K_):
    """Integer operands in [-4, 4]: every product and partial sum is exact in
    FP32, so the FP32 path must equal the exact result (pins the FP32 MFMA
    output layout independently of rounding)."""
    def ints(*shape):
        return rng.integers(-4, 5, size=shape) + 1j * rng.integers(-4, 5, size=shape)

    A = ints(*((K_, M) if opA in (1, 2) else (M, K_)))
    B = ints(*((N, K_) if opB in (1, 2) else (K_, N)))
    C = ints(M, N)
    ref = OPS[opA](A) @ OPS[opB](B) - C
    out = K.gemm(K.as_dev(A), K.as_dev(B), opA, opB, 1.0, -1.0, K.as_dev(C), precision="fp32")
    assert np.array_equal(K.to_numpy(out), ref)


@pytest.mark.parametrize("M,N,K_", [(1280, 1024, 256), (300, 77, 1000)])
def test_zgemm_fp32_rounding(K, rng, M, N, K_):
    """Gaussian operands: the FP32 product's error stays at FP32 rounding
    (normwise relative <= sqrt(K) 2^-24 up to a small factor; 4e-7 measured at
    K = 256), far above FP64's."""
    A = crandn(rng, M, K_)
    B = crandn(rng, K_, N)
    ref = A @ B
    out = K.to_numpy(K.gemm(K.as_dev(A), K.as_dev(B), precision="fp32"))
    err = rel(out, ref)
    assert err < 2e-6 * np.sqrt(K_ / 256)
    assert err > 1e-12  # really computed in FP32


def test_zlocal(K, rng):
    for (P, I, Q, J1, J2) in [(7, 20, 130, 20, 1), (33, 2, 40, 10, 5), (1, 1, 1, 1, 1), (5, 10, 64, 10, 1)]:
        x = crandn(rng, P, I, Q)
        Mop = crandn(rng, I, J1, J2)
        ref = np.einsum("piq,ijk->pjqk", x, Mop)
        out = K.local(K.as_dev(x), K.as_dev(Mop), P, I, Q, J1, J2)
        assert rel(K.to_numpy(out), ref) < 1e-14
        out2 = K.local(K.as_dev(x), K.as_dev(Mop), P, I, Q, J1, J2, out=out, accumulate=True)
        assert rel(K.to_numpy(out2), 2 * ref) < 1e-14


def test_zdotc(K, rng):
    for n in [1, 17, 4096, 100003]:
        x, y = crandn(rng, n), crandn(rng, n)
        assert abs(K.dotc(K.as_dev(x), K.as_dev(y)) - np.vdot(x, y)) < 1e-12 * np.sqrt(n) * 4


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_truncate_matches_oracle(K, rng, method):
    for trial in range(20):
        N = int(rng.integers(1, 300))
        s = np.ascontiguousarray(np.sort(np.exp(rng.normal(scale=6.0, size=N)))[::-1])
        tol = float(10.0 ** rng.uniform(-14, -1))
        mb = int(rng.integers(1, N + 2))
        st = ocore.Strategy(method=method, tolerance=tol, max_bond_dimension=mb)
        D_ref, err_ref = ocore.truncate(s, st)
        D, err = K.truncate(torch.tensor(s, device="cuda"), st.method, st.tolerance,
                            st.max_bond_dimension)
        assert D == D_ref
        assert err == err_ref  # same summation order: bit-exact


@pytest.mark.parametrize("m,n", [(1, 1), (5, 3), (3, 5), (16, 16), (17, 40), (64, 64),
                                 (128, 96), (200, 300), (512, 512), (320, 160)])
def test_svd_random(K, rng, m, n):
    A = crandn(rng, m, n)
    U, S, VH = K.svd(K.as_dev(A))
    U, S, VH = K.to_numpy(U), S.cpu().numpy(), K.to_numpy(VH)
    r = min(m, n)
    assert rel((U * S) @ VH, A) < 1e-12
    assert np.abs(U.conj().T @ U - np.eye(r)).max() < 1e-12
    assert np.abs(VH @ VH.conj().T - np.eye(r)).max() < 1e-12
    s_ref = np.linalg.svd(A, compute_uv=False)
    assert np.abs(S - s_ref).max() < 5e-13 * s_ref[0]
    assert np.all(np.diff(S) <= 0)


@pytest.mark.parametrize("n,rate", [(64, 0.3), (128, 0.25), (256, 0.12), (512, 0.07)])
def test_svd_graded_spectrum(K, rng, n, rate):
    """Singular values spanning ~1e-16..1 with random singular vectors (DMRG-like)."""
    U0, _ = np.linalg.qr(crandn(rng, n, n))
    V0, _ = np.linalg.qr(crandn(rng, n, n))
    s0 = np.exp(-rate * np.arange(n))
    A = (U0 * s0) @ V0
    U, S, VH = K.svd(K.as_dev(A))
    U, S, VH = K.to_numpy(U), S.cpu().numpy(), K.to_numpy(VH)
    assert rel((U * S) @ VH, A) < 1e-12
    assert np.abs(U.conj().T @ U - np.eye(n)).max() < 1e-12
    assert np.abs(VH @ VH.conj().T - np.eye(n)).max() < 1e-12
    s_ref = np.linalg.svd(A, compute_uv=False)
    assert np.abs(S - s_ref).max() < 5e-13 * s_ref[0]


def test_svd_rank_deficient_and_zero(K, rng):
    A = crandn(rng, 20, 3) @ crandn(rng, 3, 20)
    A[:, 5] = 0
    for M in [A, A.T.copy(), np.zeros((6, 4)), np.zeros((4, 6))]:
        U, S, VH = K.svd(K.as_dev(M))
        U, S, VH = K.to_numpy(U), S.cpu().numpy(), K.to_numpy(VH)
        r = min(M.shape)
        assert rel((U * S) @ VH, M + 0j) < 1e-12 or np.linalg.norm(M) == 0
        assert np.abs(U.conj().T @ U - np.eye(r)).max() < 1e-12
        assert np.abs(VH @ VH.conj().T - np.eye(r)).max() < 1e-12


def test_svd_split_rank_deficient_keeps_isometry(K, rng):
    """Kept columns for zero singular values (no truncation) still form an isometry."""
    for (a, d, b), rank in [((16, 2, 16), 3), ((40, 2, 24), 5), ((4, 2, 4), 1)]:
        AA = (crandn(rng, a * d, rank) @ crandn(rng, rank, d * b)).reshape(a, d, d, b)
        for mode in (0, 1):
            Ad, Bd, D, err2, _ = K.svd_split(K.as_dev(AA), a * d, d * b, mode, 0, 0.0, 10 ** 6)
            Ad, Bd = K.to_numpy(Ad), K.to_numpy(Bd)
            assert D == min(a * d, d * b)
            assert err2 == 0.0
            assert rel(Ad @ Bd, AA.reshape(a * d, d * b)) < 1e-12
            iso = Ad.conj().T @ Ad if mode == 0 else Bd @ Bd.conj().T
            assert np.abs(iso - np.eye(D)).max() < 1e-12


def test_svd_split_matches_oracle(K, rng):
    for (a, d, b), tol, mb in [((8, 2, 8), 1e-3, 100), ((30, 2, 30), 1e-10, 20), ((1, 2, 1), 1e-16, 5)]:
        AA = crandn(rng, a, d, d, b)
        AA /= np.linalg.norm(AA)
        st = ocore.Strategy(tolerance=tol, max_bond_dimension=mb)
        for mode, fn in [(0, ocore.left_orth_2site), (1, ocore.right_orth_2site)]:
            Ar, Br, er = fn(AA, st)
            Ad, Bd, D, err2, _ = K.svd_split(K.as_dev(AA), a * d, d * b, mode, st.method,
                                             st.tolerance, st.max_bond_dimension)
            assert D == Ar.shape[-1]
            assert abs(np.sqrt(err2) - er) <= 1e-12 * max(1.0, er)
            rec = K.to_numpy(Ad) @ K.to_numpy(Bd)
            assert rel(rec, Ar.reshape(a * d, -1) @ Br.reshape(D, -1)) < 1e-12


def test_two_site_matvec_and_eigensolve(K, rng):
    from oracle import algorithms as oal, builders as obl

    N, d = 6, 2
    H = obl.xxz_mpo(N)
    psi = obl.random_mps([d] * N, D=7, complex=True, rng=rng)
    T = psi.tensors
    # physical environments around the bond (2, 3)
    L = np.ones((1, 1, 1))
    for i in range(2):
        L = oal.update_left_mpo_environment(L, T[i], H[i], T[i])
    R = np.ones((1, 1, 1))
    for i in range(N - 1, 3, -1):
        R = oal.update_right_mpo_environment(R, T[i], H[i], T[i])
    O1, O2 = H[2], H[3]
    v = np.tensordot(T[2], T[3], axes=1)
    H12 = np.tensordot(O1, O2, axes=1)
    ref = oal.two_site_matvec(L, H12, R, v)
    Hm = K.two_site_operator(K.as_dev(O1), K.as_dev(O2))
    out = K.two_site_matvec(K.as_dev(L), Hm, K.as_dev(R), K.as_dev(v), d, d)
    assert rel(K.to_numpy(out), ref) < 1e-13
    n = v.size
    Hd = np.zeros((n, n), complex)
    for k in range(n):
        ek = np.zeros(n, complex)
        ek[k] = 1
        Hd[:, k] = oal.two_site_matvec(L, H12, R, ek.reshape(v.shape)).reshape(-1)
    assert np.abs(Hd - Hd.conj().T).max() < 1e-14 * np.abs(Hd).max()
    w, V = np.linalg.eigh(Hd)
    lam, vec, stats = K.eigensolve(K.as_dev(L), Hm, K.as_dev(R), K.as_dev(v), d, d, ncv=20,
                                   tol=1e-12)
    assert abs(lam - w[0]) < 1e-12 * max(1, abs(w[0]))
    x = K.to_numpy(vec).reshape(-1)
    assert abs(abs(np.vdot(V[:, 0], x)) - 1) < 1e-8
    # mixed precision: from a random start the first cycle runs FP32 products
    # after its first (FP64) one; the same FP64 stopping test at the end
    lam_m, vec_m, st_m = K.eigensolve(K.as_dev(L), Hm, K.as_dev(R), K.as_dev(v), d, d, ncv=20,
                                      tol=1e-12, precision="mixed")
    assert st_m["matvecs_fp32"] == 19
    assert st_m["matvecs"] > 20
    assert abs(lam_m - w[0]) < 1e-9 * max(1, abs(w[0]))
    assert abs(abs(np.vdot(V[:, 0], K.to_numpy(vec_m).reshape(-1))) - 1) < 1e-8
    # FP32 throughout: eigenvalue to FP32 accuracy only
    lam_s, _, st_s = K.eigensolve(K.as_dev(L), Hm, K.as_dev(R), K.as_dev(v), d, d, ncv=20,
                                  tol=1e-4, precision="fp32")
    assert st_s["matvecs_fp32"] == st_s["matvecs"]
    assert abs(lam_s - w[0]) < 1e-5 * max(1, abs(w[0]))
    # started from the eigenvector, mixed stays in FP64
    _, _, st_e = K.eigensolve(K.as_dev(L), Hm, K.as_dev(R), vec, d, d, ncv=20, tol=1e-12,
                              precision="mixed")
    assert st_e["matvecs_fp32"] == 0


def test_eigensolve_early_exit_in_cycle(K, rng):
    """Operators of dimension >= 8192 test the Ritz residual inside a cycle: a
    warm start stops after fewer than ncv products, at the same accuracy as the
    dense eigensolver."""
    from oracle import builders as obl

    H = obl.xxz_mpo(6)
    O1, O2 = H[2], H[3]
    c, w2, d = O1.shape[0], O2.shape[3], 2
    chi = 48  # n = 48 * 4 * 48 = 9216
    # Hermitian environments: L[a, c, b] = conj(L[b, c, a]) (R alike)
    L = crandn(rng, chi, c, chi)
    L = 0.5 * (L + L.conj().transpose(2, 1, 0))
    R = crandn(rng, chi, w2, chi)
    R = 0.5 * (R + R.conj().transpose(2, 1, 0))
    Hm = K.two_site_operator(K.as_dev(O1), K.as_dev(O2))
    v = crandn(rng, chi, d, d, chi)
    Ld, Rd = K.as_dev(L), K.as_dev(R)
    lam0, vec0, st0 = K.eigensolve(Ld, Hm, Rd, K.as_dev(v), d, d, ncv=20, tol=1e-12,
                                   maxiter=400)
    # the reference value: scipy's eigsh on the same operator (CPU)
    from scipy.sparse.linalg import LinearOperator, eigsh
    from oracle import algorithms as oal

    H12 = np.tensordot(O1, O2, axes=1)
    n = v.size
    op = LinearOperator((n, n), dtype=complex,
                        matvec=lambda x: oal.two_site_matvec(L, H12, R, x.reshape(v.shape)).ravel())
    ref = eigsh(op, k=1, which="SA", tol=1e-13, ncv=40, maxiter=10000)[0][0]
    assert abs(lam0 - ref) < 1e-9 * abs(ref)
    # started from the converged vector: an in-cycle test (after 4, 6, 8, ...
    # products) stops it well before ncv (random environments leave a tiny
    # spectral gap, so only a start this close converges within one cycle)
    lam1, _, st1 = K.eigensolve(Ld, Hm, Rd, vec0, d, d, ncv=20, tol=1e-10)
    assert st1["restarts"] == 1 and st1["matvecs"] <= 12 and st1["matvecs"] % 2 == 0
    assert abs(lam1 - ref) < 1e-9 * abs(ref)
    # below the size threshold the same start runs the whole cycle
    small = slice(0, 40)
    Ls, Rs = L[small, :, small], R[small, :, small]  # n = 40 * 4 * 40 = 6400
    lam_s, vec_s, _ = K.eigensolve(K.as_dev(Ls), Hm, K.as_dev(Rs),
                                   K.as_dev(v[small, :, :, small]), d, d, ncv=20, tol=1e-12,
                                   maxiter=400)
    _, _, st_s = K.eigensolve(K.as_dev(Ls), Hm, K.as_dev(Rs), vec_s, d, d, ncv=20, tol=1e-10)
    assert st_s["matvecs"] == 20


@pytest.mark.parametrize("chiL,chiR", [(64, 64), (96, 40), (256, 256)])
def test_two_site_matvec_splitk_shapes(K, rng, chiL, chiR):
    """Shapes whose output GEMMs take the split-K path (few output tiles, long K)."""
    from oracle import algorithms as oal, builders as obl

    H = obl.xxz_mpo(6)
    O1, O2 = H[2], H[3]
    c, w2, d = O1.shape[0], O2.shape[3], 2
    L = crandn(rng, chiL, c, chiL)
    R = crandn(rng, chiR, w2, chiR)
    v = crandn(rng, chiL, d, d, chiR)
    ref = oal.two_site_matvec(L, np.tensordot(O1, O2, axes=1), R, v)
    Hm = K.two_site_operator(K.as_dev(O1), K.as_dev(O2))
    out = K.two_site_matvec(K.as_dev(L), Hm, K.as_dev(R), K.as_dev(v), d, d)
    assert rel(K.to_numpy(out), ref) < 1e-13


def _decaying(rng, m, n, rate):
    U0, _ = np.linalg.qr(crandn(rng, m, min(m, n)))
    V0, _ = np.linalg.qr(crandn(rng, n, min(m, n)))
    s0 = np.exp(-rate * np.arange(min(m, n)))
    return (U0 * s0) @ V0.conj().T, s0


def _capped_certification(s, D, l):
    """(certified with a 10x margin, refused) for a sketch of width l deciding a
    cap D: the unsketched weight (at least the true tail beyond l) must be
    <= max(1e-14 ||A||^2, 1e-3 sigma_D^2)."""
    tail = float(np.sum(s[l:] ** 2))
    thr = max(1e-14 * float(np.sum(s ** 2)), 1e-3 * float(s[D - 1]) ** 2)
    return 10 * tail <= thr, tail > thr


def _sketch_bound(s, D, l, q):
    """Randomized-range accuracy of the kept rank-D part (relative to ||A||):
    ~ sigma_l (sigma_l / sigma_(D-1))^(2q), with a safety factor of 10."""
    if l >= s.size:
        return 1e-12
    return max(1e-9, 10 * s[l] / s[0] * (s[l] / s[D - 1]) ** (2 * q))


@pytest.mark.parametrize("m,n,rate,tol,mb", [(128, 128, 0.15, 1e-10, 1000), (256, 192, 0.1, 1e-8, 64),
                                             (96, 160, 0.3, 1e-12, 1000)])
def test_svd_split_sketch_matches_full(K, rng, m, n, rate, tol, mb):
    """Sketched split: same D as the full split when the sketch holds the kept rank,
    the kept part within the randomized-range bound (1e-9 when the spectrum
    has decayed past the sketch), fallback when it does not hold the rank, and
    the bond cap decided from the sketch alone."""
    A, sv = _decaying(rng, m, n, rate)
    st = ocore.Strategy(tolerance=tol, max_bond_dimension=mb)
    Ar, Br, er = ocore.left_orth_2site(A.reshape(m, 1, 1, n), st)
    Dref = Ar.shape[-1]
    l = min(Dref + 16, min(m, n) - 1)
    for mode in (0, 1):
        # accuracy of the kept part ~ sigma_l (sigma_l / sigma_D)^(2q): q = 2 reaches 1e-9 here
        Ad, Bd, D, err2, _, fb = K.svd_split_sketch(K.as_dev(A), m, n, mode, st.method,
                                                    st.tolerance, st.max_bond_dimension, l,
                                                    power_iters=2, seed=7)
        if Dref == mb:
            certified, refused = _capped_certification(sv, Dref, l)
            assert fb if refused else (not fb if certified else True)
            if fb:
                assert D == Dref  # the full split decided
                continue
        assert not fb
        assert D == Dref
        # the unsketched weight is ||A||^2 - sum s^2: an estimate, good to ~1e-3 here
        assert abs(np.sqrt(err2) - er) <= 1e-3 * max(er, 1e-12) + 1e-14
        rec = K.to_numpy(Ad) @ K.to_numpy(Bd)
        assert rel(rec, Ar.reshape(m, -1) @ Br.reshape(D, -1)) < _sketch_bound(sv, D, l, 2)
        iso = K.to_numpy(Ad).conj().T @ K.to_numpy(Ad) if mode == 0 else \
            K.to_numpy(Bd) @ K.to_numpy(Bd).conj().T
        assert np.abs(iso - np.eye(D)).max() < 1e-12
    # a sketch far below the kept rank must fall back to the full split
    Ad, Bd, D, err2, _, fb = K.svd_split_sketch(K.as_dev(A), m, n, 0, st.method, st.tolerance,
                                                st.max_bond_dimension, max(2, Dref // 3),
                                                power_iters=1, seed=7)
    assert fb and D == Dref
    # bond cap below the sketch size: decided without the full SVD
    cap = max(2, Dref // 2)
    stc = ocore.Strategy(tolerance=tol, max_bond_dimension=cap)
    Ar, Br, er = ocore.left_orth_2site(A.reshape(m, 1, 1, n), stc)
    # (the randomized range loses accuracy as sigma_l / sigma_cap grows: a wider
    # sketch and two power iterations keep this case at 1e-8)
    Ad, Bd, D, err2, _, fb = K.svd_split_sketch(K.as_dev(A), m, n, 0, stc.method, stc.tolerance,
                                                cap, min(cap + 32, min(m, n) - 1), power_iters=2,
                                                seed=3)
    lc = min(cap + 32, min(m, n) - 1)
    assert D == cap == Ar.shape[-1]
    certified, refused = _capped_certification(sv, cap, lc)
    if refused:
        assert fb  # the unseen weight may compete with the kept values: full SVD
    elif certified:
        assert not fb
    assert rel(K.to_numpy(Ad) @ K.to_numpy(Bd), Ar.reshape(m, -1) @ Br.reshape(cap, -1)) < \
        (1e-12 if fb else _sketch_bound(sv, cap, lc, 2))


def _in_span(rng, m, n, k, side, rate=0.05):
    """A (m x n) whose row (side 0) or column (side 1) space is spanned by a
    k-dimensional isometry, and that isometry in the layout the split takes."""
    g = np.exp(-rate * np.arange(k))
    if side == 0:
        Q, _ = np.linalg.qr(crandn(rng, n, k))
        basis = Q.conj().T                       # k x n, orthonormal rows
        A = (crandn(rng, m, k) * g) @ basis
    else:
        basis, _ = np.linalg.qr(crandn(rng, m, k))  # m x k, orthonormal columns
        A = basis @ (crandn(rng, k, n) * g[:, None])
    return A, basis


def _outside(rng, basis, side, m, n, r, scale):
    """Rank-r term whose row (side 0) / column (side 1) space is orthogonal to the basis."""
    if side == 0:
        X = crandn(rng, n, r)
        X -= basis.conj().T @ (basis @ X)
        return scale * crandn(rng, m, r) @ np.linalg.qr(X)[0].conj().T
    X = crandn(rng, m, r)
    X -= basis @ (basis.conj().T @ X)
    return scale * np.linalg.qr(X)[0] @ crandn(rng, r, n)


def _check_split(K, A, Ad, Bd, D, mode, Ar, Br, tol_rec):
    assert D == Ar.shape[-1]
    a, b = K.to_numpy(Ad), K.to_numpy(Bd)
    iso = a.conj().T @ a if mode == 0 else b @ b.conj().T
    assert np.abs(iso - np.eye(D)).max() < 1e-12
    ref = Ar.reshape(A.shape[0], -1) @ Br.reshape(D, -1)
    assert rel(a @ b, ref) < tol_rec


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_svd_split_basis_exact_span(K, rng, side, mode):
    """A inside the basis span: the basis-sketched split equals the full split
    (same D, reconstruction and isometry to rounding), with no fallback, also
    though the expansion block is then pure rounding noise."""
    m, n, k = 192, 160, 96
    A, basis = _in_span(rng, m, n, k, side)
    st = ocore.Strategy(tolerance=1e-12, max_bond_dimension=1000)
    Ar, Br, er = ocore.left_orth_2site(A.reshape(m, 1, 1, n), st)
    Ad, Bd, D, err2, _, fb = K.svd_split_basis(K.as_dev(A), m, n, side, K.as_dev(basis), k, 16,
                                               mode, st.method, st.tolerance,
                                               st.max_bond_dimension, seed=5)
    assert not fb
    _check_split(K, A, Ad, Bd, D, mode, Ar, Br, 1e-12)


@pytest.mark.parametrize("side", [0, 1])
def test_svd_split_basis_grows_and_caps(K, rng, side):
    """A rank-4 term outside the basis: the expansion directions let the bond
    grow to k + 4 exactly as the full split does; with the bond capped at k the
    sketch decides alone and keeps the best k directions within the 1e-9 bound;
    a generic perturbation above the tolerance falls back to the full split."""
    m, n, k = 256, 224, 100
    A0, basis = _in_span(rng, m, n, k, side, rate=0.02)
    A = A0 + _outside(rng, basis, side, m, n, 4, 0.3)
    mode = side
    st = ocore.Strategy(tolerance=1e-12, max_bond_dimension=1000)
    Ar, Br, er = ocore.left_orth_2site(A.reshape(m, 1, 1, n), st)
    assert Ar.shape[-1] == k + 4
    Ad, Bd, D, err2, _, fb = K.svd_split_basis(K.as_dev(A), m, n, side, K.as_dev(basis), k, 16,
                                               mode, st.method, st.tolerance,
                                               st.max_bond_dimension, seed=9)
    assert not fb
    _check_split(K, A, Ad, Bd, D, mode, Ar, Br, 1e-11)
    # capped at k: the sketch (k + 16 columns) holds the whole matrix, so the
    # capped split is the full one's
    stc = ocore.Strategy(tolerance=1e-12, max_bond_dimension=k)
    Ar, Br, er = ocore.left_orth_2site(A.reshape(m, 1, 1, n), stc)
    Ad, Bd, D, err2, _, fb = K.svd_split_basis(K.as_dev(A), m, n, side, K.as_dev(basis), k, 16,
                                               mode, stc.method, stc.tolerance, k, seed=9)
    assert not fb and D == k
    _check_split(K, A, Ad, Bd, D, mode, Ar, Br, 1e-11)
    assert abs(np.sqrt(err2) - er) < 1e-9 * np.linalg.norm(A)
    # generic noise outside the span, far above the tolerance: only the full split decides
    An = A0 + 1e-4 * crandn(rng, m, n)
    Ar, Br, er = ocore.left_orth_2site(An.reshape(m, 1, 1, n), st)
    Ad, Bd, D, err2, _, fb = K.svd_split_basis(K.as_dev(An), m, n, side, K.as_dev(basis), k, 16,
                                               mode, st.method, st.tolerance,
                                               st.max_bond_dimension, seed=9)
    assert fb
    _check_split(K, An, Ad, Bd, D, mode, Ar, Br, 1e-12)


@pytest.mark.parametrize("extra", [0, 16])
def test_svd_split_basis_unrelated_basis(K, rng, extra):
    """A basis unrelated to A's row space (random isometry): the sketched
    matrix's columns are far from orthogonal, so the rotation iteration gives
    up (or the block Jacobi finishes it), the truncation cannot certify the
    sketch, and the full split decides: the result is the oracle's split."""
    m, n, k = 300, 280, 120
    sv = np.exp(-0.05 * np.arange(min(m, n)))
    U, _ = np.linalg.qr(crandn(rng, m, m))
    V, _ = np.linalg.qr(crandn(rng, n, n))
    A = (U[:, : len(sv)] * sv) @ V[:, : len(sv)].conj().T
    Q, _ = np.linalg.qr(crandn(rng, n, k))
    basis = Q.conj().T.copy()  # k x n, orthonormal rows, unrelated to A
    st = ocore.Strategy(tolerance=1e-10, max_bond_dimension=k)
    Ar, Br, er = ocore.left_orth_2site(A.reshape(m, 1, 1, n), st)
    Ad, Bd, D, err2, _, fb = K.svd_split_basis(K.as_dev(A), m, n, 0, K.as_dev(basis), k, extra,
                                               0, st.method, st.tolerance,
                                               st.max_bond_dimension, seed=3)
    assert fb
    _check_split(K, A, Ad, Bd, D, 0, Ar, Br, 1e-12)
    assert abs(np.sqrt(err2) - er) < 1e-12 * np.linalg.norm(A)


def test_svd_split_basis_too_wide_falls_back(K, rng):
    m, n, k = 64, 48, 40
    A, basis = _in_span(rng, m, n, k, 0)
    st = ocore.Strategy(tolerance=1e-12, max_bond_dimension=1000)
    Ar, Br, er = ocore.left_orth_2site(A.reshape(m, 1, 1, n), st)
    Ad, Bd, D, err2, _, fb = K.svd_split_basis(K.as_dev(A), m, n, 0, K.as_dev(basis), k, 16,
                                               0, st.method, st.tolerance, 1000)
    assert fb  # k + p > min(m, n)
    _check_split(K, A, Ad, Bd, D, 0, Ar, Br, 1e-12)
