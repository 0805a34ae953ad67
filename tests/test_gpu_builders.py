"""Function-loading input builders and the Euler driver on the device
(native kernels), against closed forms (tests/builder_checks.py), and a
reduced function-loading workload against the oracle's MPO.apply + simplify."""

import numpy as np
import pytest

from tests import builder_checks as bc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def S():
    import seemps2_amd

    return seemps2_amd


def test_interleaving_and_intervals(S):
    bc.check_interleaving(S)
    bc.check_intervals(S)


def test_shifts(S, rng):
    bc.check_shifts(S, rng)


def test_finite_differences(S):
    bc.check_finite_differences(S)


def test_tridiagonal(S):
    bc.check_tridiagonal(S)


def test_extend_and_join(S, rng):
    bc.check_extend_and_join(S, rng)


def test_tensor_product(S, rng):
    bc.check_tensor_product(S, rng)


def test_euler(S, rng):
    bc.check_euler(S, rng)


def test_function_loading_laplacian_euler(S):
    """f(x, y) = exp(-(x^2 + y^2)/0.1) cos(20 x) on [-1, 1)^2 with 5 + 5 qubits
    in "B" order, the 2-D Laplacian from extended 1-D finite differences joined
    into one MPO, Euler steps: against the dense closed form and against the
    oracle's MPO.apply + simplify on the same MPO and state."""
    from oracle import algorithms as oa, core as oc

    n = 5
    I = S.QuantizedInterval(-1.0, 1.0, n)
    x = I.to_vector()
    fx = np.exp(-x ** 2 / 0.1) * np.cos(20 * x)
    fy = np.exp(-x ** 2 / 0.1)
    st = S.Strategy(tolerance=1e-12, simplification_tolerance=1e-12, max_bond_dimension=64)
    psi = S.mps_tensor_product([S.MPS.from_vector(fx, [2] * n, normalize=False),
                                S.MPS.from_vector(fy, [2] * n, normalize=False)], "B")
    perm = S.interleaving_permutation([n, n])
    D1 = S.finite_differences_mpo(2, I)
    # "B" order puts x's sites at even and y's at odd positions
    Lx = D1.extend(2 * n, sites=[2 * k for k in range(n)])
    Ly = D1.extend(2 * n, sites=[2 * k + 1 for k in range(n)])
    Lap = (Lx + Ly).join(strategy=st)
    dense = Lap.to_matrix()
    # the dense operator is the Kronecker sum in the interleaved site order
    d1 = D1.to_matrix()
    ksum = (np.kron(d1, np.eye(2 ** n)) + np.kron(np.eye(2 ** n), d1)).reshape([2] * (4 * n))
    axes = list(perm) + [2 * n + p for p in perm]
    assert np.abs(dense - ksum.transpose(axes).reshape(4 ** n, 4 ** n)).max() < 1e-13 * np.abs(dense).max()
    f = np.kron(fx, fy).reshape([2] * (2 * n)).transpose(perm).reshape(-1)
    assert np.abs(psi.to_vector() - f).max() < 1e-12
    dt, steps = 1e-6, 5
    out = S.euler(Lap, (0.0, dt * steps), psi, steps=steps, strategy=st)
    ref = f.astype(complex)
    for _ in range(steps):
        ref = ref + dt * (dense @ ref)
    err = np.linalg.norm(out.to_vector() - ref) / np.linalg.norm(ref)
    assert err < 1e-5  # sqrt(simplification tolerance) per step
    # the oracle's MPO.apply + simplify on the same inputs
    ost = oc.Strategy(tolerance=1e-12, simplification_tolerance=1e-12, max_bond_dimension=64)
    W = [A.cpu().numpy() for A in Lap]
    o = oa.OMPS([A.cpu().numpy() for A in psi], 0.0)
    for _ in range(steps):
        Lo = oa.mpo_apply(W, o, ost, simplify_flag=True)
        o = oa.simplify(oa.OMPSSum([1.0, dt], [o, Lo]), ost)
    ov = o.to_vector()
    assert np.linalg.norm(out.to_vector() - ov) / np.linalg.norm(ov) < 1e-5


def test_trotter(S, rng):
    bc.check_trotter(S, rng)


def test_gate_layer(S, rng):
    bc.check_gate_layer(S, rng)


def _haar(rng, n):
    Q, R = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    return Q * (np.diag(R) / np.abs(np.diag(R)))


def test_haar_gate_layers_learned_vs_full(S):
    """Reduced two-site-gate workload (d = 4, N = 10, product state, 3 sweeps of
    seeded Haar 16 x 16 gates, bond cap 48, tolerance 1e-8): the learned-rank
    splits (randomized sketches; the basis sketch cannot hold a gate-rotated
    row space and falls back) against plain full splits: the fallback rate is
    reported, truncation errors agree, and the final states agree to 1e-9 in
    fidelity."""
    from seemps2_amd import learned
    from seemps2_amd.evolution.trotter import _sweep_gates

    rng = np.random.default_rng(5)
    N, d = 10, 4
    layers = [[_haar(rng, d * d) for _ in range(N - 1)] for _ in range(3)]
    st = S.Strategy(tolerance=1e-8, max_bond_dimension=48)

    def run(ctx=None):
        psi = S.CanonicalMPS(S.product_state(np.eye(d)[0], N), center=0, strategy=st)
        if ctx is not None:
            with ctx:
                for gates in layers:
                    psi = _sweep_gates(psi, gates, st)
        else:
            for gates in layers:
                psi = _sweep_gates(psi, gates, st)
        return psi

    plain = run()
    X = torch.stack([learned.spectral_features(K_rand(rng, 64), 64, 64, seed=i)
                     for i in range(16)])
    net = learned.RankPredictor(seed=3)
    # a predictor that asks for narrow sketches: l = cap + 16 = 64 < min(m, n)
    net.train(X, torch.full((16,), 0.2, device=X.device), epochs=200)
    lt = learned.LearnedTruncation(net, min_size=32, oversample=16)
    out = run(learned.learned_truncation(lt))
    assert out.bond_dimensions() == plain.bond_dimensions()
    a, b = plain.to_vector(), out.to_vector()
    fid = abs(np.vdot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b))
    assert 1 - fid < 1e-9
    assert abs(out.error() - plain.error()) <= 1e-6 * max(plain.error(), 1e-12)
    # flat spectra at the cap cannot be certified from a sketch: those splits
    # fall back to the full SVD
    assert lt.calls > 0 and lt.fallbacks > 0
    print(f"fallback rate {lt.fallbacks}/{lt.calls - lt.full_splits}")


def K_rand(rng, n):
    from seemps2_amd import kernels as K

    return K.as_dev(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))


def test_expmv(S, rng):
    bc.check_expmv(S, rng)


def test_tdvp(S, rng):
    bc.check_tdvp(S, rng)


def test_dmrg_solve(S, rng):
    bc.check_dmrg_solve(S, rng)
