"""Host-side logic of the drop-in API on CPU, with the kernels replaced by the
NumPy test double (tests/fake_kernels.py).  Checks control flow, shapes and
error bookkeeping against the oracle; the device kernels themselves are
covered by the gpu-marked tests."""

import numpy as np
import pytest

from oracle import core as oc, algorithms as oa, builders as ob, exact as ex


@pytest.fixture(scope="module")
def S():
    import seemps2_amd
    from seemps2_amd import kernels
    from tests import fake_kernels as fk

    saved = {}
    for name in dir(fk):
        if not name.startswith("_") and hasattr(kernels, name) and callable(getattr(fk, name)):
            saved[name] = getattr(kernels, name)
            setattr(kernels, name, getattr(fk, name))
    yield seemps2_amd
    for name, f in saved.items():
        setattr(kernels, name, f)


def dev(S, o):
    return S.MPS(o.tensors, error=o.error)


def test_canonical_and_recenter(S, rng):
    a = ob.random_mps([2] * 7, D=5, complex=True, rng=rng)
    for c in (0, 3, 6):
        r = oa.canonical(a, c, oc.Strategy(max_bond_dimension=3))
        o = S.CanonicalMPS(dev(S, a), center=c, strategy=S.Strategy(max_bond_dimension=3))
        assert o.bond_dimensions() == r.bond_dimensions()
        assert abs(o.error() - r.error) < 1e-10
        assert np.allclose(o.to_vector(), r.to_vector(), atol=1e-12)
    o.recenter(1)
    assert o.center == 1


def test_apply_simplify(S, rng):
    N = 8
    H = ob.heisenberg_mpo(N)
    a = ob.random_mps([2] * N, D=6, complex=True, rng=rng)
    ost = oc.Strategy(tolerance=1e-10, simplification_tolerance=1e-10, max_bond_dimension=8,
                      max_sweeps=2)
    st = S.Strategy(tolerance=1e-10, simplification_tolerance=1e-10, max_bond_dimension=8,
                    max_sweeps=2)
    r = oa.mpo_apply(H, a, ost, simplify_flag=True)
    o = S.MPO(H).apply(dev(S, a), strategy=st)
    assert o.bond_dimensions() == r.bond_dimensions()
    assert abs(o.error() - r.error) < 1e-10 * max(1, r.error)
    assert np.allclose(o.to_vector(), r.to_vector(), atol=1e-10)
    e = oa.mpo_expectation(H, a)
    assert abs(S.MPO(H).expectation(dev(S, a)) - e) < 1e-12 * abs(e)


def test_simplify_right_to_left(S, rng):
    """simplify(direction=-1) starts at the last site (the reference passes
    start = -1 to AntilinearForm, meant as the last site)."""
    N = 8
    a = ob.random_mps([2] * N, D=6, complex=True, rng=rng)
    b = ob.random_mps([2] * N, D=5, complex=True, rng=rng)
    ost = oc.Strategy(tolerance=1e-10, simplification_tolerance=1e-10, max_bond_dimension=7)
    st = S.Strategy(tolerance=1e-10, simplification_tolerance=1e-10, max_bond_dimension=7)
    r = oa.simplify(oa.OMPSSum([1.0, 0.5], [a, b]), ost, direction=-1)
    o = S.simplify(S.MPSSum([1.0, 0.5], [dev(S, a), dev(S, b)]), st, direction=-1)
    assert o.bond_dimensions() == r.bond_dimensions()
    assert np.allclose(o.to_vector(), r.to_vector(), atol=1e-10)
    r1 = oa.simplify(a, ost, direction=-1)
    o1 = S.simplify(dev(S, a), st, direction=-1)
    assert o1.bond_dimensions() == r1.bond_dimensions()
    assert np.allclose(o1.to_vector(), r1.to_vector(), atol=1e-10)


def test_simplify_sum_and_ops(S, rng):
    a = ob.random_mps([2] * 6, D=3, complex=True, rng=rng)
    b = ob.random_mps([2] * 6, D=4, complex=True, rng=rng)
    st = S.Strategy(tolerance=1e-9, simplification_tolerance=1e-9, normalize=True)
    r = oa.simplify(oa.OMPSSum([2.0, -1j], [a, b]), oc.Strategy(tolerance=1e-9,
                    simplification_tolerance=1e-9, normalize=True))
    s = 2.0 * dev(S, a) - 1j * dev(S, b)
    o = S.simplify(s, st)
    assert o.bond_dimensions() == r.bond_dimensions()
    assert np.allclose(o.to_vector(), r.to_vector(), atol=1e-10)
    assert abs(o.error() - r.error) < 1e-10
    assert np.allclose(s.join().to_vector(), oa.join(oa.OMPSSum([2.0, -1j], [a, b])).to_vector())


def test_dmrg_host_logic(S):
    N = 6
    rng = np.random.default_rng(4)
    g = ob.random_mps([2] * N, D=2, complex=True, rng=rng)
    H = ob.xxz_mpo(N)
    r = oa.dmrg(H, g.copy(), maxiter=6)
    o = S.dmrg(S.MPO(H), dev(S, g), maxiter=6)
    assert abs(o.energy - r.energy) < 1e-10
    assert len(o.trajectory) == len(o.variances)
    with pytest.raises(ValueError):
        S.dmrg(S.MPO(H), dev(S, g), maxiter=0)
    with pytest.raises(ValueError):
        S.dmrg(S.MPO(H), dev(S, g), tol=-1.0)


def test_strategy_api(S):
    st = S.Strategy(method=S.Truncation.RELATIVE_SINGULAR_VALUE, tolerance=0.0)
    assert st.get_method() == S.Truncation.ABSOLUTE_SINGULAR_VALUE
    assert S.NO_TRUNCATION.get_method() == 0 and not S.NO_TRUNCATION.get_simplify_flag()
    assert S.DEFAULT_STRATEGY.replace(max_bond_dimension=5).get_max_bond_dimension() == 5
    with pytest.raises(AssertionError):
        S.Strategy(tolerance=2.0)
    assert "RelativeNorm" in str(S.DEFAULT_STRATEGY)


def test_learned_truncation_backs_off_after_fallbacks(monkeypatch):
    """Sketched splits that keep falling back pause the sketching: after
    `fallback_streak` fallbacks in a row the next `backoff` eligible splits go
    to the full SVD, the pause doubling with every further streak and
    resetting after a sketched split that holds."""
    import torch
    from seemps2_amd import learned

    calls = []

    def full(AA, m, n, mode, *a):
        calls.append("full")
        return None, None, 1, 0.0, None

    outcome = {"fb": True}

    def sketch(AA, m, n, mode, *a, **k):
        calls.append("sketch")
        return None, None, 1, 0.0, None, outcome["fb"]

    class Half:
        def predict(self, X):
            return torch.full((X.shape[0],), 0.5)

    monkeypatch.setattr(learned.K, "svd_split", full)
    monkeypatch.setattr(learned.K, "svd_split_sketch", sketch)
    monkeypatch.setattr(learned, "spectral_features", lambda *a, **k: torch.zeros(32))
    lt = learned.LearnedTruncation(Half(), min_size=8, fallback_streak=2, backoff=3)
    strategy = type("St", (), {"method": 2, "tolerance": 1e-8, "max_bond_dimension": 100})()
    AA = torch.zeros(1)

    def run(k):
        for _ in range(k):
            lt.split(AA, 100, 100, 0, strategy)

    run(2)                       # two fallbacks: streak reached
    assert calls == ["sketch", "sketch"]
    run(3)                       # paused for 3
    assert calls[2:] == ["full"] * 3
    run(2)                       # two more fallbacks: pause doubles to 6
    run(6)
    assert calls[5:] == ["sketch", "sketch"] + ["full"] * 6
    outcome["fb"] = False
    run(1)                       # a sketch that holds resets the pause
    outcome["fb"] = True
    run(2)
    run(3)
    assert calls[13:] == ["sketch", "sketch", "sketch"] + ["full"] * 3
    assert lt.fallbacks == 6 and lt.full_splits == 12
    # splits below min_size do not consume the pause
    lt._skip = 2
    lt.split(AA, 4, 4, 0, strategy)
    assert lt._skip == 2
