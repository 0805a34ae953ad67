"""Pin the oracle's core (truncation, SVD split, canonical form, environments).

Checks: the reference's own known-answer tests (golden JSON), a brute-force
restatement of each truncation rule in exact arithmetic, and dense closed forms.
"""

import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import core, algorithms as al, builders as bl

GOLDEN = json.loads((Path(__file__).parent / "golden" / "truncation_known_answers.json").read_text())


@pytest.mark.parametrize("case", GOLDEN["cases"], ids=lambda c: c["source"])
def test_truncation_known_answers(case):
    st = core.Strategy(method=case["method"], tolerance=case["tolerance"],
                       max_bond_dimension=case["max_bond"])
    D, err = core.truncate(np.array(case["s"]), st)
    assert D == case["expected_size"]
    if case["exact"]:
        assert err == case["expected_err"]
    else:
        assert abs(err - case["expected_err"]) < 1e-7  # assertAlmostEqual in the reference


@pytest.mark.parametrize("case", GOLDEN["method_changes"], ids=lambda c: c["source"])
def test_zero_tolerance_changes_method(case):
    assert core.Strategy(method=case["method"], tolerance=case["tolerance"]).method == \
        case["expected_method"]


def test_strategy_validation():
    for kw in [dict(tolerance=-1.0), dict(tolerance=1.0), dict(max_bond_dimension=0),
               dict(max_sweeps=-1), dict(simplify=7), dict(method=9)]:
        with pytest.raises(AssertionError):
            core.Strategy(**kw)


def _brute_force(s, st):
    """Exact-arithmetic statement of the rules (independent formulation)."""
    N = len(s)
    if st.method == 0:
        return N
    if st.method == 2:
        sq = [Fraction(float(x)) ** 2 for x in s]
        total = sum(sq)
        tol = Fraction(st.tolerance)
        # largest number of trailing values whose squares sum to <= tol * total, keep >= 1
        drop = 0
        for k in range(1, N):
            if sum(sq[N - k:]) <= tol * total:
                drop = k
        return min(N - drop, st.max_bond_dimension)
    thr = st.tolerance * s[0] if st.method == 1 else st.tolerance
    D = min(N, st.max_bond_dimension)
    for i in range(1, D):
        if s[i] <= thr:
            return i
    return D


def test_truncation_rules_vs_bruteforce(rng):
    for method in (1, 2, 3):
        for _ in range(200):
            N = int(rng.integers(1, 40))
            s = np.sort(np.exp(rng.normal(scale=4.0, size=N)))[::-1]
            st = core.Strategy(method=method, tolerance=float(10 ** rng.uniform(-12, -0.5)),
                               max_bond_dimension=int(rng.integers(1, 45)))
            D, err = core.truncate(s, st)
            assert D == _brute_force(s, st)
            assert abs(err - float(sum(Fraction(float(x)) ** 2 for x in s[D:]))) <= 1e-14 * max(
                1.0, float(np.sum(s ** 2)))


def crandn(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def test_orth_2site_split(rng):
    AA = crandn(rng, 5, 2, 3, 4)
    for fn, left in [(core.left_orth_2site, True), (core.right_orth_2site, False)]:
        A, B, err = fn(AA, core.NO_TRUNCATION)
        assert err == 0.0
        assert np.allclose(np.tensordot(A, B, axes=1), AA, atol=1e-13)
        if left:
            M = A.reshape(-1, A.shape[-1])
            assert np.allclose(M.conj().T @ M, np.eye(M.shape[1]), atol=1e-13)
        else:
            M = B.reshape(B.shape[0], -1)
            assert np.allclose(M @ M.conj().T, np.eye(M.shape[0]), atol=1e-13)
    A, B, err = core.left_orth_2site(AA, core.Strategy(max_bond_dimension=3))
    s = np.linalg.svd(AA.reshape(10, 12), compute_uv=False)
    assert A.shape[-1] == 3 and abs(err - np.sqrt(np.sum(s[3:] ** 2))) < 1e-12


def test_canonicalize_preserves_state(rng):
    psi = bl.random_mps([2] * 7, D=5, complex=True, rng=rng)
    v = psi.to_vector()
    for center in (0, 3, 6):
        st = list(psi.tensors)
        err = core.canonicalize(st, center, core.DEFAULT_STRATEGY)
        w = al.OMPS(st).to_vector()
        assert np.linalg.norm(w - v) < 1e-12 * np.linalg.norm(v)
        assert err < 1e-7
        for i, A in enumerate(st):
            if i < center:
                M = A.reshape(-1, A.shape[-1])
                assert np.allclose(M.conj().T @ M, np.eye(M.shape[1]), atol=1e-12)
            elif i > center:
                M = A.reshape(A.shape[0], -1)
                assert np.allclose(M @ M.conj().T, np.eye(M.shape[0]), atol=1e-12)


def test_scprod_and_environments_dense(rng):
    a = bl.random_mps([2, 3, 2, 2], D=4, complex=True, rng=rng)
    b = bl.random_mps([2, 3, 2, 2], D=3, complex=True, rng=rng)
    assert abs(core.scprod(a.tensors, b.tensors) - np.vdot(a.to_vector(), b.to_vector())) < 1e-12
    # left/right environments joined at every cut give the same scalar product
    for cut in range(1, 4):
        L = np.eye(1)
        for i in range(cut):
            L = core.update_left_environment(a.tensors[i], b.tensors[i], L)
        R = np.eye(1)
        for i in range(3, cut - 1, -1):
            R = core.update_right_environment(a.tensors[i], b.tensors[i], R)
        assert abs(core.join_environments(L, R) - core.scprod(a.tensors, b.tensors)) < 1e-12


def test_two_site_gate_dense(rng):
    A = crandn(rng, 3, 2, 4)
    B = crandn(rng, 4, 2, 5)
    U = crandn(rng, 4, 4)
    out = core.contract_nrjl_ijk_klm(U, A, B)
    theta = np.tensordot(A, B, axes=1).reshape(3, 4, 5)
    ref = np.einsum("xy,ayb->axb", U, theta).reshape(3, 2, 2, 5)
    assert np.allclose(out, ref, atol=1e-13)
