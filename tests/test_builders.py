"""Function-loading input builders on the CPU (kernels replaced by the NumPy
test double), against closed forms (tests/builder_checks.py)."""

import pytest

from tests import builder_checks as bc


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


def test_interleaving(S):
    bc.check_interleaving(S)


def test_intervals(S):
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


def test_trotter(S, rng):
    bc.check_trotter(S, rng)


def test_gate_layer(S, rng):
    bc.check_gate_layer(S, rng)


def test_expmv(S, rng):
    bc.check_expmv(S, rng)


def test_tdvp(S, rng):
    bc.check_tdvp(S, rng)


def test_dmrg_solve(S, rng):
    bc.check_dmrg_solve(S, rng)
