"""Pin the oracle's simplify bookkeeping (simplification.py:111-142, 250-280):
the per-sweep error measure, the stopping sweep and the accumulated error.

The oracle evaluates each sweep's measure 2 |1 - Re <xi|psi> / sqrt(|xi|^2 |psi|^2)|
through its environments: the overlap from the antilinear forms at the sweep's
last site and |xi|^2 from the centre tensor alone (canonical form).  Here the
same quantity is computed independently from dense vectors, as the squared
distance of the normalised vectors |xi/|xi| - psi/|psi||^2 (equal to the
measure whenever Re <xi|psi> <= |xi||psi|, i.e. always), after every sweep.
The stopping rule is then replayed on the dense values, and the final error
bound on them.  Planted mistakes in the oracle's environment bookkeeping must
make the comparison fail.
"""

import math

import numpy as np
import pytest

from oracle import core, algorithms as al, builders as bl


def _dense_measure(tensors, psi_vec):
    xi = al.OMPS(tensors).to_vector()
    a = xi / np.linalg.norm(xi)
    b = psi_vec / np.linalg.norm(psi_vec)
    return float(np.linalg.norm(a - b) ** 2)


def _run(state, psi_vec, strategy, direction):
    seen = []

    def observer(tensors, err, last):
        # the centre sits at `last`: |xi|^2 from that tensor alone equals the dense norm
        T = tensors[last]
        nrm = float(np.vdot(T, T).real)
        xi = al.OMPS(tensors).to_vector()
        seen.append((err, _dense_measure(tensors, psi_vec), nrm, float(np.vdot(xi, xi).real)))

    out = al.simplify(state, strategy, direction=direction, observer=observer)
    return out, seen


def _check(state, psi_vec, input_error, strategy, direction):
    out, seen = _run(state, psi_vec, strategy, direction)
    assert seen, "simplify ran no sweep"
    errs = [e for e, _, _, _ in seen]
    dense = [d for _, d, _, _ in seen]
    for err, d, nrm, nrm_dense in seen:
        assert abs(nrm - nrm_dense) <= 1e-12 * nrm_dense
        assert abs(err - d) <= 1e-12 + 1e-9 * d, (err, d)
    # replay the stopping rule on the dense values: stop at the first sweep
    # below the tolerance or above the previous one, else after max_sweeps
    prev, stop = 2.0, None
    for k, d in enumerate(dense):
        if d < strategy.simplification_tolerance or d > prev:
            stop = k + 1
            break
        prev = d
    if stop is None:
        stop = max(1, strategy.max_sweeps)
    assert len(errs) == stop
    assert out.sweep_errors == errs
    # accumulated bound: the input's error plus sqrt of the last sweep's
    # measure (the measure itself is pinned to the dense value above; its
    # square root is taken from the recorded value, since near zero the two
    # differ by rounding that the root amplifies)
    assert abs(out.error - (input_error + math.sqrt(errs[-1]))) <= 1e-12
    return len(errs)


@pytest.mark.parametrize("direction", [+1, -1])
def test_simplify_mps_sweep_errors_match_dense(direction):
    rng = np.random.default_rng(314 + direction)
    N = 12
    psi = bl.random_mps([2] * N, D=10, complex=True, rng=rng)
    psi.error = 0.25
    st = core.Strategy(tolerance=1e-6, simplification_tolerance=1e-14, max_bond_dimension=6,
                       max_sweeps=5)
    sweeps = _check(psi, psi.to_vector(), 0.25, st, direction)
    assert sweeps >= 2  # the capped bond keeps it sweeping: the rule is exercised


def test_simplify_sum_sweep_errors_match_dense():
    rng = np.random.default_rng(2718)
    N = 13
    a = bl.random_mps([2] * N, D=7, complex=True, rng=rng)
    b = bl.random_mps([2] * N, D=5, complex=True, rng=rng)
    a.error, b.error = 0.1, 0.2
    s = al.OMPSSum([1.0, 0.5j], [a, b])
    dense = s.to_vector()
    st = core.Strategy(tolerance=1e-8, simplification_tolerance=1e-14, max_bond_dimension=8,
                       max_sweeps=4)
    _check(s, dense, s.error(), st, +1)


def test_simplify_stops_on_tolerance():
    rng = np.random.default_rng(99)
    psi = bl.random_mps([2] * 10, D=4, complex=True, rng=rng)
    st = core.Strategy(tolerance=1e-12, simplification_tolerance=1e-10, max_sweeps=6)
    assert _check(psi, psi.to_vector(), 0.0, st, +1) == 1  # exact at the first sweep


def test_planted_bookkeeping_mistakes_are_caught(monkeypatch):
    """A right-to-left sweep that forgets to move the environments, or a
    measure taken with the norm of the wrong tensor, must disagree with the
    dense recomputation."""
    rng = np.random.default_rng(314)
    psi = bl.random_mps([2] * 10, D=8, complex=True, rng=rng)
    st = core.Strategy(tolerance=1e-6, simplification_tolerance=1e-14, max_bond_dimension=5,
                       max_sweeps=4)
    v = psi.to_vector()

    def mismatch():
        _, seen = _run(psi, v, st, +1)
        return any(abs(e - d) > 1e-12 + 1e-9 * d for e, d, _, _ in seen)

    assert not mismatch()
    # 1: update_left leaves the right environment where it was
    with monkeypatch.context() as m:
        m.setattr(al.AntilinearForm, "update_left",
                  lambda self: setattr(self, "center", self.center - 1))
        assert mismatch()
    # 2: the one-site tensor is read one site off the centre
    with monkeypatch.context() as m:
        orig = al.AntilinearForm.tensor1site

        def shifted(self):
            c = self.center
            self.center = min(c + 1, self.size - 1) if c == 0 else c - 1
            try:
                return orig(self)
            finally:
                self.center = c

        m.setattr(al.AntilinearForm, "tensor1site", shifted)
        try:
            assert mismatch()
        except (ValueError, IndexError):
            pass  # shape mismatch: caught as well
