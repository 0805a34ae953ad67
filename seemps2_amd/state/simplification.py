"""Variational MPS simplification on the device.

Reference: src/seemps/state/simplification.py:22-296.  Same control flow:
canonical guess centred at the starting end, AntilinearForm environments,
two-site sweeps with truncated SVD splits, the stopping rule
``err < simplification_tolerance or err > old_err`` with
``err = 2 |1 - Re<xi|psi> / sqrt(|xi|^2 |psi|^2)|``, the ``_error``
bookkeeping ``state._error + sqrt(err)`` and last-tensor normalisation only
when ``strategy.normalize``.
"""

from __future__ import annotations

from math import sqrt

from .. import kernels as K
from ..core import (
    DEFAULT_TOLERANCE,
    MAX_BOND_DIMENSION,
    Simplification,
    Strategy,
    Truncation,
)
from .antilinear import AntilinearForm
from .canonical_mps import CanonicalMPS
from .mps import MPS
from .mpssum import MPSSum

SIMPLIFICATION_STRATEGY = Strategy(
    method=Truncation.RELATIVE_NORM_SQUARED_ERROR,
    tolerance=DEFAULT_TOLERANCE,
    max_bond_dimension=MAX_BOND_DIMENSION,
    normalize=True,
    max_sweeps=4,
    simplify=Simplification.VARIATIONAL,
)


def simplify(state, strategy: Strategy = SIMPLIFICATION_STRATEGY, direction: int = +1,
             guess: MPS | None = None) -> CanonicalMPS:
    """Approximate ``state`` (MPS or MPSSum) by a CanonicalMPS (simplification.py:32-144)."""
    if isinstance(state, MPSSum):
        return simplify_mps_sum(state, strategy, direction, guess)
    size = state.size
    start = 0 if direction > 0 else -1
    method = strategy.get_simplification_method()
    if method == Simplification.CANONICAL_FORM:
        return CanonicalMPS(state, center=start, strategy=strategy).recenter(-1 - start, strategy)
    if method == Simplification.DO_NOT_SIMPLIFY:
        return CanonicalMPS(state, center=-1 - start, strategy=strategy)
    mps = CanonicalMPS(state if guess is None else guess, center=start, normalize=False,
                       strategy=strategy)
    norm_state_sqr = state.norm_squared()
    if not norm_state_sqr:
        return CanonicalMPS(state.zero_state(), is_canonical=True)
    forms = [AntilinearForm(mps, state, center=start)]
    err, last, norm_mps_sqr = _sweeps(mps, [1.0], forms, strategy, direction, size,
                                      norm_state_sqr)
    return _finish(mps, state._error + sqrt(err), strategy, last, norm_mps_sqr)


def simplify_mps_sum(sum_state: MPSSum, strategy: Strategy = SIMPLIFICATION_STRATEGY,
                     direction: int = +1, guess: MPS | None = None) -> CanonicalMPS:
    """Simplify sum_i w_i psi_i (simplification.py:150-280)."""
    norm_state_sqr = sum_state.delete_zero_components()
    if not norm_state_sqr:
        return CanonicalMPS(sum_state.states[0].zero_state(), is_canonical=True)
    start = 0 if direction > 0 else -1
    method = strategy.get_simplification_method()
    if method == Simplification.CANONICAL_FORM:
        return CanonicalMPS(sum_state.join(), center=start, strategy=strategy).recenter(
            -1 - start, strategy)
    if method == Simplification.DO_NOT_SIMPLIFY:
        return CanonicalMPS(sum_state.join(), center=-1 - start, strategy=strategy)
    mps = CanonicalMPS(sum_state.join() if guess is None else guess, center=start,
                       normalize=False, strategy=strategy)
    size = mps.size
    forms = [AntilinearForm(mps, s, center=start) for s in sum_state.states]
    err, last, norm_mps_sqr = _sweeps(mps, sum_state.weights, forms, strategy, direction, size,
                                      norm_state_sqr)
    return _finish(mps, sum_state.error() + sqrt(err), strategy, last, norm_mps_sqr)


def _theta(weights, forms, direction):
    if len(forms) == 1 and weights[0] == 1.0:
        return forms[0].tensor2site(direction)
    out = None
    for w, f in zip(weights, forms):
        if out is None:
            out = f.tensor2site(direction, alpha=w)
        else:
            f.tensor2site(direction, out=out, alpha=w, beta=1.0)
    return out


def _sweeps(mps, weights, forms, strategy, direction, size, norm_state_sqr):
    err = 2.0
    norm_mps_sqr = 0.0
    last = 0
    for _ in range(max(1, strategy.get_max_sweeps())):
        if direction > 0:
            for n in range(0, size - 1):
                mps.update_2site_right(_theta(weights, forms, direction), n, strategy)
                for f in forms:
                    f.update_right()
            last = size - 1
        else:
            for n in reversed(range(0, size - 1)):
                mps.update_2site_left(_theta(weights, forms, direction), n, strategy)
                for f in forms:
                    f.update_left()
            last = 0
        T = mps[last]
        norm_mps_sqr = K.dotc(T, T).real
        one = None
        for w, f in zip(weights, forms):
            t1 = f.tensor1site()
            one = K.scale(w, t1) if one is None else K.axpby(w, t1, 1.0, one)
        sc = K.dotc(T, one)
        old_err = err
        err = 2 * abs(1.0 - sc.real / sqrt(norm_mps_sqr * norm_state_sqr))
        if err < strategy.get_simplification_tolerance() or err > old_err:
            break
        direction = -direction
    return err, last, norm_mps_sqr


def _finish(mps, total_error_bound, strategy, last, norm_mps_sqr):
    if strategy.get_normalize_flag() and norm_mps_sqr:
        factor = sqrt(norm_mps_sqr)
        mps._data[last] = K.scale(1.0 / factor, mps._data[last])
        total_error_bound /= factor
    mps._error = total_error_bound
    return mps


def combine(weights, states, strategy: Strategy = SIMPLIFICATION_STRATEGY, direction: int = +1,
            guess: MPS | None = None) -> CanonicalMPS:
    """Deprecated alias kept for API parity (simplification.py:283-291)."""
    return simplify_mps_sum(MPSSum(weights, states))


simplify_mps = simplify

__all__ = ["simplify", "simplify_mps", "simplify_mps_sum", "combine", "SIMPLIFICATION_STRATEGY"]
