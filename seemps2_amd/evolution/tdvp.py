"""Two-site TDVP (reference: src/seemps/evolution/tdvp.py:13-141).

Per bond, the two-site tensor is evolved by exp(dt/2 H_eff(i, i+1)) and split
with the Strategy's truncation; the new centre tensor is then evolved back by
exp(-dt/2 H_eff(i+1)) (sweep right), and the mirror image on the way back.
The effective operators are the device matvecs of ``QuadraticForm`` (the
one-site one is the two-site kernel with an identity partner tensor).
``exp(tau A) v`` is an Arnoldi iteration on device vectors: each step one
matvec, two classical Gram-Schmidt passes as complex GEMMs against the basis,
one norm; only the small Hessenberg matrix comes to the host, where its
exponential (shifted by the Rayleigh quotient of v) and Saad's error estimate
decide convergence, as in the reference.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.linalg

from .. import kernels as K
from ..core import DEFAULT_STRATEGY, Strategy, _contract_last_and_first
from ..operators.quadratic import QuadraticForm
from ..state import CanonicalMPS, MPS
from .euler import ode_solver

_OP_T, _OP_C = 1, 3


def _expmv(apply, v, tau: complex, tol: float = 1e-12, m_max: int = 30):
    """exp(tau A) v for the operator `apply` on device tensors (tdvp.py:13-68)."""
    v = K.as_dev(v)
    shape = v.shape
    flat = v.reshape(-1)
    n = flat.numel()
    norm_v = math.sqrt(max(K.dotc(flat, flat).real, 0.0))
    if norm_v == 0.0:
        return v
    Q = K.empty(m_max + 1, n)
    K.axpby(1.0 / norm_v, flat, 0.0, Q[0])
    H = np.zeros((m_max + 1, m_max), dtype=complex)
    mu = None
    tol_v = tol * norm_v
    abs_tau = abs(tau)
    abs_tau_pow = inv_fact = prod_h = 1.0
    m = 0
    e1p = None
    for j in range(m_max):
        w = apply(Q[j].view(shape)).reshape(n, 1)
        Qj = Q[: j + 1]
        h = np.zeros(j + 1, dtype=complex)
        for _ in range(2):  # classical Gram-Schmidt, twice
            c = K.gemm(Qj, w, opA=_OP_C)                        # conj(Q) w
            K.gemm(Qj, c, opA=_OP_T, alpha=-1.0, beta=1.0, C=w)  # w -= Q^T c
            h += K.to_numpy(c).reshape(-1)
        H[: j + 1, j] = h
        if mu is None:
            mu = H[0, 0]
        hn = math.sqrt(max(K.dotc(w, w).real, 0.0))
        H[j + 1, j] = hn
        m = j + 1
        # leading-order Saad bound: skip the exponential while surely unconverged
        if hn * abs_tau_pow * inv_fact * prod_h > tol_v and hn >= 1e-14:
            prod_h *= hn
            abs_tau_pow *= abs_tau
            inv_fact /= m
            K.axpby(1.0 / hn, w.view(-1), 0.0, Q[j + 1])
            continue
        Hm = H[:m, :m].copy()
        Hm.flat[:: m + 1] -= mu
        e1p = scipy.linalg.expm(tau * Hm)[:, 0]
        if hn * abs(e1p[m - 1]) <= tol_v or hn < 1e-14:
            break
        prod_h *= hn
        abs_tau_pow *= abs_tau
        inv_fact /= m
        K.axpby(1.0 / hn, w.view(-1), 0.0, Q[j + 1])
    if e1p is None:
        Hm = H[:m, :m].copy()
        Hm.flat[:: m + 1] -= mu
        e1p = scipy.linalg.expm(tau * Hm)[:, 0]
    coef = K.as_dev((np.exp(tau * mu) * norm_v * e1p).reshape(m, 1))
    return K.gemm(Q[:m], coef, opA=_OP_T).view(shape)                # Q^T coef


def tdvp_step(L, state: MPS, dt, strategy: Strategy = DEFAULT_STRATEGY) -> CanonicalMPS:
    """One symmetric two-site TDVP sweep pair (tdvp.py:71-103)."""
    if not isinstance(state, CanonicalMPS):
        state = CanonicalMPS(state, center=0, strategy=strategy)
    QF = QuadraticForm(L, state, start=0)
    size = L.size
    for i in range(size - 1):
        A2 = _contract_last_and_first(QF.state[i], QF.state[i + 1])
        A2 = _expmv(lambda x: QF.two_site_operator_matvec(i, x), A2, 0.5 * dt)
        QF.update_2site_right(A2, i, strategy)
        if i < size - 2:
            QF.state[i + 1] = _expmv(lambda x: QF.one_site_matvec(i + 1, x), QF.state[i + 1],
                                     -0.5 * dt)
    for i in range(size - 2, -1, -1):
        A2 = _contract_last_and_first(QF.state[i], QF.state[i + 1])
        A2 = _expmv(lambda x: QF.two_site_operator_matvec(i, x), A2, 0.5 * dt)
        QF.update_2site_left(A2, i, strategy)
        if i > 0:
            QF.state[i] = _expmv(lambda x: QF.one_site_matvec(i, x), QF.state[i], -0.5 * dt)
    return QF.state


def tdvp(L, time, state: MPS, steps: int = 1000, strategy: Strategy = DEFAULT_STRATEGY,
         callback=None):
    """Solve d|state>/dt = L|state> by two-site TDVP steps (tdvp.py:106-141)."""

    def step(t, psi, dt, strategy):
        return tdvp_step(L, psi, dt, strategy)

    return ode_solver(step, time, state, steps, strategy, callback)


__all__ = ["tdvp", "tdvp_step"]
