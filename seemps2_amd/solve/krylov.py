"""Krylov solvers on device vectors for the local systems of dmrg_solve.

Same stopping rule as scipy.sparse.linalg's solvers that the reference calls
(src/seemps/solve/dmrg.py:17-23): ||b - A x|| <= max(rtol ||b||, atol), and
scipy's default iteration limits.  Vectors stay on the device; every step is
one operator application (the device two-site matvec) plus device dot
products and axpys; only scalars come to the host.
"""

from __future__ import annotations

import math

import numpy as np

from .. import kernels as K

_OP_T, _OP_C = 1, 3


def _dot(x, y) -> complex:
    return K.dotc(x, y)


def _norm(x) -> float:
    return math.sqrt(max(K.dotc(x, x).real, 0.0))


def cg(apply, b, x0, atol: float = 0.0, rtol: float = 1e-5, maxiter: int | None = None):
    """Conjugate gradients (Hermitian positive definite operators)."""
    n = b.numel()
    maxiter = 10 * n if maxiter is None else maxiter
    tol = max(atol, rtol * _norm(b))
    x = K.copy(x0)
    r = K.copy(b)
    K.axpby(-1.0, apply(x), 1.0, r)
    if _norm(r) <= tol:
        return x, 0
    p = K.copy(r)
    rr = _dot(r, r).real
    for it in range(maxiter):
        Ap = apply(p)
        alpha = rr / _dot(p, Ap).real
        K.axpby(alpha, p, 1.0, x)
        K.axpby(-alpha, Ap, 1.0, r)
        rr_new = _dot(r, r).real
        if math.sqrt(rr_new) <= tol:
            return x, 0
        K.axpby(1.0, r, rr_new / rr, p)
        rr = rr_new
    return x, maxiter


def bicgstab(apply, b, x0, atol: float = 0.0, rtol: float = 1e-5, maxiter: int | None = None):
    """Stabilised biconjugate gradients (van der Vorst)."""
    n = b.numel()
    maxiter = 10 * n if maxiter is None else maxiter
    tol = max(atol, rtol * _norm(b))
    x = K.copy(x0)
    r = K.copy(b)
    K.axpby(-1.0, apply(x), 1.0, r)
    if _norm(r) <= tol:
        return x, 0
    rhat = K.copy(r)
    rho = alpha = omega = 1.0
    v = K.zeros(*b.shape)
    p = K.zeros(*b.shape)
    for it in range(maxiter):
        rho_new = _dot(rhat, r)
        if rho_new == 0:
            return x, -10
        beta = (rho_new / rho) * (alpha / omega)
        # p = r + beta (p - omega v)
        K.axpby(-omega, v, 1.0, p)
        K.axpby(1.0, r, beta, p)
        v = apply(p)
        alpha = rho_new / _dot(rhat, v)
        s = K.copy(r)
        K.axpby(-alpha, v, 1.0, s)
        if _norm(s) <= tol:
            K.axpby(alpha, p, 1.0, x)
            return x, 0
        t = apply(s)
        tt = _dot(t, t).real
        omega = _dot(t, s) / tt if tt else 0.0
        K.axpby(alpha, p, 1.0, x)
        K.axpby(omega, s, 1.0, x)
        r = s
        K.axpby(-omega, t, 1.0, r)
        if _norm(r) <= tol:
            return x, 0
        if omega == 0:
            return x, -11
        rho = rho_new
    return x, maxiter


def gmres(apply, b, x0, atol: float = 0.0, rtol: float = 1e-5, restart: int = 20,
          maxiter: int | None = None):
    """Restarted GMRES with Arnoldi (classical Gram-Schmidt twice, as complex
    GEMMs against the basis) and Givens rotations on the host."""
    n = b.numel()
    maxiter = n if maxiter is None else maxiter
    restart = min(restart, n)
    tol = max(atol, rtol * _norm(b))
    x = K.copy(x0)
    shape = b.shape
    for outer in range(max(1, maxiter)):
        r = K.copy(b)
        K.axpby(-1.0, apply(x), 1.0, r)
        beta = _norm(r)
        if beta <= tol:
            return x, 0
        V = K.empty(restart + 1, n)
        K.axpby(1.0 / beta, r.reshape(-1), 0.0, V[0])
        H = np.zeros((restart + 1, restart), dtype=complex)
        cs = np.zeros(restart, dtype=complex)
        sn = np.zeros(restart, dtype=complex)
        g = np.zeros(restart + 1, dtype=complex)
        g[0] = beta
        k_used = 0
        for j in range(restart):
            w = apply(V[j].view(shape)).reshape(n, 1)
            Vj = V[: j + 1]
            h = np.zeros(j + 1, dtype=complex)
            for _ in range(2):
                c = K.gemm(Vj, w, opA=_OP_C)
                K.gemm(Vj, c, opA=_OP_T, alpha=-1.0, beta=1.0, C=w)
                h += K.to_numpy(c).reshape(-1)
            hn = _norm(w)
            H[: j + 1, j] = h
            H[j + 1, j] = hn
            # apply the previous rotations, then make a new one
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + np.conj(cs[i]) * H[i + 1, j]
                H[i, j] = t
            a, bb = H[j, j], H[j + 1, j]
            den = math.sqrt(abs(a) ** 2 + abs(bb) ** 2)
            cs[j], sn[j] = (a / den, bb / den) if den else (1.0, 0.0)
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] = np.conj(cs[j]) * g[j]
            k_used = j + 1
            if abs(g[j + 1]) <= tol or hn == 0.0:
                break
            K.axpby(1.0 / hn, w.view(-1), 0.0, V[j + 1])
        y = np.linalg.solve(np.triu(H[:k_used, :k_used]), g[:k_used])
        upd = K.gemm(V[:k_used], K.as_dev(y.reshape(k_used, 1)), opA=_OP_T)
        K.axpby(1.0, upd.view(shape), 1.0, x)
        if abs(g[k_used]) <= tol:
            return x, 0
    return x, maxiter


SOLVERS = {"cg": cg, "bicgstab": bicgstab, "gmres": gmres, "lgmres": gmres}

__all__ = ["SOLVERS", "bicgstab", "cg", "gmres"]
