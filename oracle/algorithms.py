"""Oracle: MPS/MPSSum/MPO containers, simplify, MPO.apply, MPO environments, DMRG.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Containers are thin records around Python lists of NumPy tensors:
``OMPS(tensors, error, center)`` (centre ``None`` = not canonical),
``OMPSSum(weights, states)``, and an MPO as a list of 4-tensors.

Restates:
* ``state/simplification.py:32-280``  -> ``simplify``
* ``state/antilinear.py:38-146``       -> ``AntilinearForm``
* ``state/mpssum.py:130-228``          -> ``join``, ``delete_zero_components``
* ``operators/mpo.py:26-45,229-281,378-416`` -> ``mpo_multiply_tensor``, ``mpo_apply``,
  ``mpo_expectation``
* ``state/environments.py:15-51``      -> MPO environments
* ``operators/quadratic.py:40-78,187-237,333-348`` -> ``QuadraticForm``
* ``optimization/dmrg.py:15-199``      -> ``dmrg`` (ARPACK through scipy, as the reference)

Pinning (tests/test_oracle_algorithms.py): dense brute force (state vectors,
operator matrices) for MPO.apply / expectation / environments / the two-site
matvec; exactness of simplify without truncation, dense fidelity with
truncation; exact diagonalisation (N <= 10) and free fermions (XX chain) for
dmrg.  simplify's per-sweep error measure, its stopping sweep and the
accumulated error bound are pinned sweep by sweep against dense vectors, with
planted bookkeeping mistakes that the check catches
(tests/test_oracle_simplify_bookkeeping.py).  The DMRG trajectory before
convergence depends on ARPACK iterates: parity unpinned for those
intermediate values (final converged quantities are pinned).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from math import sqrt

import numpy as np
import scipy.sparse.linalg

from . import core
from .core import Strategy, Simplification


@dataclass
class OMPS:
    tensors: list
    error: float = 0.0
    center: int | None = None

    @property
    def size(self):
        return len(self.tensors)

    def copy(self):
        return OMPS(list(self.tensors), self.error, self.center)

    def norm_squared(self):
        # canonical_mps.py:161-165 / mps.py:240-242
        if self.center is not None:
            A = self.tensors[self.center]
            return float(np.vdot(A, A).real)
        return abs(core.scprod(self.tensors, self.tensors).real)

    def to_vector(self):
        psi = np.ones((1, 1))
        for A in self.tensors:
            a, d, b = A.shape
            psi = (psi @ A.reshape(a, d * b)).reshape(-1, b)
        return psi.reshape(-1)

    def bond_dimensions(self):
        return [A.shape[0] for A in self.tensors] + [self.tensors[-1].shape[-1]]


@dataclass
class OMPSSum:
    weights: list
    states: list

    @property
    def size(self):
        return self.states[0].size

    def error(self):
        return sum(abs(w) * s.error for w, s in zip(self.weights, self.states))

    def to_vector(self):
        return sum(w * s.to_vector() for w, s in zip(self.weights, self.states))


def interpret_center(center, size):
    if 0 <= center < size:
        return center
    center += size
    if 0 <= center < size:
        return center
    raise IndexError()


def canonical(state: OMPS, center: int, strategy: Strategy, normalize=None):
    """``canonical_mps.py:76-108`` (new list, original tensors untouched)."""
    out = OMPS(list(state.tensors), 0.0, None)
    center = interpret_center(center, out.size)
    if state.center is not None:
        out.error = state.error
        out.center = state.center
        out.error += core.recanonicalize(out.tensors, state.center, center, strategy)
    else:
        out.error += core.canonicalize(out.tensors, center, strategy)
    out.center = center
    if normalize is True or (normalize is None and strategy.normalize):
        A = out.tensors[center]
        n = np.linalg.norm(A.reshape(-1))
        if n:
            out.tensors[center] = A / n
    return out


def recenter(state: OMPS, center: int, strategy: Strategy):
    center = interpret_center(center, state.size)
    if center != state.center:
        state.error += core.recanonicalize(state.tensors, state.center, center, strategy)
        state.center = center
    return state


def update_2site_right(mps: OMPS, AA, site, strategy):
    # canonical_mps.py:326-345
    mps.tensors[site], mps.tensors[site + 1], err = core.left_orth_2site(AA, strategy)
    mps.center = site + 1
    mps.error += err


def update_2site_left(mps: OMPS, AA, site, strategy):
    # canonical_mps.py:347-366
    mps.tensors[site], mps.tensors[site + 1], err = core.right_orth_2site(AA, strategy)
    mps.center = site
    mps.error += err


class AntilinearForm:
    """``antilinear.py:13-146``: environments of <bra|ket> around ``center``."""

    def __init__(self, bra: OMPS, ket: OMPS, center: int = 0):
        size = bra.size
        rho = np.eye(1)
        self.R = [rho] * size
        for i in range(size - 1, center, -1):
            rho = core.update_right_environment(bra.tensors[i], ket.tensors[i], rho)
            self.R[i - 1] = rho
        rho = np.eye(1)
        self.L = [rho] * size
        for i in range(0, center):
            rho = core.update_left_environment(bra.tensors[i], ket.tensors[i], rho)
            self.L[i + 1] = rho
        self.bra, self.ket, self.size, self.center = bra, ket, size, center

    def tensor1site(self):
        c = self.center
        return np.einsum("li,ijk,kn->ljn", self.L[c], self.ket.tensors[c], self.R[c],
                         optimize=True)

    def tensor2site(self, direction):
        if direction > 0:
            i, j = self.center, self.center + 1
        else:
            i, j = self.center - 1, self.center
        return np.einsum(
            "li,ijk,kmo,op->ljmp",
            self.L[i], self.ket.tensors[i], self.ket.tensors[j], self.R[j], optimize=True,
        )

    def update_right(self):
        p = self.center
        self.L[p + 1] = core.update_left_environment(
            self.bra.tensors[p], self.ket.tensors[p], self.L[p]
        )
        self.center = p + 1

    def update_left(self):
        p = self.center
        self.R[p - 1] = core.update_right_environment(
            self.bra.tensors[p], self.ket.tensors[p], self.R[p]
        )
        self.center = p - 1


def join(s: OMPSSum) -> OMPS:
    """``mpssum.py:130-167``: block-diagonal joined tensors, weights in site 0."""
    L = s.size
    out = []
    for i in range(L):
        As = [st.tensors[i] for st in s.states]
        if i == 0:
            out.append(np.concatenate([w * A for w, A in zip(s.weights, As)], axis=2))
        elif i == L - 1:
            out.append(np.concatenate(As, axis=0))
        else:
            DL = sum(A.shape[0] for A in As)
            DR = sum(A.shape[2] for A in As)
            dtype = np.result_type(*As)
            B = np.zeros((DL, As[0].shape[1], DR), dtype=dtype)
            l = r = 0
            for A in As:
                B[l : l + A.shape[0], :, r : r + A.shape[2]] = A
                l += A.shape[0]
                r += A.shape[2]
            out.append(B)
    return OMPS(out)


def delete_zero_components(s: OMPSSum) -> float:
    """``mpssum.py:204-228``."""
    c = 0.0
    fw, fs = [], []
    for wi, si in zip(s.weights, s.states):
        wic = np.conj(wi)
        ni = (wic * wi).real * si.norm_squared()
        if ni:
            for wj, sj in zip(fw, fs):
                c += 2 * (wic * wj * core.scprod(si.tensors, sj.tensors)).real
            fs.append(si)
            fw.append(wi)
            c += ni
    if len(fw) < len(s.states):
        if not fw:
            s.weights = [0.0]
            s.states = [zero_state(s.states[0])]
            return 0.0
        s.weights, s.states = fw, fs
    return abs(c)


def zero_state(state: OMPS) -> OMPS:
    return OMPS([np.zeros((1, A.shape[1], 1)) for A in state.tensors], 0.0, 0)


def simplify(state, strategy: Strategy, direction: int = +1, guess=None, observer=None):
    """``simplification.py:32-144`` and ``:150-280`` (MPSSum).

    ``observer`` (test hook, not in the reference): called after every sweep
    as observer(tensors, err, last) with a copy of the current tensors, the
    sweep's error measure and the site that holds the centre."""
    if isinstance(state, OMPSSum):
        return _simplify_sum(state, strategy, direction, guess, observer)
    size = state.size
    start = 0 if direction > 0 else size - 1
    method = strategy.simplify
    if method == Simplification.CANONICAL_FORM:
        mps = canonical(state, start, strategy)
        return recenter(mps, size - 1 - start, strategy)
    if method == Simplification.DO_NOT_SIMPLIFY:
        return canonical(state, size - 1 - start, strategy)
    mps = canonical(state if guess is None else guess, start, strategy, normalize=False)
    norm_state_sqr = state.norm_squared()
    if not norm_state_sqr:
        return zero_state(state)
    form = AntilinearForm(mps, state, center=start)
    sweeps, err = _sweeps(mps, [1.0], [form], strategy, direction, norm_state_sqr, observer)
    return _finish(mps, state.error + sqrt(err), strategy, direction, sweeps, err)


def _simplify_sum(s: OMPSSum, strategy, direction, guess, observer=None):
    norm_state_sqr = delete_zero_components(s)
    if not norm_state_sqr:
        return zero_state(s.states[0])
    size = s.size
    start = 0 if direction > 0 else size - 1
    method = strategy.simplify
    if method == Simplification.CANONICAL_FORM:
        mps = canonical(join(s), start, strategy)
        return recenter(mps, size - 1 - start, strategy)
    if method == Simplification.DO_NOT_SIMPLIFY:
        return canonical(join(s), size - 1 - start, strategy)
    mps = canonical(join(s) if guess is None else guess, start, strategy, normalize=False)
    forms = [AntilinearForm(mps, si, center=start) for si in s.states]
    sweeps, err = _sweeps(mps, s.weights, forms, strategy, direction, norm_state_sqr, observer)
    return _finish(mps, s.error() + sqrt(err), strategy, direction, sweeps, err)


def _sweeps(mps, weights, forms, strategy, direction, norm_state_sqr, observer=None):
    size = mps.size
    err = 2.0
    history = []
    for _ in range(max(1, strategy.max_sweeps)):
        if direction > 0:
            for n in range(0, size - 1):
                theta = sum(w * f.tensor2site(direction) for w, f in zip(weights, forms))
                update_2site_right(mps, theta, n, strategy)
                for f in forms:
                    f.update_right()
            last = size - 1
        else:
            for n in reversed(range(0, size - 1)):
                theta = sum(w * f.tensor2site(direction) for w, f in zip(weights, forms))
                update_2site_left(mps, theta, n, strategy)
                for f in forms:
                    f.update_left()
            last = 0
        T = mps.tensors[last]
        norm_mps_sqr = np.vdot(T, T).real
        sc = np.vdot(T, sum(w * f.tensor1site() for w, f in zip(weights, forms)))
        old_err = err
        err = 2 * abs(1.0 - sc.real / sqrt(norm_mps_sqr * norm_state_sqr))
        history.append(err)
        if observer is not None:
            observer([A.copy() for A in mps.tensors], err, last)
        if err < strategy.simplification_tolerance or err > old_err:
            break
        direction = -direction
    mps._last = last
    mps._norm_sqr = norm_mps_sqr
    return history, err


def _finish(mps, total_error_bound, strategy, direction, history, err):
    if strategy.normalize and mps._norm_sqr:
        factor = sqrt(mps._norm_sqr)
        mps.tensors[mps._last] = mps.tensors[mps._last] / factor
        total_error_bound /= factor
    mps.error = total_error_bound
    mps.sweep_errors = history
    return mps


# ---------------------------------------------------------------- MPO side


def mpo_multiply_tensor(W: np.ndarray, B: np.ndarray) -> np.ndarray:
    """``mpo.py:26-45``: out[(c,a), i, (d,b)] = sum_j B[c,j,d] W[a,i,j,b]."""
    a, i, j, b = W.shape
    c, _, d = B.shape
    return np.einsum("cjd,aijb->caidb", B, W).reshape(c * a, i, d * b)


def mpo_apply(mpo: list, state, strategy: Strategy, simplify_flag=None):
    """``mpo.py:229-281``."""
    if simplify_flag is None:
        simplify_flag = strategy.simplify != Simplification.DO_NOT_SIMPLIFY
    if isinstance(state, OMPSSum):
        states = [
            OMPS([mpo_multiply_tensor(W, A) for W, A in zip(mpo, st.tensors)], st.error)
            for st in state.states
        ]
        out = OMPSSum(list(state.weights), states)
    else:
        out = OMPS([mpo_multiply_tensor(W, A) for W, A in zip(mpo, state.tensors)], state.error)
    if simplify_flag:
        out = simplify(out, strategy)
    return out


def update_left_mpo_environment(rho, A, O, B):
    """``environments.py:19-30``: acb,ajd,cjie,bif->def (A = bra), as three
    pairwise contractions (bra, MPO, ket), the order the reference uses."""
    t = np.tensordot(rho, A.conj(), axes=([0], [0]))  # c b j d
    t = np.tensordot(t, O, axes=([0, 2], [0, 1]))  # b d i e
    return np.tensordot(t, B, axes=([0, 2], [0, 1]))  # d e f


def update_right_mpo_environment(rho, A, O, B):
    """``environments.py:33-42``: def,ajd,cjie,bif->acb (A = bra), as three
    pairwise contractions (bra, MPO, ket)."""
    t = np.tensordot(A.conj(), rho, axes=([2], [0]))  # a j e f
    t = np.tensordot(t, O, axes=([1, 2], [1, 3]))  # a f c i
    return np.tensordot(t, B, axes=([1, 3], [2, 1]))  # a c b


def mpo_expectation(mpo: list, bra: OMPS, ket: OMPS | None = None):
    """``mpo.py:378-416``."""
    center = bra.center if bra.center is not None else len(mpo) - 1
    ket = bra if ket is None else ket
    left = right = np.ones((1, 1, 1))
    for i in range(0, center):
        left = update_left_mpo_environment(left, bra.tensors[i], mpo[i], ket.tensors[i])
    for i in range(len(mpo) - 1, center - 1, -1):
        right = update_right_mpo_environment(right, bra.tensors[i], mpo[i], ket.tensors[i])
    return np.dot(left.reshape(-1), right.reshape(-1))


def two_site_matvec(L, H12, R, v):
    """``quadratic.py:187-214``: out[a,j,l,g] = L[a,c,b] H[c,j,k,l,m,d] v[b,k,m,f] R[g,d,f].

    Same three-GEMM order as the reference (L first, then the MPO pair, then R)."""
    a, c, b = L.shape
    _, ko, k, lo, l, d = H12.shape
    g, _, f = R.shape
    x = (L.reshape(a * c, b) @ v.reshape(b, k * l * f)).reshape(a, c, k, l, f)
    Hm = H12.transpose(0, 2, 4, 1, 3, 5).reshape(c * k * l, ko * lo * d)
    x = x.transpose(0, 4, 1, 2, 3).reshape(a * f, c * k * l) @ Hm
    x = x.reshape(a, f, ko, lo, d).transpose(0, 2, 3, 1, 4).reshape(a * ko * lo, f * d)
    return (x @ R.transpose(2, 1, 0).reshape(f * d, g)).reshape(a, ko, lo, g)


class QuadraticForm:
    """``quadratic.py:31-348`` (the parts DMRG uses)."""

    def __init__(self, O: list, state: OMPS, start: int = 0):
        size = state.size
        one = np.ones((1, 1, 1))
        self.left_env = [one] * size
        self.right_env = [one] * size
        env = one
        for i in range(size - 1, max(0, start - 1), -1):
            env = update_right_mpo_environment(env, state.tensors[i], O[i], state.tensors[i])
            self.right_env[i - 1] = env
        env = one
        for i in range(0, start):
            env = update_left_mpo_environment(env, state.tensors[i], O[i], state.tensors[i])
            self.left_env[i + 1] = env
        self.O, self.state, self.size, self.site = O, state, size, start

    def two_site_block(self, i):
        H12 = np.tensordot(self.O[i], self.O[i + 1], axes=1)
        return self.left_env[i], H12, self.right_env[i + 1]

    def update_2site_right(self, AB, i, strategy):
        update_2site_right(self.state, AB, i, strategy)
        if i < self.size - 2:
            self.site = i + 1
            T = self.state.tensors[i]
            self.left_env[i + 1] = update_left_mpo_environment(self.left_env[i], T, self.O[i], T)

    def update_2site_left(self, AB, i, strategy):
        update_2site_left(self.state, AB, i, strategy)
        if i > 0:
            T = self.state.tensors[i + 1]
            self.right_env[i] = update_right_mpo_environment(
                self.right_env[i + 1], T, self.O[i + 1], T
            )
            self.site = i - 1


def diagonalize_two_site(QF: QuadraticForm, i, tol_eigs):
    """``dmrg.py:15-23`` (ARPACK via scipy, ``which='SA'``, ``v0`` = current theta)."""
    L, H, R = QF.two_site_block(i)
    v = np.tensordot(QF.state.tensors[i], QF.state.tensors[i + 1], axes=1)
    shape = (L.shape[0], H.shape[1], H.shape[3], R.shape[0])
    n = int(np.prod(v.shape))
    dtype = np.result_type(L, H, R, v)
    op = scipy.sparse.linalg.LinearOperator(
        (n, n), dtype=dtype,
        matvec=lambda x: two_site_matvec(L, H, R, x.reshape(v.shape)).reshape(-1),
    )
    vals, vecs = scipy.sparse.linalg.eigsh(op, 1, which="SA", v0=v.reshape(-1), tol=tol_eigs)
    return float(vals[0]), vecs[:, 0].reshape(shape)


@dataclass
class OptimizeResults:
    state: OMPS
    energy: float
    converged: bool
    message: str
    trajectory: list = field(default_factory=list)
    variances: list = field(default_factory=list)


def energy_and_variance(H: list, state: OMPS):
    """``dmrg.py:55-59``: energy is NOT divided by <psi|psi>."""
    Hpsi = OMPS([mpo_multiply_tensor(W, A) for W, A in zip(H, state.tensors)])
    energy = float(mpo_expectation(H, state).real)
    variance = abs(Hpsi.norm_squared() - energy * energy)
    return energy, variance


def convergence_reason(energy_change, energy_scale, tol, tol_up):
    """``dmrg.py:45-52``."""
    if energy_change > tol_up * energy_scale:
        return f"Energy fluctuation above tolerance {tol_up}"
    if -tol * energy_scale <= energy_change <= 0:
        return f"Energy decrease slower than tolerance {tol}"
    return None


def dmrg(H: list, guess: OMPS, maxiter=20, tol=1e-10, tol_up=None, tol_eigs=None,
         strategy: Strategy = core.DEFAULT_STRATEGY):
    """``dmrg.py:76-199``."""
    if maxiter < 1:
        raise ValueError("maxiter must be positive")
    if len(H) < 2:
        raise ValueError("DMRG requires at least two sites")
    if tol < 0:
        raise ValueError("tol must be non-negative")
    tol_up = tol if tol_up is None else tol_up
    if tol_up < 0:
        raise ValueError("tol_up must be non-negative")
    tol_eigs = tol if tol_eigs is None else tol_eigs
    strat = strategy.replace(normalize=True)
    size = len(H)
    if guess.center not in (0, size - 1):
        guess = canonical(guess, 0, strat)
    A = guess.tensors[guess.center]
    n = np.linalg.norm(A.reshape(-1))
    if n:
        guess.tensors[guess.center] = A / n
    if guess.center == 0:
        direction = +1
        QF = QuadraticForm(H, guess, start=0)
    else:
        direction = -1
        QF = QuadraticForm(H, guess, start=size - 2)
    energy, variance = energy_and_variance(H, QF.state)
    results = OptimizeResults(
        state=_deepcopy(QF.state), energy=energy, converged=False,
        message=f"Exceeded maximum number of steps {maxiter}",
        trajectory=[energy], variances=[variance],
    )
    last_energy = energy
    for _sweep in range(1, maxiter + 1):
        sites = range(size - 1) if direction > 0 else range(size - 2, -1, -1)
        for i in sites:
            _, AB = diagonalize_two_site(QF, i, tol_eigs)
            if direction > 0:
                QF.update_2site_right(AB, i, strat)
            else:
                QF.update_2site_left(AB, i, strat)
        direction = -direction
        energy, variance = energy_and_variance(H, QF.state)
        results.trajectory.append(energy)
        results.variances.append(variance)
        if energy < results.energy:
            results.energy, results.state = energy, _deepcopy(QF.state)
        reason = convergence_reason(energy - last_energy, max(1.0, abs(energy)), tol, tol_up)
        if reason is not None:
            results.converged = True
            results.message = reason
            break
        last_energy = energy
    return results


def _deepcopy(state: OMPS) -> OMPS:
    return OMPS([A.copy() for A in state.tensors], state.error, state.center)
