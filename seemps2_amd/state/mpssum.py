"""MPSSum on the device (reference: src/seemps/state/mpssum.py:11-262)."""

from __future__ import annotations

import math
from collections.abc import Iterable

import numpy as np

from .. import kernels as K
from ..core import scprod
from .mps import MPS


class MPSSum:
    """Weighted sum of MPS, kept as (weights, states) until joined or simplified."""

    __array_priority__ = 10000

    def __init__(self, weights: Iterable, states: Iterable, check_args: bool = True):
        if check_args:
            self.weights, self.states = [], []
            for w, s in zip(weights, states):
                if isinstance(s, MPS):
                    self.weights.append(w)
                    self.states.append(s)
                elif isinstance(s, MPSSum):
                    self.weights.extend(w * wi for wi in s.weights)
                    self.states.extend(s.states)
                else:
                    raise ValueError(s)
        else:
            self.weights = list(weights)
            self.states = list(states)
        self.size = self.states[0].size

    def as_mps(self) -> MPS:
        return self.join()

    def copy(self) -> "MPSSum":
        return MPSSum(self.weights.copy(), self.states.copy())

    __copy__ = copy

    def __add__(self, state):
        if isinstance(state, MPS):
            return MPSSum(self.weights + [1.0], self.states + [state], check_args=False)
        if isinstance(state, MPSSum):
            return MPSSum(self.weights + state.weights, self.states + state.states,
                          check_args=False)
        raise TypeError(f"Invalid operation + between {type(self)} and {type(state)}")

    def __sub__(self, state):
        if isinstance(state, MPS):
            return MPSSum(self.weights + [-1], self.states + [state], check_args=False)
        if isinstance(state, MPSSum):
            return MPSSum(self.weights + [-w for w in state.weights],
                          self.states + state.states, check_args=False)
        raise TypeError(f"Invalid operation - between {type(self)} and {type(state)}")

    def __mul__(self, n):
        if isinstance(n, (int, float, complex)):
            return MPSSum([n * w for w in self.weights], self.states, check_args=False)
        raise TypeError(f"Invalid operation * between {type(self)} and {type(n)}")

    __rmul__ = __mul__

    def __truediv__(self, n):
        if isinstance(n, (int, float, complex)):
            return self.__mul__(1.0 / n)
        raise TypeError(f"Invalid operation / between {type(self)} and {type(n)}")

    def physical_dimensions(self) -> list[int]:
        return self.states[0].physical_dimensions()

    def to_vector(self) -> np.ndarray:
        return sum(w * s.to_vector() for w, s in zip(self.weights, self.states))

    def to_tensor(self) -> np.ndarray:
        return self.to_vector().reshape(self.physical_dimensions())

    def _joined_tensors(self, i: int, L: int):
        """Block-diagonal join (mpssum.py:130-155); weights fold into site 0."""
        As = [s[i] for s in self.states]
        d = As[0].shape[1]
        if i == 0:
            DR = sum(A.shape[2] for A in As)
            B = K.zeros(As[0].shape[0], d, DR)
            off = 0
            for w, A in zip(self.weights, As):
                K.embed_block(B, K.scale(w, A), 0, off)
                off += A.shape[2]
            return B
        if i == L - 1:
            DL = sum(A.shape[0] for A in As)
            B = K.empty(DL, d, As[0].shape[2])
            off = 0
            for A in As:
                K.embed_block(B, A, off, 0)
                off += A.shape[0]
            return B
        DL = sum(A.shape[0] for A in As)
        DR = sum(A.shape[2] for A in As)
        B = K.zeros(DL, d, DR)
        l = r = 0
        for A in As:
            K.embed_block(B, A, l, r)
            l += A.shape[0]
            r += A.shape[2]
        return B

    def join(self) -> MPS:
        L = self.size
        return MPS([self._joined_tensors(i, L) for i in range(L)])

    def join_canonical(self, *args, **kwargs):
        from .canonical_mps import CanonicalMPS

        return CanonicalMPS(self.join(), *args, **kwargs)

    def conj(self) -> "MPSSum":
        return MPSSum([np.conj(w) for w in self.weights], [s.conj() for s in self.states])

    def norm_squared(self) -> float:
        w, s = self.weights, self.states
        L = len(w)
        return abs(sum(
            (np.conj(w[i]) * w[j] * scprod(s[i], s[j])).real * (1 if i == j else 2)
            for i in range(L) for j in range(i, L)
        ))

    def norm(self) -> float:
        return math.sqrt(self.norm_squared())

    def error(self) -> float:
        return sum(abs(w) * s._error for w, s in zip(self.weights, self.states))

    def delete_zero_components(self) -> float:
        """Norm squared, dropping zero terms (mpssum.py:204-228)."""
        c = 0.0
        fw, fs = [], []
        for wi, si in zip(self.weights, self.states):
            wic = np.conj(wi)
            ni = (wic * wi).real * si.norm_squared()
            if ni:
                for wj, sj in zip(fw, fs):
                    c += 2 * (wic * wj * scprod(si, sj)).real
                fs.append(si)
                fw.append(wi)
                c += ni
        if len(fw) < len(self.states):
            if not fw:
                self.weights = [0.0]
                self.states = [self.states[0].zero_state()]
                return 0.0
            self.weights, self.states = fw, fs
        return abs(c)

    def reverse(self) -> "MPSSum":
        return MPSSum(self.weights, [s.reverse() for s in self.states])


def to_mps(mps_or_sum) -> MPS:
    if isinstance(mps_or_sum, MPSSum):
        return mps_or_sum.join()
    return mps_or_sum
