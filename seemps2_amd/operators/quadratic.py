"""QuadraticForm <psi|O|psi> on the device (reference: src/seemps/operators/quadratic.py:31-348).

Holds the MPO environments ``left_env[i]`` / ``right_env[i]`` as device tensors
and the per-bond two-site operator ``Hm`` (the H12 reshape of
quadratic.py:103,198-201), built once per bond on the device.
"""

from __future__ import annotations

import numpy as np

from .. import kernels as K
from ..state import CanonicalMPS, MPS, Strategy
from ..state.environments import (
    begin_mpo_environment,
    update_left_mpo_environment,
    update_right_mpo_environment,
)
from .mpo import MPO


class QuadraticForm:
    def __init__(self, O: MPO, state: CanonicalMPS, start: int = 0, ket: MPS | None = None):
        self.O = O
        if not isinstance(state, CanonicalMPS):
            raise TypeError("QuadraticForm: state must be a CanonicalMPS")
        if state.center not in [start, start + 1]:
            raise ValueError(
                f"QuadraticForm: state must be centered at start={start} or {start + 1}")
        self.state = state
        self.ket = ket if ket is not None else state
        self.size = size = state.size
        if size != O.size:
            raise ValueError("QuadraticForm: MPO and MPS do not have the same size")
        if any(Op.shape[1] != A.shape[1] for Op, A in zip(O, state)):
            raise ValueError("QuadraticForm: MPO and MPS dimensions do not match")
        one = begin_mpo_environment()
        left_env = [one] * size
        right_env = [one] * size
        env = one
        for i in range(size - 1, max(0, start - 1), -1):
            right_env[i - 1] = env = update_right_mpo_environment(env, state[i], O[i], self.ket[i])
        env = one
        for i in range(0, start):
            left_env[i + 1] = env = update_left_mpo_environment(env, state[i], O[i], self.ket[i])
        self.left_env = left_env
        self.right_env = right_env
        self.site = start
        self._hm: dict[int, object] = {}

    def two_site_Hm(self, i: int):
        """Hm[(c,k,l),(ko,lo,d)] for the MPO pair (i, i+1), cached."""
        Hm = self._hm.get(i)
        if Hm is None:
            Hm = self._hm[i] = K.two_site_operator(self.O[i], self.O[i + 1])
        return Hm

    def two_site_block(self, i: int):
        return self.left_env[i], self.two_site_Hm(i), self.right_env[i + 1]

    def two_site_matvec(self, i: int, v):
        """H_eff(i, i+1) v on a device tensor v[b, k, l, f] (quadratic.py:204-214)."""
        L, Hm, R = self.two_site_block(i)
        return K.two_site_matvec(L, Hm, R, K.as_dev(v), self.O[i].shape[1], self.O[i + 1].shape[1])

    def one_site_Hm(self, i: int):
        """Hm for site i alone: the MPO tensor paired with an identity tensor of
        its right bond (a one-site operator in the two-site layout, d2 = 1)."""
        key = ("one", i)
        Hm = self._hm.get(key)
        if Hm is None:
            w = self.O[i].shape[3]
            eye = K.as_dev(np.eye(w, dtype=np.complex128).reshape(w, 1, 1, w))
            Hm = self._hm[key] = K.two_site_operator(self.O[i], eye)
        return Hm

    def one_site_matvec(self, i: int, v):
        """H_eff(i) v on a device tensor v[b, k, f] (quadratic.py:239-271): the
        environments left of site i and right of site i around the MPO tensor."""
        v = K.as_dev(v)
        b, d, f = v.shape
        out = K.two_site_matvec(self.left_env[i], self.one_site_Hm(i), self.right_env[i],
                                v.reshape(b, d, 1, f), d, 1)
        return out.view(b, d, f)

    def two_site_operator_matvec(self, i: int, v):
        """H_eff(i, i+1) v keeping v's shape."""
        v = K.as_dev(v)
        return self.two_site_matvec(i, v).view(v.shape)

    def move_right(self, i: int):
        j = self.site
        self.left_env[i] = update_left_mpo_environment(self.left_env[j], self.state[j], self.O[j],
                                                       self.ket[j])
        self.site = i

    def move_left(self, i: int):
        j = self.site
        self.right_env[i] = update_right_mpo_environment(self.right_env[j], self.state[j],
                                                         self.O[j], self.ket[j])
        self.site = i

    def ket_2site(self, i: int):
        from ..core import _contract_last_and_first

        return _contract_last_and_first(self.ket[i], self.ket[i + 1])

    def update_2site_right(self, AB, i: int, strategy: Strategy) -> None:
        """Split and extend the left environment (quadratic.py:333-339)."""
        self.state.update_2site_right(AB, i, strategy)
        if i < self.size - 2:
            self.site = j = i + 1
            self.left_env[j] = update_left_mpo_environment(self.left_env[i], self.state[i],
                                                           self.O[i], self.ket[i])

    def update_2site_left(self, AB, i: int, strategy: Strategy) -> None:
        """Split and extend the right environment (quadratic.py:341-348)."""
        self.state.update_2site_left(AB, i, strategy)
        if i > 0:
            j = i + 1
            self.right_env[i] = update_right_mpo_environment(self.right_env[j], self.state[j],
                                                             self.O[j], self.ket[j])
            self.site = i - 1
