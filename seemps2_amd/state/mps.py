"""MPS container on the device (reference: src/seemps/state/mps.py:22-504).

Tensors ``A[alpha, i, beta]`` are complex128 torch tensors on the HIP device.
Host NumPy arrays appear only at the API edges (``to_vector``, inputs).
"""

from __future__ import annotations

import math
from collections.abc import Iterable, Sequence

import numpy as np
import torch

from .. import kernels as K
from ..core import DEFAULT_STRATEGY, Strategy, scprod


class TensorArray(Sequence):
    """List of tensors copied on construction (src/seemps/state/array.py:10-72)."""

    def __init__(self, data: Iterable):
        self._data = [K.as_dev(A) for A in data]
        self.size = len(self._data)

    def __getitem__(self, k):
        return self._data[k]

    def __setitem__(self, k, value):
        if isinstance(k, slice):
            self._data[k] = [K.as_dev(v) for v in value]
        else:
            self._data[k] = K.as_dev(value)
        return value

    def __iter__(self):
        return self._data.__iter__()

    def __reversed__(self):
        return self._data.__reversed__()

    def __len__(self):
        return self.size


class MPS(TensorArray):
    """Matrix product state with open boundaries (mps.py:22-58)."""

    __array_priority__ = 10000

    def __init__(self, data: Iterable, error: float = 0):
        super().__init__(data)
        self._error = error

    def copy(self) -> "MPS":
        """Shallow copy: new list, same tensors (mps.py:55-58)."""
        return MPS(self._data, self._error)

    def as_mps(self) -> "MPS":
        return self

    def dimension(self) -> int:
        return math.prod(self.physical_dimensions())

    def physical_dimensions(self) -> list[int]:
        return [a.shape[1] for a in self._data]

    def bond_dimensions(self) -> list[int]:
        return [a.shape[0] for a in self._data] + [self._data[-1].shape[-1]]

    def max_bond_dimension(self) -> int:
        return max(a.shape[-1] for a in self._data)

    def to_vector(self) -> np.ndarray:
        """Dense state vector (host), contracted on the device (mps.py:488-504)."""
        psi = None
        for A in reversed(self._data):
            a, d, b = A.shape
            if psi is None:
                psi = A.reshape(a * d, b)
            else:
                psi = K.gemm(A.reshape(a * d, b), psi)
            psi = psi.reshape(a, -1)
        return K.to_numpy(psi).reshape(-1)

    def to_tensor(self) -> np.ndarray:
        return self.to_vector().reshape(self.physical_dimensions())

    @classmethod
    def from_vector(cls, psi, dimensions: Sequence[int], strategy: Strategy = DEFAULT_STRATEGY,
                    normalize: bool = True, center: int = -1) -> "MPS":
        from .schmidt import _vector2mps

        data, error = _vector2mps(psi, dimensions, strategy, normalize, center)
        return MPS(data, error)

    @classmethod
    def from_tensor(cls, state, strategy: Strategy = DEFAULT_STRATEGY,
                    normalize: bool = True) -> "MPS":
        state = np.asarray(state)
        return cls.from_vector(state.reshape(-1), state.shape, strategy, normalize)

    def __add__(self, state):
        from .mpssum import MPSSum

        if isinstance(state, MPS):
            return MPSSum([1.0, 1.0], [self, state], check_args=False)
        if isinstance(state, MPSSum):
            return MPSSum([1.0] + state.weights, [self] + state.states, check_args=False)
        raise TypeError(f"Invalid operation + between {type(self)} and {type(state)}")

    def __sub__(self, state):
        from .mpssum import MPSSum

        if isinstance(state, MPS):
            return MPSSum([1, -1], [self, state], check_args=False)
        if isinstance(state, MPSSum):
            return MPSSum([1] + [-w for w in state.weights], [self] + state.states,
                          check_args=False)
        raise TypeError(f"Invalid operation - between {type(self)} and {type(state)}")

    def __mul__(self, n) -> "MPS":
        """n * self for a scalar n: scales the first tensor; for an MPS n, the
        site-wise (Hadamard) product with bonds (self, n) (mps.py:191-212)."""
        if isinstance(n, (int, float, complex, np.number)):
            if n:
                out = self.copy()
                out._data[0] = K.scale(complex(n), out._data[0])
                out._error *= abs(n)
                return out
            return self.zero_state()
        if isinstance(n, MPS):
            if n.size != self.size:
                raise ValueError("MPS product: sizes differ")
            return MPS([K.site_product(A, B) for A, B in zip(self._data, n._data)])
        raise TypeError(f"Invalid operation * between {type(self)} and {type(n)}")

    def __rmul__(self, n) -> "MPS":
        if isinstance(n, (int, float, complex, np.number)):
            return self.__mul__(n)
        raise TypeError(f"Invalid operation * between {type(n)} and {type(self)}")

    def __truediv__(self, n) -> "MPS":
        if isinstance(n, (int, float, complex)):
            return self.__mul__(1.0 / n)
        raise TypeError(f"Invalid operation / between {type(self)} and {type(n)}")

    def norm_squared(self) -> float:
        """|<psi|psi>| real part (mps.py:240-242)."""
        return abs(scprod(self, self).real)

    def norm(self) -> float:
        return math.sqrt(abs(scprod(self, self)))

    def zero_state(self) -> "MPS":
        return MPS([np.zeros((1, A.shape[1], 1)) for A in self._data])

    def left_environment(self, site: int):
        from ..core import _begin_environment, _update_left_environment

        rho = _begin_environment()
        for A in self._data[:site]:
            rho = _update_left_environment(A, A, rho)
        return rho

    def right_environment(self, site: int):
        from ..core import _begin_environment, _update_right_environment

        rho = _begin_environment()
        for A in self._data[-1:site:-1]:
            rho = _update_right_environment(A, A, rho)
        return rho

    def error(self) -> float:
        return self._error

    def update_error(self, delta: float) -> None:
        self._error += delta

    def conj(self) -> "MPS":
        out = self.copy()
        out._data = [K.conj(A) for A in out._data]
        return out

    def reverse(self) -> "MPS":
        return MPS([K.permute(A, (2, 1, 0)) for A in reversed(self._data)])
