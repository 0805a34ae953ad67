"""MPO on the device (reference: src/seemps/operators/mpo.py:26-416).

``MPO.apply`` builds each output site with one local-operator contraction
(``cjd,aijb->caidb``: the MPS bond is the outer index, the MPO bond the inner
one, as in the reference) and then simplifies with the MPO's strategy unless
told otherwise.
"""

from __future__ import annotations

import numpy as np

from .. import kernels as K
from ..core import DEFAULT_STRATEGY, Strategy
from ..state import MPS, CanonicalMPS, MPSSum
from ..state.environments import (
    begin_mpo_environment,
    join_mpo_environments,
    update_left_mpo_environment,
    update_right_mpo_environment,
)
from ..state.mps import TensorArray

def _apply_op(W):
    # Mop[j, (a, i), b] = W[a, i, j, b], materialised once per MPO tensor
    return K.permuted(W, (2, 0, 1, 3))


def _mpo_multiply_tensor(A, B):
    """out[(c,a), i, (d,b)] = sum_j B[c,j,d] A[a,i,j,b] (mpo.py:26-45)."""
    A, B = K.as_dev(A), K.as_dev(B)
    a, i, j, b = A.shape
    c, j2, d = B.shape
    if j2 != j:
        raise ValueError("MPO and MPS physical dimensions differ")
    out = K.local(B, _apply_op(A), c, j, d, a * i, b)   # [c, (a,i), d, b]
    return out.view(c * a, i, d * b)


class MPO(TensorArray):
    """Matrix product operator W[a, i_out, j_in, b] (mpo.py:48-75)."""

    __array_priority__ = 10000

    def __init__(self, data, strategy: Strategy = DEFAULT_STRATEGY):
        super().__init__(data)
        self.strategy = strategy

    def copy(self) -> "MPO":
        return MPO(self._data, self.strategy)

    @classmethod
    def from_local_operators(cls, operators, strategy: Strategy = DEFAULT_STRATEGY) -> "MPO":
        return MPO([np.asarray(o).reshape((1,) + np.asarray(o).shape + (1,)) for o in operators],
                   strategy)

    def __mul__(self, n) -> "MPO":
        """n * self: the modulus spread evenly over the sites, the phase on site 0 (mpo.py:103-120)."""
        if isinstance(n, (int, float, complex)):
            absn = abs(n)
            if absn:
                phase = n / absn
                factor = np.exp(np.log(absn) / self.size)
            else:
                phase = 0.0
                factor = 0.0
            return MPO([K.scale(factor if i > 0 else factor * phase, A)
                        for i, A in enumerate(self._data)], self.strategy)
        raise TypeError(f"Invalid operation * between {type(self)} and {type(n)}")

    __rmul__ = __mul__

    def __truediv__(self, n) -> "MPO":
        if isinstance(n, (int, float, complex)):
            return self.__mul__(1.0 / n)
        raise TypeError(f"Invalid operation / between {type(self)} and {type(n)}")

    def physical_dimensions(self) -> list[int]:
        return [A.shape[2] for A in self._data]

    def bond_dimensions(self) -> list[int]:
        return [A.shape[-1] for A in self._data][:-1]

    def max_bond_dimension(self) -> int:
        return max(A.shape[0] for A in self._data)

    @property
    def T(self) -> "MPO":
        return MPO([K.permute(A, (0, 2, 1, 3)) for A in self._data], self.strategy)

    def conj(self) -> "MPO":
        return MPO([K.conj(A) for A in self._data], self.strategy)

    def to_matrix(self) -> np.ndarray:
        """Dense operator (host), contracted on the device (mpo.py:192-207)."""
        out = None
        Di = Dj = 1
        for A in self._data:
            _, i, j, b = A.shape
            if out is None:
                out = A.reshape(i, j, b)
            else:
                x, y, a = out.shape
                # [x, y, a] [a, i, j, b] -> [x, y, i, j, b] -> [x, i, y, j, b]
                t = K.gemm(out.reshape(x * y, a), A.reshape(a, i * j * b)).view(x, y, i, j, b)
                out = K.permute(t, (0, 2, 1, 3, 4)).view(x * i, y * j, b)
            Di *= i
            Dj *= j
        return K.to_numpy(out)[:, :, 0]

    def set_strategy(self, strategy: Strategy) -> "MPO":
        return MPO(self._data, strategy)

    def apply(self, state, strategy: Strategy | None = None, simplify: bool | None = None):
        """MPO x MPS (or MPSSum), then optional simplify (mpo.py:229-281)."""
        from ..state.simplification import simplify as simplify_mps

        if strategy is None:
            strategy = self.strategy
        if strategy is None:
            raise RuntimeError("MPO has no truncation strategy")
        if simplify is None:
            simplify = strategy.get_simplify_flag()
        if isinstance(state, MPSSum):
            if self.size != state.size:
                raise ValueError("MPO and state must have the same size")
            out = None
            for i, (w, mps) in enumerate(zip(state.weights, state.states)):
                Ostate = w * MPS([_mpo_multiply_tensor(A, B) for A, B in zip(self._data, mps)],
                                 error=mps.error())
                out = Ostate if i == 0 else out + Ostate
            state = out
        elif isinstance(state, MPS):
            if self.size != state.size:
                raise ValueError("MPO and state must have the same size")
            state = MPS([_mpo_multiply_tensor(A, B) for A, B in zip(self._data, state)],
                        error=state.error())
        else:
            raise TypeError(f"Cannot multiply MPO with {state}")
        if simplify:
            state = simplify_mps(state, strategy=strategy)
        return state

    def __matmul__(self, b):
        if isinstance(b, (MPS, MPSSum)):
            return self.apply(b)
        raise TypeError(f"Cannot multiply MPO with {b}")

    def __rmatmul__(self, b):
        if isinstance(b, (MPS, MPSSum)):
            return self.T.apply(b)
        raise TypeError(f"Cannot multiply {b} with MPO")

    def expectation(self, bra: MPS, ket: MPS | None = None) -> complex:
        """<bra|O|ket> via MPO environments split at the bra's centre (mpo.py:378-416)."""
        if isinstance(bra, CanonicalMPS):
            center = bra.center
        elif isinstance(bra, MPS):
            center = self.size - 1
        else:
            raise TypeError("MPS required")
        if ket is None:
            ket = bra
        elif not isinstance(ket, MPS):
            raise TypeError("MPS required")
        left = right = begin_mpo_environment()
        for i in range(0, center):
            left = update_left_mpo_environment(left, bra[i], self[i], ket[i])
        for i in range(self.size - 1, center - 1, -1):
            right = update_right_mpo_environment(right, bra[i], self[i], ket[i])
        return join_mpo_environments(left, right)

    def reverse(self) -> "MPO":
        return MPO([K.permute(A, (3, 1, 2, 0)) for A in reversed(self._data)], self.strategy)

    def __add__(self, A):
        """self + A as an MPOSum (mpo.py:87-101)."""
        from .mposum import MPOSum

        if isinstance(A, MPO):
            return MPOSum([self, A], [1.0, 1.0], self.strategy)
        if isinstance(A, MPOSum):
            return MPOSum([self] + A.mpos, [1.0] + A.weights, self.strategy)
        raise TypeError(f"Cannot add MPO and {type(A)}")

    def __sub__(self, A):
        from .mposum import MPOSum

        if isinstance(A, MPO):
            return MPOSum([self, A], [1.0, -1.0], self.strategy)
        if isinstance(A, MPOSum):
            return MPOSum([self] + A.mpos, [1.0] + [-w for w in A.weights], self.strategy)
        raise TypeError(f"Cannot subtract {type(A)} from MPO")

    def extend(self, L: int, sites=None, dimensions=2) -> "MPO":
        """The operator on L sites: this MPO's tensors at `sites` (default the
        first ones), identities of the given physical dimensions elsewhere,
        each carrying the bond of the tensor to its left (mpo.py:326-376)."""
        if isinstance(dimensions, int):
            extra = [dimensions] * max(L - self.size, 0)
        else:
            extra = list(dimensions)
            if len(extra) != L - self.size:
                raise ValueError("len(dimensions) must equal L - self.size")
        if L < self.size:
            raise ValueError("New size L must be at least the current size")
        sites = list(range(self.size)) if sites is None else list(sites)
        if len(sites) != self.size:
            raise ValueError("len(sites) must equal the current size")
        placed = dict(zip(sites, self._data))
        data, D, k = [], 1, 0
        for i in range(L):
            if i in placed:
                A = placed[i]
                D = A.shape[-1]
            else:
                d = extra[k]
                k += 1
                A = np.einsum("ab,ij->aijb", np.eye(D), np.eye(d)).astype(np.complex128)
            data.append(A)
        return MPO(data, strategy=self.strategy)


__all__ = ["MPO", "_mpo_multiply_tensor"]
