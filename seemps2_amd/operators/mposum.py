"""Linear combinations of MPOs (reference: src/seemps/operators/mposum.py:11-262).

``MPOSum.join`` builds one MPO whose tensors hold the terms' tensors as
blocks: weighted along the right bond of the first site, stacked along the
left bond of the last one, block-diagonal in between.  The blocks are
written on the device by the strided copy kernel (``seemps_zscale2d``).
``apply`` applies every term and adds the results as an MPSSum, simplified
with the sum's strategy unless told otherwise, as the reference does.
"""

from __future__ import annotations

import numpy as np

from .. import kernels as K
from ..core import DEFAULT_STRATEGY, Strategy
from ..state import MPS, MPSSum
from .mpo import MPO


class MPOSum:
    """sum_k weights[k] * mpos[k] (mposum.py:11-58)."""

    __array_priority__ = 10000

    def __init__(self, mpos, weights=None, strategy: Strategy = DEFAULT_STRATEGY):
        self.mpos = list(mpos)
        if len(self.mpos) < 1:
            raise ValueError("MPOSum requires at least one MPO")
        self.size = self.mpos[0].size
        self.weights = [1.0] * len(self.mpos) if weights is None else list(weights)
        self.strategy = strategy

    def physical_dimensions(self) -> list[int]:
        return self.mpos[0].physical_dimensions()

    def copy(self) -> "MPOSum":
        return MPOSum(self.mpos, self.weights, self.strategy)

    def __add__(self, A):
        if isinstance(A, MPO):
            return MPOSum(self.mpos + [A], self.weights + [1.0], self.strategy)
        if isinstance(A, MPOSum):
            return MPOSum(self.mpos + A.mpos, self.weights + A.weights, self.strategy)
        raise TypeError(f"Cannot add an MPOSum to an object of type {type(A)}")

    def __sub__(self, A):
        if isinstance(A, MPO):
            return MPOSum(self.mpos + [A], self.weights + [-1.0], self.strategy)
        if isinstance(A, MPOSum):
            return MPOSum(self.mpos + A.mpos, self.weights + [-w for w in A.weights],
                          self.strategy)
        raise TypeError(f"Cannot subtract an object of type {type(A)} from an MPOSum")

    def __mul__(self, n) -> "MPOSum":
        if not isinstance(n, (int, float, complex, np.number)):
            raise TypeError(f"Cannot multiply MPOSum by {n}")
        return MPOSum(self.mpos, [n * w for w in self.weights], self.strategy)

    __rmul__ = __mul__

    def __truediv__(self, n) -> "MPOSum":
        if isinstance(n, (int, float, complex, np.number)):
            return self.__mul__(1.0 / n)
        raise TypeError(f"Cannot divide MPOSum by {n}")

    @property
    def T(self) -> "MPOSum":
        return MPOSum([A.T for A in self.mpos], self.weights, self.strategy)

    def conj(self) -> "MPOSum":
        return MPOSum([A.conj() for A in self.mpos], [np.conj(w) for w in self.weights],
                      self.strategy)

    def to_matrix(self) -> np.ndarray:
        out = self.weights[0] * self.mpos[0].to_matrix()
        for w, A in zip(self.weights[1:], self.mpos[1:]):
            out = out + w * A.to_matrix()
        return out

    def set_strategy(self, strategy: Strategy, strategy_components: Strategy | None = None):
        mpos = self.mpos if strategy_components is None else \
            [A.set_strategy(strategy_components) for A in self.mpos]
        return MPOSum(mpos, self.weights, strategy)

    def extend(self, L: int, sites=None, dimensions=2) -> "MPOSum":
        return MPOSum([A.extend(L, sites=sites, dimensions=dimensions) for A in self.mpos],
                      self.weights, self.strategy)

    def _joined_tensor(self, i: int):
        As = [A[i] for A in self.mpos]
        L = self.size
        _, d, e, _ = As[0].shape
        if i == 0:
            DR = sum(A.shape[3] for A in As)
            out = K.empty(1, d * e, DR)
            off = 0
            for w, A in zip(self.weights, As):
                b = A.shape[3]
                K.embed_block(out, K.scale(complex(w), A).reshape(1, d * e, b), 0, off)
                off += b
            return out.view(1, d, e, DR)
        if i == L - 1:
            DL = sum(A.shape[0] for A in As)
            out = K.empty(DL, d * e, 1)
            off = 0
            for A in As:
                a = A.shape[0]
                K.embed_block(out, A.reshape(a, d * e, 1), off, 0)
                off += a
            return out.view(DL, d, e, 1)
        DL = sum(A.shape[0] for A in As)
        DR = sum(A.shape[3] for A in As)
        out = K.zeros(DL, d * e, DR)
        l = r = 0
        for A in As:
            a, _, _, b = A.shape
            K.embed_block(out, A.reshape(a, d * e, b), l, r)
            l += a
            r += b
        return out.view(DL, d, e, DR)

    def join(self, strategy: Strategy | None = None) -> MPO:
        """One MPO with the terms' tensors as blocks (mposum.py:221-262)."""
        return MPO([self._joined_tensor(i) for i in range(self.size)],
                   strategy=self.strategy if strategy is None else strategy)

    def apply(self, state, strategy: Strategy | None = None, simplify: bool | None = None):
        """sum_k w_k O_k state, simplified by default (mposum.py:166-187)."""
        from ..state.simplification import simplify as simplify_mps

        terms = [w * O.apply(state, simplify=False) for w, O in zip(self.weights, self.mpos)]
        out = MPSSum([1.0] * len(terms), terms, check_args=False)
        if strategy is None:
            strategy = self.strategy
        if simplify is None:
            simplify = strategy.get_simplify_flag()
        return simplify_mps(out, strategy=strategy) if simplify else out

    def __matmul__(self, b):
        if isinstance(b, (MPS, MPSSum)):
            return self.apply(b)
        raise TypeError(f"Cannot multiply MPOSum with {b}")


__all__ = ["MPOSum"]
