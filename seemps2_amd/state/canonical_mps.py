"""CanonicalMPS on the device (reference: src/seemps/state/canonical_mps.py:39-428)."""

from __future__ import annotations

import math
import warnings
from collections.abc import Iterable

import numpy as np

from .. import kernels as K
from ..core import (
    DEFAULT_STRATEGY,
    Strategy,
    _begin_environment,
    _canonicalize,
    _left_orth_2site,
    _recanonicalize,
    _right_orth_2site,
    _update_in_canonical_form_left,
    _update_in_canonical_form_right,
    _update_left_environment,
    _update_right_environment,
)
from .mps import MPS


def _norm(A) -> float:
    return math.sqrt(max(K.dotc(A, A).real, 0.0))


class CanonicalMPS(MPS):
    """MPS in canonical form with respect to ``center`` (canonical_mps.py:39-108).

    Same constructor semantics as the reference: an MPS input is canonicalised
    with ``strategy`` and its error is NOT inherited (``error`` argument only);
    a CanonicalMPS input keeps its error and is re-centred.
    """

    def __init__(self, data: Iterable, center: int | None = None,
                 normalize: bool | None = None, strategy: Strategy = DEFAULT_STRATEGY,
                 is_canonical: bool = False, error: float = 0):
        super().__init__(data, error)
        self.strategy = strategy
        if isinstance(data, CanonicalMPS):
            actual_center = self._center = data.center
            self._error = data._error
            if center is not None:
                actual_center = center
                self.recenter(actual_center)
        else:
            self._center = actual_center = self._interpret_center(0 if center is None else center)
            if not is_canonical:
                self._error += _canonicalize(self._data, actual_center, self.strategy)
        if normalize is True or (normalize is None and self.strategy.get_normalize_flag()):
            A = self[actual_center]
            N = _norm(A)
            if N:
                self._data[actual_center] = K.scale(1.0 / N, A)
            else:
                warnings.warn("Refusing to noramlize zero vector")

    @property
    def center(self) -> int:
        return self._center

    @classmethod
    def from_vector(cls, psi, dimensions, strategy: Strategy = DEFAULT_STRATEGY,
                    normalize: bool = True, center: int = 0) -> "CanonicalMPS":
        from .schmidt import _vector2mps

        data, error = _vector2mps(psi, dimensions, strategy, normalize, center)
        return CanonicalMPS(data, error=error, center=center, is_canonical=True)

    def zero_state(self) -> "CanonicalMPS":
        return CanonicalMPS([np.zeros((1, A.shape[1], 1)) for A in self._data], error=0.0,
                            center=0, is_canonical=True)

    def norm_squared(self) -> float:
        """vdot of the centre tensor (canonical_mps.py:161-165)."""
        A = self._data[self.center]
        return K.dotc(A, A).real

    def left_environment(self, site: int):
        start = min(site, self.center)
        rho = _begin_environment(self[start].shape[0])
        for A in self._data[start:site]:
            rho = _update_left_environment(A, A, rho)
        return rho

    def right_environment(self, site: int):
        start = max(site, self.center)
        rho = _begin_environment(self[start].shape[-1])
        for A in self._data[start:site:-1]:
            rho = _update_right_environment(A, A, rho)
        return rho

    def Schmidt_weights(self, site: int | None = None) -> np.ndarray:
        """Normalised squared singular values across the centre (canonical_mps.py:225-247)."""
        site = self.center if site is None else self._interpret_center(site)
        if site != self.center:
            return self.copy().recenter(site).Schmidt_weights()
        A = self._data[site]
        d1, d2, d3 = A.shape
        _, s, _ = K.svd(A.reshape(d1 * d2, d3))
        s = s.cpu().numpy()
        s = s * s
        return s / np.sum(s)

    def entanglement_entropy(self, site: int | None = None) -> float:
        s = self.Schmidt_weights(site)
        return -np.sum(s * np.log2(s))

    def Renyi_entropy(self, site: int | None = None, alpha: float = 2.0) -> float:
        s = self.Schmidt_weights(site)
        if alpha < 0:
            raise ValueError("Invalid Renyi entropy power")
        if alpha == 0:
            alpha = 1e-9
        elif alpha == 1:
            alpha = 1 - 1e-9
        return np.log(np.sum(s**alpha)) / (1 - alpha)

    def update_canonical(self, A, direction: int, truncation: Strategy) -> None:
        if direction > 0:
            self._center, error = _update_in_canonical_form_right(self._data, A, self.center,
                                                                  truncation)
        else:
            self._center, error = _update_in_canonical_form_left(self._data, A, self.center,
                                                                 truncation)
        self._error += error

    def update_2site_right(self, AA, site: int, strategy: Strategy) -> None:
        """Left-orthogonal split of AA into sites (site, site+1) (canonical_mps.py:326-345)."""
        # with the centre at `site`, site + 1 is right-canonical: a sketch basis
        basis = self._data[site + 1] if self._center == site else None
        self._data[site], self._data[site + 1], error = _left_orth_2site(AA, strategy, basis)
        self._center = site + 1
        self._error += error

    def update_2site_left(self, AA, site: int, strategy: Strategy) -> None:
        """Right-orthogonal split of AA into sites (site, site+1) (canonical_mps.py:347-366)."""
        # with the centre at `site + 1`, site is left-canonical: a sketch basis
        basis = self._data[site] if self._center == site + 1 else None
        self._data[site], self._data[site + 1], error = _right_orth_2site(AA, strategy, basis)
        self._center = site
        self._error += error

    def _interpret_center(self, center: int) -> int:
        size = self.size
        if 0 <= center < size:
            return center
        center += size
        if 0 <= center < size:
            return center
        raise IndexError()

    def recenter(self, center: int, strategy: Strategy | None = None) -> "CanonicalMPS":
        newcenter = self._interpret_center(center)
        oldcenter = self.center
        if newcenter != oldcenter:
            self._error += _recanonicalize(self._data, oldcenter, newcenter,
                                           self.strategy if strategy is None else strategy)
            self._center = newcenter
        return self

    def normalize_inplace(self) -> "CanonicalMPS":
        n = self.center
        A = self._data[n]
        N = _norm(A)
        if N:
            self._data[n] = K.scale(1.0 / N, A)
        return self

    def __copy__(self):
        return type(self)(self, center=self.center, strategy=self.strategy, error=self._error)

    def copy(self):
        return self.__copy__()
