"""Vector -> MPS by successive Schmidt splits (reference: src/seemps/state/schmidt.py:39-81)."""

from __future__ import annotations

import math

import numpy as np

from .. import kernels as K
from ..core import DEFAULT_STRATEGY, Strategy, _left_orth_2site, _right_orth_2site


def _vector2mps(state, dimensions, strategy: Strategy = DEFAULT_STRATEGY,
                normalize: bool = True, center: int = -1):
    psi = K.as_dev(np.asarray(state)).reshape(1, -1, 1)
    L = len(dimensions)
    if math.prod(dimensions) != psi.numel():
        raise ValueError("Wrong dimensions specified when converting a vector to MPS")
    output = [psi] * L
    if center < 0:
        center = L + center
    if center < 0 or center >= L:
        raise ValueError("Invalid value of center in _vector2mps")
    err = 0.0
    for i in range(center):
        output[i], psi, new_err = _left_orth_2site(
            psi.reshape(psi.shape[0], dimensions[i], -1, psi.shape[-1]), strategy)
        err += new_err
    for i in range(L - 1, center, -1):
        psi, output[i], new_err = _right_orth_2site(
            psi.reshape(psi.shape[0], -1, dimensions[i], psi.shape[-1]), strategy)
        err += new_err
    if normalize:
        N = math.sqrt(K.dotc(psi, psi).real)
        psi = K.scale(1.0 / N, psi)
        err /= N
    output[center] = psi
    return output, err
