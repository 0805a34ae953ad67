"""MPS and MPO environments on the device (reference: src/seemps/state/environments.py).

MPO environments ``L[bra, mpo, ket]`` / ``R[bra, mpo, ket]`` are updated with
two complex GEMMs on the FP64 matrix cores and one local-operator contraction
over the small (mpo bond x physical) index pair, never materialising a
transposed copy of the large tensors.
"""

from __future__ import annotations

import numpy as np

from .. import kernels as K
from .._lib import OP_C, OP_H, OP_T
from ..core import (
    _begin_environment,
    _end_environment,
    _join_environments,
    _update_left_environment,
    _update_right_environment,
    scprod,
    vdot,
)


def _left_op(O):
    # Mop[(c,i), (j,e)] = O[c,j,i,e], materialised once per MPO tensor
    return K.permuted(O, (0, 2, 1, 3))


def _right_op(O):
    # Mop[(j,e), (c,i)] = O[c,j,i,e], materialised once per MPO tensor
    return K.permuted(O, (1, 3, 0, 2))


def begin_mpo_environment():
    return K.as_dev(np.ones((1, 1, 1)))


def update_left_mpo_environment(rho, A, O, B):
    """L'[d,e,f] = sum rho[a,c,b] A*[a,j,d] O[c,j,i,e] B[b,i,f] (environments.py:19-30)."""
    rho, A, O, B = K.as_dev(rho), K.as_dev(A), K.as_dev(O), K.as_dev(B)
    a, c, b = rho.shape
    _, j, d = A.shape
    c2, jo, i, e = O.shape
    _, i2, f = B.shape
    T1 = K.gemm(rho.reshape(a * c, b), B.reshape(b, i * f))          # [a, (c,i), f]
    T2 = K.local(T1, _left_op(O), a, c * i, f, jo * e, 1)             # [a, (j,e), f]
    return K.gemm(A.reshape(a * j, d), T2.view(a * j, e * f), opA=OP_H).view(d, e, f)


def update_right_mpo_environment(rho, A, O, B):
    """R'[a,c,b] = sum rho[d,e,f] A*[a,j,d] O[c,j,i,e] B[b,i,f] (environments.py:33-42)."""
    rho, A, O, B = K.as_dev(rho), K.as_dev(A), K.as_dev(O), K.as_dev(B)
    d, e, f = rho.shape
    a, j, _ = A.shape
    c, jo, i, e2 = O.shape
    b, i2, _ = B.shape
    T1 = K.gemm(A.reshape(a * j, d), rho.reshape(d, e * f), opA=OP_C)  # [a, (j,e), f]
    T2 = K.local(T1, _right_op(O), a, j * e, f, c * i, 1)              # [a, (c,i), f]
    return K.gemm(T2.view(a * c, i * f), B.reshape(b, i * f), opB=OP_T).view(a, c, b)


def begin_mpo2_environment():
    return K.as_dev(np.ones((1, 1, 1, 1)))


def update_left_mpo2_environment(E, A, O):
    """E'[a',w1',w2',b'] = sum E[a,w1,w2,b] conj(A[a,i,a']) conj(O[w1,k,i,w1'])
    O[w2,k,j,w2'] A[b,j,b']: the left environment of <O psi|O psi> with both
    MPO layers (no counterpart in the reference, which applies O and folds the
    MPS product of bond w chi).  Two GEMMs of chi^3 d w^2 complex MACs around
    two small local contractions."""
    E, A, O = K.as_dev(E), K.as_dev(A), K.as_dev(O)
    a, w1, w2, b = E.shape
    _, j, b2 = A.shape
    c, k, i, e = O.shape
    T1 = K.gemm(E.reshape(a * w1 * w2, b), A.reshape(b, j * b2))             # [a,w1,w2,j,b']
    T2 = K.local(T1, _left_op(O), a * w1, w2 * j, b2, k * e, 1)               # [a,w1,k,w2',b']
    T3 = K.local(T2, K.conj(O).view(c * k, i * e), a, w1 * k, e * b2, i * e, 1)  # [a,i,w1',w2',b']
    return K.gemm(A.reshape(a * i, A.shape[2]), T3.view(a * i, e * e * b2),
                  opA=OP_H).view(A.shape[2], e, e, b2)


def end_mpo_environment(rho) -> complex:
    return complex(rho.reshape(-1)[0].item())


def join_mpo_environments(left, right) -> complex:
    """sum_k left[k] right[k] (environments.py:49-51)."""
    return K.dotc(K.conj(left.reshape(-1)), right.reshape(-1))


__all__ = [
    "_begin_environment",
    "_update_left_environment",
    "_update_right_environment",
    "_end_environment",
    "_join_environments",
    "scprod",
    "vdot",
    "begin_mpo_environment",
    "begin_mpo2_environment",
    "update_left_mpo2_environment",
    "update_left_mpo_environment",
    "update_right_mpo_environment",
    "end_mpo_environment",
    "join_mpo_environments",
]
