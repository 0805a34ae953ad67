"""TEST DOUBLE: NumPy stand-ins for ``seemps2_amd.kernels`` on CPU tensors.

Used only by tests/test_host_logic.py to exercise the drop-in API's host-side
control flow (shapes, sweeps, error bookkeeping) on machines without a GPU.
It is installed by monkeypatching inside that test; the product package never
imports it and has no CPU fallback of its own.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg
import torch

from oracle import core as oc, algorithms as oa

CDTYPE = torch.complex128


def _np(x):
    return x.detach().numpy()


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex128)))


def as_dev(x):
    if isinstance(x, torch.Tensor):
        return x.to(CDTYPE).contiguous()
    return _t(np.asarray(x))


def to_numpy(x):
    return _np(x)


def empty(*shape):
    return torch.empty(*shape, dtype=CDTYPE)


def zeros(*shape):
    return torch.zeros(*shape, dtype=CDTYPE)


def fill_zero(t):
    t.zero_()
    return t


_OPS = {0: lambda x: x, 1: lambda x: x.T, 2: lambda x: x.conj().T, 3: lambda x: x.conj()}


def gemm(A, B, opA=0, opB=0, alpha=1.0, beta=0.0, C=None):
    r = complex(alpha) * (_OPS[opA](_np(A)) @ _OPS[opB](_np(B)))
    if C is None:
        return _t(r)
    C.copy_(_t(r + complex(beta) * _np(C)) if beta != 0 else _t(r))
    return C


def local(x, Mop, P, I, Q, J1, J2, out=None, accumulate=False):
    r = np.einsum("piq,ijk->pjqk", _np(x).reshape(P, I, Q), _np(Mop).reshape(I, J1, J2))
    if out is None:
        return _t(r)
    if accumulate:
        r = r + _np(out).reshape(r.shape)
    out.view(-1).copy_(_t(r).view(-1))
    return out


def axpby(a, x, b, y, conj_x=False):
    xv = 0 if x is None else (np.conj(_np(x)) if conj_x else _np(x))
    r = complex(a) * xv + (complex(b) * _np(y) if b != 0 else 0)
    y.copy_(_t(np.broadcast_to(r, y.shape)))
    return y


def scale(a, x):
    return _t(complex(a) * _np(x))


def copy(x):
    return x.clone()


def conj(x):
    return _t(np.conj(_np(x)))


def scale2d(src, rows, cols, ld_src, sr=None, sc=None, out=None):
    s = _np(src).reshape(-1)
    M = np.stack([s[r * ld_src: r * ld_src + cols] for r in range(rows)])
    if sr is not None:
        M = M * sr.numpy()[:rows, None]
    if sc is not None:
        M = M * sc.numpy()[None, :cols]
    if out is None:
        return _t(M)
    out.view(rows, cols).copy_(_t(M))
    return out


def embed_block(B, A, l_off, r_off):
    a, d, b = A.shape
    B[l_off:l_off + a, :, r_off:r_off + b] = A


def site_product(A, B):
    A, B = _np(A), _np(B)
    a, d, b = A.shape
    c, _, e = B.shape
    return _t((A[:, None, :, :, None] * B[None, :, :, None, :]).reshape(a * c, d, b * e))


def dotc_dev(x, y):
    return _t(np.array([np.vdot(_np(x), _np(y))]))


def dotc(x, y):
    return complex(np.vdot(_np(x), _np(y)))


def truncate(s, method, tol, max_bond):
    st = oc.Strategy(method=method, tolerance=tol, max_bond_dimension=max_bond)
    if tol == 0 and method != 0:
        st.method = 3
    return oc.truncate(s.numpy(), st)


def svd(A):
    U, s, Vh = scipy.linalg.svd(_np(A), full_matrices=False)
    return _t(U), torch.from_numpy(s.copy()), _t(Vh)


def svd_split(AA, m, n, mode, method, tol, max_bond, want_s=False):
    U, s, Vh = scipy.linalg.svd(_np(AA).reshape(m, n), full_matrices=False)
    D, err = truncate(torch.from_numpy(s), method, tol, max_bond)
    if mode == 0:
        A, B = U[:, :D], s[:D, None] * Vh[:D]
    else:
        A, B = U[:, :D] * s[:D], Vh[:D]
    return _t(A), _t(B), D, err, torch.from_numpy(s[:D].copy()) if want_s else None


def permute(x, perm):
    return _t(np.transpose(_np(x), perm))


def two_site_operator(O1, O2):
    H = np.tensordot(_np(O1), _np(O2), axes=1)  # [c,ko,k,lo,l,d]
    c, ko, k, lo, l, d = H.shape
    return _t(H.transpose(0, 2, 4, 1, 3, 5).reshape(c * k * l, ko * lo * d))


def _h12(Hm, c, d1, d2, w2):
    return _np(Hm).reshape(c, d1, d2, d1, d2, w2).transpose(0, 3, 1, 4, 2, 5)


def two_site_matvec(L, Hm, R, v, d1, d2):
    c = L.shape[1]
    w2 = R.shape[1]
    return _t(oa.two_site_matvec(_np(L), _h12(Hm, c, d1, d2, w2), _np(R), _np(v)))


def eigensolve(L, Hm, R, v0, d1, d2, ncv=20, tol=0.0, maxiter=1000, precision="fp64"):
    c, w2 = L.shape[1], R.shape[1]
    H12 = _h12(Hm, c, d1, d2, w2)
    shape = (L.shape[0], d1, d2, R.shape[0])
    n = int(np.prod(shape))
    Hd = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros(n, complex)
        e[k] = 1
        Hd[:, k] = oa.two_site_matvec(_np(L), H12, _np(R), e.reshape(shape)).reshape(-1)
    w, V = np.linalg.eigh(0.5 * (Hd + Hd.conj().T))
    return float(w[0]), _t(V[:, 0].reshape(shape)), {"residual": 0.0, "matvecs": n,
                                                     "restarts": 1, "matvecs_fp32": 0}
