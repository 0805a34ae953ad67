"""Python-facing wrappers of the C-ABI kernels on torch device tensors.

Every function enqueues HIP kernels from ``libseemps_hip.so`` on the current
stream.  Tensors are complex128 (FP64) and contiguous; real or complex64
inputs are promoted on upload, matching the reference's promotion rules
(src/seemps/cython/gemm.pxi:29-58, svd.pxi:74-77).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, stream, workspace

CDTYPE = torch.complex128


def as_dev(x) -> torch.Tensor:
    """Upload / convert to a contiguous complex128 tensor on the current device."""
    if isinstance(x, torch.Tensor):
        if x.device.type != "cuda":
            x = x.to(_lib.device())
        if x.dtype != CDTYPE:
            x = x.to(CDTYPE)
        return x.contiguous()
    arr = np.asarray(x)
    t = torch.from_numpy(np.ascontiguousarray(arr.astype(np.complex128, copy=False)))
    return t.to(_lib.device(), non_blocking=False)


def need(*tensors: torch.Tensor | None) -> None:
    """Raise ValueError unless every tensor (None is skipped) is a contiguous
    complex128 tensor on a HIP device: the kernels read raw device pointers
    with row-major complex128 strides, so anything else must be converted
    (:func:`as_dev`) before a launch."""
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"expected a device tensor, got {type(t).__name__}")
        if t.device.type != "cuda":
            raise ValueError("expected a tensor on the HIP device, got one on "
                             f"{t.device.type}")
        if t.dtype != CDTYPE:
            raise ValueError(f"expected a complex128 tensor, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError("expected a contiguous tensor")


def to_numpy(x: torch.Tensor) -> np.ndarray:
    return x.detach().cpu().numpy()


def empty(*shape) -> torch.Tensor:
    return torch.empty(*shape, dtype=CDTYPE, device=_lib.device())


def zeros(*shape) -> torch.Tensor:
    t = empty(*shape)
    fill_zero(t)
    return t


def fill_zero(t: torch.Tensor) -> torch.Tensor:
    check(lib.seemps_zaxpby(t.numel(), 0.0, 0.0, None, 0, 0.0, 0.0, ptr(t), stream()))
    return t


PRECISIONS = {"fp64": 0, "fp32": 1, "mixed": 2}


def gemm(A: torch.Tensor, B: torch.Tensor, opA: int = 0, opB: int = 0, alpha=1.0,
         beta=0.0, C: torch.Tensor | None = None, precision: str = "fp64") -> torch.Tensor:
    """C = alpha op(A) op(B) + beta C for 2-D contiguous complex128 tensors
    (``precision`` "fp32": FP32 products and accumulation)."""
    need(A, B, C)
    if A.dim() != 2 or B.dim() != 2:
        raise ValueError("gemm expects matrices")
    M = A.shape[1] if opA in (1, 2) else A.shape[0]
    K = A.shape[0] if opA in (1, 2) else A.shape[1]
    K2 = B.shape[1] if opB in (1, 2) else B.shape[0]
    N = B.shape[0] if opB in (1, 2) else B.shape[1]
    if K != K2:
        raise ValueError(f"gemm: inner dimensions differ ({K} vs {K2})")
    if C is None:
        C = empty(M, N)
        beta = 0.0
    alpha, beta = complex(alpha), complex(beta)
    if precision == "fp64":
        # split K over the workspace when the output has too few tiles
        wb = lib.seemps_zgemm_workspace(M, N, K)
        if wb > 0:
            wp, wn = workspace.get(wb)
            check(lib.seemps_zgemm_ws(opA, opB, M, N, K, alpha.real, alpha.imag, ptr(A),
                                      A.shape[1], ptr(B), B.shape[1], beta.real, beta.imag,
                                      ptr(C), N, wp, wn, stream()))
            return C
    check(lib.seemps_zgemm_prec(PRECISIONS[precision], opA, opB, M, N, K, alpha.real,
                                alpha.imag, ptr(A), A.shape[1], 0, ptr(B), B.shape[1], 0,
                                beta.real, beta.imag, ptr(C), N, 0, 1, stream()))
    return C


def local(x: torch.Tensor, Mop: torch.Tensor, P: int, I: int, Q: int, J1: int, J2: int,
          out: torch.Tensor | None = None, accumulate: bool = False) -> torch.Tensor:
    """out[p, j1, q, j2] (+)= sum_i x[p, i, q] Mop[i, j1, j2]."""
    need(x, Mop, out)
    if x.numel() != P * I * Q or Mop.numel() != I * J1 * J2:
        raise ValueError("local: operand sizes do not match the dimensions")
    if out is None:
        out = empty(P, J1, Q, J2)
    check(lib.seemps_zlocal(P, I, Q, J1, J2, ptr(x), ptr(Mop), ptr(out), int(accumulate),
                            stream()))
    return out


def axpby(a, x: torch.Tensor | None, b, y: torch.Tensor, conj_x: bool = False) -> torch.Tensor:
    """y = a op(x) + b y in place."""
    need(x, y)
    if x is not None and x.numel() != y.numel():
        raise ValueError("axpby: sizes differ")
    a, b = complex(a), complex(b)
    check(lib.seemps_zaxpby(y.numel(), a.real, a.imag, ptr(x), int(conj_x), b.real, b.imag,
                            ptr(y), stream()))
    return y


def scale(a, x: torch.Tensor) -> torch.Tensor:
    """New tensor a * x."""
    y = empty(*x.shape)
    return axpby(a, x, 0.0, y)


def copy(x: torch.Tensor) -> torch.Tensor:
    return scale(1.0, x)


def conj(x: torch.Tensor) -> torch.Tensor:
    y = empty(*x.shape)
    return axpby(1.0, x, 0.0, y, conj_x=True)


def scale2d(src: torch.Tensor, rows: int, cols: int, ld_src: int, sr=None, sc=None,
            out: torch.Tensor | None = None) -> torch.Tensor:
    if out is None:
        out = empty(rows, cols)
    check(lib.seemps_zscale2d(rows, cols, ptr(sr), ptr(sc), ptr(src), ld_src, ptr(out), cols,
                              stream()))
    return out


def embed_block(B: torch.Tensor, A: torch.Tensor, l_off: int, r_off: int) -> None:
    """B[l_off:l_off+a, :, r_off:r_off+b] = A for B (DL, d, DR), A (a, d, b)."""
    DL, d, DR = B.shape
    a, d2, b = A.shape
    if d2 != d or l_off + a > DL or r_off + b > DR:
        raise ValueError("embed_block: block does not fit")
    dst = B.data_ptr() + 16 * (l_off * d * DR + r_off)
    check(lib.seemps_zscale2d(a * d, b, None, None, ptr(A), b, dst, DR, stream()))


def site_product(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """out[(x,y), k, (u,v)] = A[x,k,u] B[y,k,v] for MPS tensors A (a,d,b), B (c,d,e)."""
    need(A, B)
    a, d, b = A.shape
    c, d2, e = B.shape
    if d2 != d:
        raise ValueError("site_product: physical dimensions differ")
    out = empty(a * c, d, b * e)
    check(lib.seemps_zsite_product(a, d, b, c, e, ptr(A), ptr(B), ptr(out), stream()))
    return out


def dotc_dev(x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """Device scalar sum(conj(x) * y)."""
    need(x, y)
    n = x.numel()
    if y.numel() != n:
        raise ValueError("dotc: sizes differ")
    out = empty(1)
    wb = lib.seemps_reduce_workspace(n)
    wp, wn = workspace.get(wb)
    check(lib.seemps_zdotc(n, ptr(x), ptr(y), ptr(out), wp, wn, stream()))
    return out


def dotc(x: torch.Tensor, y: torch.Tensor) -> complex:
    return complex(dotc_dev(x, y).item())


def truncate(s: torch.Tensor, method: int, tol: float, max_bond: int) -> tuple[int, float]:
    """Kept size and squared error of the Strategy rule on a device spectrum."""
    n = s.numel()
    buf = torch.empty(2, dtype=torch.float64, device=s.device)
    Dbuf = torch.empty(1, dtype=torch.int64, device=s.device)
    check(lib.seemps_truncate(n, ptr(s), method, tol, max_bond, ptr(Dbuf), ptr(buf), stream()))
    return int(Dbuf.item()), float(buf[0].item())


def svd(A: torch.Tensor):
    """Economy SVD of a device matrix: (U, S, VH), S descending FP64."""
    need(A)
    m, n = A.shape
    r = min(m, n)
    U = empty(m, r)
    VH = empty(r, n)
    S = torch.empty(r, dtype=torch.float64, device=A.device)
    wp, wn = workspace.get(lib.seemps_svd_workspace(m, n))
    info = (ctypes.c_int32 * 2)()
    check(lib.seemps_zgesvd(m, n, ptr(A), n, ptr(U), ptr(S), ptr(VH), wp, wn, info, stream()))
    return U, S, VH


def svd_split(AA: torch.Tensor, m: int, n: int, mode: int, method: int, tol: float,
              max_bond: int, want_s: bool = False):
    """SVD + truncation + split; returns (A (m x D), B (D x n), D, err^2, s or None)."""
    need(AA)
    if AA.numel() != m * n:
        raise ValueError("svd_split: matrix size does not match m x n")
    r = min(m, n)
    outA = empty(m * r)
    outB = empty(r * n)
    s = torch.empty(r, dtype=torch.float64, device=AA.device) if want_s else None
    D = ctypes.c_int64(0)
    err = ctypes.c_double(0.0)
    info = (ctypes.c_int32 * 2)()
    wp, wn = workspace.get(lib.seemps_svd_split_workspace(m, n))
    check(lib.seemps_svd_split(m, n, ptr(AA), mode, method, tol, max_bond, ptr(outA), ptr(outB),
                               ptr(s), ctypes.byref(D), ctypes.byref(err), wp, wn, info,
                               stream()))
    Dv = D.value
    return (outA[: m * Dv].view(m, Dv), outB[: Dv * n].view(Dv, n), Dv, err.value,
            None if s is None else s[:Dv])


def svd_split_sketch(AA: torch.Tensor, m: int, n: int, mode: int, method: int, tol: float,
                     max_bond: int, sketch: int, power_iters: int = 1, seed: int = 0,
                     want_s: bool = False):
    """Randomized fixed-rank variant of :func:`svd_split` with sketch size `sketch`.

    Returns (A, B, D, err^2, s or None, fallback) where fallback is True when the
    full SVD decided the split (the sketch missed more weight than the rule
    may discard, or `sketch` >= min(m, n))."""
    r = min(m, n)
    outA = empty(m * r)
    outB = empty(r * n)
    s = torch.empty(r, dtype=torch.float64, device=AA.device) if want_s else None
    D = ctypes.c_int64(0)
    err = ctypes.c_double(0.0)
    fb = ctypes.c_int32(0)
    info = (ctypes.c_int32 * 2)()
    wp, wn = workspace.get(lib.seemps_svd_split_sketch_workspace(m, n, sketch))
    check(lib.seemps_svd_split_sketch(m, n, ptr(AA), mode, method, tol, max_bond, sketch,
                                      power_iters, seed & 0xFFFFFFFFFFFFFFFF, ptr(outA),
                                      ptr(outB), ptr(s), ctypes.byref(D), ctypes.byref(err),
                                      ctypes.byref(fb), wp, wn, info, stream()))
    Dv = D.value
    return (outA[: m * Dv].view(m, Dv), outB[: Dv * n].view(Dv, n), Dv, err.value,
            None if s is None else s[:Dv], bool(fb.value))


def svd_split_basis(AA: torch.Tensor, m: int, n: int, side: int, basis: torch.Tensor, k: int,
                    extra: int, mode: int, method: int, tol: float, max_bond: int, seed: int = 0,
                    want_s: bool = False):
    """Split of the m x n matrix AA sketched by an isometry of the bond.

    side 0: `basis` is k x n with orthonormal rows (the right neighbour of a
    centre moving right); side 1: m x k with orthonormal columns (the left
    neighbour of a centre moving left).  `extra` expansion directions are
    added to the sketch.  Returns (A, B, D, err^2, s or None, fallback) like
    :func:`svd_split_sketch`."""
    r = min(m, n)
    basis = as_dev(basis)
    outA = empty(m * r)
    outB = empty(r * n)
    s = torch.empty(r, dtype=torch.float64, device=AA.device) if want_s else None
    D = ctypes.c_int64(0)
    err = ctypes.c_double(0.0)
    fb = ctypes.c_int32(0)
    info = (ctypes.c_int32 * 2)()
    wp, wn = workspace.get(lib.seemps_svd_split_basis_workspace(m, n, side, k, extra))
    check(lib.seemps_svd_split_basis(m, n, ptr(AA), side, ptr(basis), k, extra,
                                     seed & 0xFFFFFFFFFFFFFFFF, mode, method, tol, max_bond,
                                     ptr(outA), ptr(outB), ptr(s), ctypes.byref(D),
                                     ctypes.byref(err), ctypes.byref(fb), wp, wn, info, stream()))
    Dv = D.value
    return (outA[: m * Dv].view(m, Dv), outB[: Dv * n].view(Dv, n), Dv, err.value,
            None if s is None else s[:Dv], bool(fb.value))


def permute(x: torch.Tensor, perm) -> torch.Tensor:
    need(x)
    dims = (ctypes.c_int64 * x.dim())(*x.shape)
    p = (ctypes.c_int * x.dim())(*perm)
    out = empty(*[x.shape[k] for k in perm])
    check(lib.seemps_zpermute(x.dim(), dims, p, ptr(x), ptr(out), stream()))
    return out


def permuted(x: torch.Tensor, perm) -> torch.Tensor:
    """``x.permute(perm)`` materialised once per tensor.

    The copy is kept in a dictionary attached to ``x`` itself, so it is freed
    with ``x``, and it is recomputed when ``x`` has been modified in place
    (torch's version counter of ``x`` moved)."""
    key = tuple(perm)
    cache = getattr(x, "_seemps_permuted", None)
    if cache is None:
        cache = {}
        x._seemps_permuted = cache
    hit = cache.get(key)
    if hit is not None and hit[0] == x._version:
        return hit[1]
    y = permute(x, perm)
    cache[key] = (x._version, y)
    return y


def two_site_operator(O1: torch.Tensor, O2: torch.Tensor) -> torch.Tensor:
    """Hm[(c,k,l),(ko,lo,d)] from MPO tensors O1[c,ko,k,e], O2[e,lo,l,d]."""
    need(O1, O2)
    c, d1, d1b, e = O1.shape
    e2, d2, d2b, w2 = O2.shape
    if d1 != d1b or d2 != d2b or e != e2:
        raise ValueError("two_site_operator: incompatible MPO tensors")
    Hm = empty(c * d1 * d2, d1 * d2 * w2)
    check(lib.seemps_two_site_operator(c, d1, e, d2, w2, ptr(O1), ptr(O2), ptr(Hm), stream()))
    return Hm


def two_site_matvec(L: torch.Tensor, Hm: torch.Tensor, R: torch.Tensor, v: torch.Tensor,
                    d1: int, d2: int) -> torch.Tensor:
    need(L, Hm, R, v)
    a, c, b = L.shape
    g, w2, f = R.shape
    out = empty(a, d1, d2, g)
    wb = lib.seemps_two_site_matvec_workspace(a, c, b, d1, d2, w2, f)
    wp, wn = workspace.get(wb)
    check(lib.seemps_two_site_matvec(a, c, b, d1, d2, w2, f, g, ptr(L), ptr(Hm), ptr(R), ptr(v),
                                     ptr(out), wp, wn, stream()))
    return out


def eigensolve(L: torch.Tensor, Hm: torch.Tensor, R: torch.Tensor, v0: torch.Tensor,
               d1: int, d2: int, ncv: int = 20, tol: float = 0.0, maxiter: int = 1000,
               precision: str = "fp64"):
    """Lowest eigenpair of the two-site operator; returns (eigenvalue, vector, stats).

    ``precision`` selects the arithmetic of the operator's contractions: "fp64",
    "fp32", or "mixed" (FP32 restarts down to a 1e-5 relative residual, FP64 to
    ``tol``)."""
    need(L, Hm, R, v0)
    chiL, c, _ = L.shape
    chiR, w2, _ = R.shape
    evec = empty(chiL, d1, d2, chiR)
    stats = (ctypes.c_double * 5)()
    wb = lib.seemps_eigensolve_workspace(chiL, c, d1, d2, w2, chiR, ncv)
    wp, wn = workspace.get(wb)
    check(lib.seemps_two_site_eigensolve_prec(PRECISIONS[precision], chiL, c, d1, d2, w2, chiR,
                                              ptr(L), ptr(Hm), ptr(R), ptr(v0), ptr(evec), ncv,
                                              tol, maxiter, stats, wp, wn, stream()))
    return stats[0], evec, {"residual": stats[1], "matvecs": int(stats[2]),
                            "restarts": int(stats[3]), "matvecs_fp32": int(stats[4])}
