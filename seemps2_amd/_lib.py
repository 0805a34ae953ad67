"""ctypes binding of libseemps_hip.so and device-buffer plumbing.

The shared library is built in-tree by ``seemps2_amd/build.py``.  Importing
this module without it raises ``ImportError``: there is no CPU fallback.
Device buffers are torch tensors on the current HIP device; kernels are
enqueued on torch's current stream.
"""

from __future__ import annotations

import atexit
import ctypes
from ctypes import c_double, c_int, c_int32, c_int64, c_void_p, POINTER
from pathlib import Path

import numpy as np
import torch

LIB_PATH = Path(__file__).resolve().parent / "libseemps_hip.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python seemps2_amd/build.py` "
        "(the HIP kernels are the only implementation of this package)"
    )

lib = ctypes.CDLL(str(LIB_PATH))

OK, ERR_INVALID, ERR_NOCONVERGE, ERR_RUNTIME, ERR_WORKSPACE = 0, 1, 2, 3, 4
OP_N, OP_T, OP_H, OP_C = 0, 1, 2, 3

_P = c_void_p
_I = c_int64
_D = c_double

_SIGNATURES = {
    "seemps_last_error": (ctypes.c_char_p, []),
    "seemps_abi_version": (c_int, []),
    "seemps_release": (c_int, []),
    "seemps_zgemm": (
        c_int,
        [c_int, c_int, _I, _I, _I, _D, _D, _P, _I, _I, _P, _I, _I, _D, _D, _P, _I, _I, _I, _P],
    ),
    "seemps_zgemm_workspace": (_I, [_I, _I, _I]),
    "seemps_zgemm_ws": (
        c_int,
        [c_int, c_int, _I, _I, _I, _D, _D, _P, _I, _P, _I, _D, _D, _P, _I, _P, _I, _P],
    ),
    "seemps_zgemm_prec": (
        c_int,
        [c_int, c_int, c_int, _I, _I, _I, _D, _D, _P, _I, _I, _P, _I, _I, _D, _D, _P, _I, _I, _I,
         _P],
    ),
    "seemps_zsite_product": (c_int, [_I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "seemps_grouped_tiles": (_I, [c_int, _I, _I]),
    "seemps_zgemm_grouped": (c_int, [c_int, c_int, c_int, _I, _P, _P, _I, _P]),
    "seemps_zlocal": (c_int, [_I, _I, _I, _I, _I, _P, _P, _P, c_int, _P]),
    "seemps_zaxpby": (c_int, [_I, _D, _D, _P, c_int, _D, _D, _P, _P]),
    "seemps_zscale2d": (c_int, [_I, _I, _P, _P, _P, _I, _P, _I, _P]),
    "seemps_reduce_workspace": (_I, [_I]),
    "seemps_zdotc": (c_int, [_I, _P, _P, _P, _P, _I, _P]),
    "seemps_truncate": (c_int, [_I, _P, c_int, _D, _I, _P, _P, _P]),
    "seemps_svd_workspace": (_I, [_I, _I]),
    "seemps_zgesvd": (c_int, [_I, _I, _P, _I, _P, _P, _P, _P, _I, POINTER(c_int32), _P]),
    "seemps_svd_split_workspace": (_I, [_I, _I]),
    "seemps_svd_split": (
        c_int,
        [_I, _I, _P, c_int, c_int, _D, _I, _P, _P, _P, POINTER(c_int64), POINTER(c_double),
         _P, _I, POINTER(c_int32), _P],
    ),
    "seemps_two_site_operator": (c_int, [_I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "seemps_zpermute": (c_int, [c_int, POINTER(c_int64), POINTER(c_int), _P, _P, _P]),
    "seemps_two_site_matvec_workspace": (_I, [_I, _I, _I, _I, _I, _I, _I]),
    "seemps_two_site_matvec": (
        c_int, [_I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _P]
    ),
    "seemps_eigensolve_workspace": (_I, [_I, _I, _I, _I, _I, _I, _I]),
    "seemps_two_site_eigensolve": (
        c_int,
        [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _D, _I, POINTER(c_double), _P, _I, _P],
    ),
    "seemps_two_site_eigensolve_prec": (
        c_int,
        [c_int, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _D, _I, POINTER(c_double), _P,
         _I, _P],
    ),
    "seemps_svd_split_sketch_workspace": (_I, [_I, _I, _I]),
    "seemps_svd_split_sketch": (
        c_int,
        [_I, _I, _P, c_int, c_int, _D, _I, _I, _I, ctypes.c_uint64, _P, _P, _P,
         POINTER(c_int64), POINTER(c_double), POINTER(c_int32), _P, _I, POINTER(c_int32), _P],
    ),
    "seemps_svd_split_basis_workspace": (_I, [_I, _I, c_int, _I, _I]),
    "seemps_svd_split_basis": (
        c_int,
        [_I, _I, _P, c_int, _P, _I, _I, ctypes.c_uint64, c_int, c_int, _D, _I, _P, _P, _P,
         POINTER(c_int64), POINTER(c_double), POINTER(c_int32), _P, _I, POINTER(c_int32), _P],
    ),
    "seemps_spectral_sketch_workspace": (_I, [_I, _I]),
    "seemps_spectral_sketch": (c_int, [_I, _I, _P, ctypes.c_uint64, _P, _P, _I, _P]),
    "seemps_mlp_param_count": (_I, [c_int, c_int]),
    "seemps_mlp_scratch_count": (_I, [c_int, c_int]),
    "seemps_mlp_forward": (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, _P]),
    "seemps_mlp_backward": (
        c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    ),
    "seemps_adam_step": (
        c_int, [_I, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                ctypes.c_float, _I, _P]
    ),
    "seemps_timing_enable": (c_int, [ctypes.c_char_p]),
    "seemps_timing_calibrate": (c_int, [c_int, _P]),
    "seemps_timing_count": (c_int, []),
    "seemps_timing_get": (
        c_int,
        [c_int, ctypes.c_char_p, _I, POINTER(c_int64), POINTER(c_double), POINTER(c_double)],
    ),
}

for _name, (_res, _args) in _SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

ABI_VERSION = lib.seemps_abi_version()

# Free the pinned staging buffers and pooled events while the HIP runtime is
# still up (interpreter exit runs atexit handlers before the runtime's own
# teardown).
atexit.register(lib.seemps_release)


def last_error() -> str:
    return lib.seemps_last_error().decode(errors="replace")


def check(status: int) -> None:
    """Map a library status code to the reference's exception types."""
    if status == OK:
        return
    msg = last_error()
    if status == ERR_NOCONVERGE:
        raise np.linalg.LinAlgError(msg)
    if status in (ERR_INVALID, ERR_WORKSPACE):
        raise ValueError(msg)
    raise RuntimeError(msg)


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("seemps2_amd needs a HIP device (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


class _Workspace:
    """Per-device scratch arena, grown on demand (stream-ordered reuse)."""

    def __init__(self):
        self.buf: dict[int, torch.Tensor] = {}

    def get(self, nbytes: int) -> tuple[int, int]:
        dev = torch.cuda.current_device()
        b = self.buf.get(dev)
        if b is None or b.numel() < nbytes:
            size = max(int(nbytes), 1 << 20)
            if b is not None:
                size = max(size, int(b.numel() * 1.5))
            b = torch.empty(size, dtype=torch.uint8, device=device())
            self.buf[dev] = b
        return b.data_ptr(), b.numel()


workspace = _Workspace()


def timing_enable(kernels: str | None) -> None:
    """Time the named kernels' launches with HIP events ("*" = all, None = off)."""
    check(lib.seemps_timing_enable(kernels.encode() if kernels else None))


def timing_bracket_us(launches: int = 400) -> float:
    """Average HIP-event time (us) of an empty kernel bracketed the way the
    instrumented launches are: the bracketing's own cost in each record.
    Leaves timing off."""
    timing_enable("timing_calibration")
    try:
        check(lib.seemps_timing_calibrate(launches, stream()))
        rec = timing_records().get("timing_calibration")
    finally:
        timing_enable(None)
    if not rec or rec["launches"] == 0:
        return 0.0
    return rec["ms"] * 1e3 / rec["launches"]


def timing_records() -> dict[str, dict]:
    """{kernel: {launches, ms, flops}} accumulated since timing_enable."""
    out = {}
    buf = ctypes.create_string_buffer(128)
    for i in range(lib.seemps_timing_count()):
        n, ms, fl = c_int64(), c_double(), c_double()
        check(lib.seemps_timing_get(i, buf, 128, ctypes.byref(n), ctypes.byref(ms),
                                    ctypes.byref(fl)))
        out[buf.value.decode()] = dict(launches=n.value, ms=ms.value, flops=fl.value)
    return out
