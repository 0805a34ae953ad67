"""seemps_zdotc_noalloc (partials in the library's static device scratch)
against seemps_zdotc (partials in a caller workspace), through the C ABI:
identical results on two calls in a row, the second on a smaller input."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_zdotc_noalloc_matches_workspace_entry_point():
    from seemps2_amd import _lib

    lib = _lib.lib
    lib.seemps_zdotc_noalloc.restype = ctypes.c_int
    lib.seemps_zdotc_noalloc.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]
    dev = _lib.device()
    stream = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device=dev).manual_seed(77)
    for n in (3_000_000, 40_000, 1):   # partial counts 512 (the cap), 10, 1
        x = torch.randn(n, dtype=torch.complex128, device=dev, generator=g)
        y = torch.randn(n, dtype=torch.complex128, device=dev, generator=g)
        ref = torch.empty(1, dtype=torch.complex128, device=dev)
        out = torch.full((1,), complex(np.nan, np.nan), dtype=torch.complex128, device=dev)
        nbytes = lib.seemps_reduce_workspace(n)
        work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        assert lib.seemps_zdotc(n, x.data_ptr(), y.data_ptr(), ref.data_ptr(), work.data_ptr(),
                                nbytes, stream) == 0
        assert lib.seemps_zdotc_noalloc(n, x.data_ptr(), y.data_ptr(), out.data_ptr(),
                                        stream) == 0, _lib.last_error()
        torch.cuda.synchronize()
        # same kernels, same partition and summation order: bit-identical
        assert complex(out.item()) == complex(ref.item()), (n, out.item(), ref.item())
        exact = np.vdot(x.cpu().numpy(), y.cpu().numpy())
        assert abs(complex(out.item()) - exact) <= 1e-12 * np.sqrt(n) * max(abs(exact), 1.0)
    assert lib.seemps_zdotc_noalloc(-1, None, None, None, stream) == _lib.ERR_INVALID
