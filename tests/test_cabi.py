"""The C-ABI library loads and exports every symbol declared in include/seemps_hip.h.

CPU-only: no compute call is made (there is no device here)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "seemps_hip.h"
LIB = ROOT / "seemps2_amd" / "libseemps_hip.so"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(seemps_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ["seemps_zgemm", "seemps_zgesvd", "seemps_svd_split", "seemps_truncate",
                 "seemps_two_site_eigensolve", "seemps_two_site_matvec", "seemps_zlocal"]:
        assert name in syms


@pytest.mark.skipif(not LIB.exists(), reason="libseemps_hip.so not built")
def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(str(LIB))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not LIB.exists(), reason="libseemps_hip.so not built")
def test_host_only_entry_points():
    lib = ctypes.CDLL(str(LIB))
    lib.seemps_abi_version.restype = ctypes.c_int
    assert lib.seemps_abi_version() == 1
    lib.seemps_svd_workspace.restype = ctypes.c_int64
    lib.seemps_svd_workspace.argtypes = [ctypes.c_int64, ctypes.c_int64]
    assert lib.seemps_svd_workspace(512, 512) > 2 * 512 * 512 * 16
    lib.seemps_eigensolve_workspace.restype = ctypes.c_int64
    lib.seemps_eigensolve_workspace.argtypes = [ctypes.c_int64] * 7
    assert lib.seemps_eigensolve_workspace(256, 5, 2, 2, 5, 256, 20) > 21 * 262144 * 16


@pytest.mark.skipif(not LIB.exists(), reason="libseemps_hip.so not built")
def test_timing_registry_host_side():
    """Enabling, reading and disabling the kernel timers needs no device when nothing ran."""
    lib = ctypes.CDLL(str(LIB))
    lib.seemps_timing_enable.argtypes = [ctypes.c_char_p]
    lib.seemps_timing_get.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
    lib.seemps_last_error.restype = ctypes.c_char_p
    assert lib.seemps_timing_enable(b"svd_jacobi,zgemm_kernel") == 0
    assert lib.seemps_timing_count() == 0
    n, ms, fl = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    buf = ctypes.create_string_buffer(32)
    assert lib.seemps_timing_get(0, buf, 32, ctypes.byref(n), ctypes.byref(ms),
                                 ctypes.byref(fl)) == 1  # SEEMPS_ERR_INVALID
    assert b"out of range" in lib.seemps_last_error()
    assert lib.seemps_timing_enable(None) == 0
    assert lib.seemps_timing_count() == 0


def test_product_package_does_not_import_oracle():
    for py in (ROOT / "seemps2_amd").rglob("*.py"):
        src = py.read_text()
        assert "import oracle" not in src and "from oracle" not in src, py
