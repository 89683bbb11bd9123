"""The C-ABI library loads on a CPU-only host and exports every symbol that
include/bamf_gpu.h declares (no compute calls)."""
import ctypes

import pytest

from bamf_amd import _lib


def test_header_parses():
    names = _lib.header_functions()
    assert "bamf_gauss_mle_xynb" in names and "bamf_find_roi" in names


def test_library_exports_header_symbols():
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in _lib.header_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_arg_errors():
    L = _lib.lib()
    assert L.bamf_version() == 1
    # invalid arguments are rejected on the host, before any device work
    rc = L.bamf_gauss_mle_xynb(None, 1.3, 0, 20, None, None, None, 10, None)
    assert rc == -1 and b"sz" in L.bamf_last_error()
    rc = L.bamf_find_roi(None, 8, 8, 1, 1.3, 2.6, 7, 1.0, None, 0, None, None, None, None, None, 0, None, None)
    assert rc == -1


def test_gauss_coeffs_host_matches_oracle(oracle):
    L = _lib.lib()
    for s in (0.9, 1.3, 2.6, 3.1):
        out = (ctypes.c_float * 5)()
        L.bamf_gauss_coeffs(s, out)
        assert list(out) == oracle.gauss_coeffs(s)
