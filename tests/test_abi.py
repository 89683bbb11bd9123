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


def test_scmos_gateway_refuses_box_off_the_map():
    """bamf_gauss_mle_scmos_fit_host checks every box corner against the gain map
    on the host before any device work (the kernels read the map at the caller's
    corners)."""
    import numpy as np

    L = _lib.lib()
    sz, mapsz, n = 7, 16, 3
    data = np.zeros((n, sz, sz), np.float32)
    gain = np.ones((mapsz, mapsz), np.float32)
    out = [np.zeros(4 * n, np.float32), np.zeros(4 * n, np.float32), np.zeros(n, np.float32)]
    for bad in ([10.0, 0.0], [0.0, 9.5], [-1.0, 2.0], [3.0, -0.5]):
        coords = np.array([[0, 0], [9, 9], bad], np.float32)
        rc = L.bamf_gauss_mle_scmos_fit_host(data.ctypes.data, sz, n, coords.ctypes.data, gain.ctypes.data, mapsz,
                                             1.3, 20, 0, None, None, out[0].ctypes.data, out[1].ctypes.data,
                                             out[2].ctypes.data, None)
        assert rc == -1, bad
        assert b"not on the" in L.bamf_last_error()


def test_find_roi_noalloc_refuses_before_device_work():
    L = _lib.lib()
    assert L.bamf_find_roi_workspace_bytes(256, 256, 115) <= 64 << 20 < L.bamf_find_roi_workspace_bytes(256, 256, 116)
    rc = L.bamf_find_roi_noalloc(None, 8, 8, 1, 1.3, 2.6, 7, 1.0, None, None, None, None, None, 0, None, None)
    assert rc == -1
