"""ctypes binding of libbamf_gpu.so (the C ABI declared in include/bamf_gpu.h).

The library must exist: there is no CPU fallback.  Build it with
``python -m bamf_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libbamf_gpu.so"
HEADER = Path(__file__).resolve().parent.parent / "include" / "bamf_gpu.h"

_lib = None


class BamfError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise BamfError(f"{LIB_PATH} is missing: build it with `python -m bamf_amd.build`")
    L = ctypes.CDLL(str(LIB_PATH))
    _declare(L)
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().bamf_last_error().decode(errors="replace")
        raise BamfError(f"{what}: {msg} (code {rc})")


def _declare(L: ctypes.CDLL) -> None:
    fp, ip = POINTER(c_float), POINTER(c_int)
    L.bamf_last_error.restype = c_char_p
    L.bamf_version.restype = c_int
    L.bamf_device_info.argtypes = [c_int, ip, ip, c_char_p, c_int]
    L.bamf_gauss_mle_xynb.argtypes = [c_void_p, c_float, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p]
    L.bamf_gauss_mle_xynbs.argtypes = [c_void_p, c_float, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p]
    L.bamf_gauss_mle_xynbsxsy.argtypes = [c_void_p, c_float, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                          c_void_p]
    L.bamf_gauss_mle_fit_host.argtypes = [c_void_p, c_int, c_int64, c_float, c_int, c_int, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p]
    L.bamf_gauss_mle_xynbz.argtypes = [c_void_p] + [c_float] * 8 + [c_int, c_int, c_void_p, c_void_p, c_void_p,
                                                                     c_int, c_void_p]
    sc = [c_void_p, c_void_p, c_void_p, c_float, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p]
    for name in ("bamf_gauss_mle_scmos_xynb", "bamf_gauss_mle_scmos_xynbs", "bamf_gauss_mle_scmos_xynbsxsy"):
        getattr(L, name).argtypes = sc
    L.bamf_gauss_mle_scmos_xynbz.argtypes = [c_void_p] * 4 + [c_float] * 8 + [c_int, c_int, c_int, c_void_p,
                                                                               c_void_p, c_void_p, c_int, c_void_p]
    L.bamf_gauss_mle_scmos_fit_host.argtypes = [c_void_p, c_int, c_int64, c_void_p, c_void_p, c_int, c_float, c_int,
                                                c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.bamf_gauss_coeffs.argtypes = [c_float, fp]
    L.bamf_gauss_coeffs.restype = None
    coef = [c_float] * 5
    L.bamf_find_roi_gauss_x.argtypes = [c_void_p, c_int, c_int, c_int] + coef + [c_void_p]
    L.bamf_find_roi_gauss_y.argtypes = [c_void_p, c_int, c_int, c_int] + coef + [c_void_p]
    L.bamf_find_roi_subtract.argtypes = [c_void_p, c_void_p, c_int64, c_void_p]
    L.bamf_find_roi_max_x.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_float, c_void_p]
    L.bamf_find_roi_max_y.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_float, c_void_p]
    L.bamf_find_roi_workspace_bytes.argtypes = [c_int, c_int, c_int]
    L.bamf_find_roi_workspace_bytes.restype = c_size_t
    L.bamf_find_roi.argtypes = [c_void_p, c_int, c_int, c_int, c_float, c_float, c_int, c_float,
                                c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_int64, c_void_p, c_void_p]
    L.bamf_find_roi_noalloc.argtypes = [c_void_p, c_int, c_int, c_int, c_float, c_float, c_int, c_float,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                        c_void_p]
    L.bamf_find_roi_crop.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int64, c_int,
                                     c_void_p, c_void_p]
    L.bamf_fit_keep.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_float, c_float, c_float, c_void_p,
                                c_void_p]
    L.bamf_kde_pdf.argtypes = [c_void_p, c_int64, c_double, c_double, c_double, c_int, c_void_p, c_void_p]
    L.bamf_psf_shift_average.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p]
    L.bamf_psf_workspace_bytes.argtypes = [c_int64]
    L.bamf_psf_workspace_bytes.restype = c_int64
    L.bamf_find_psf_host.argtypes = [c_void_p, c_int, c_int64, c_float, c_float, c_float, c_void_p, c_void_p,
                                     c_void_p, c_void_p]
    L.bamf_find_priors_host.argtypes = [c_void_p, c_int, c_int, c_int] + [c_float] * 7 + [c_int] + \
        [c_void_p] * 3 + [c_int64, c_void_p, c_void_p]
    L.bamf_subregion_count.argtypes = [c_int, c_int, c_int, c_int, c_int]
    L.bamf_subregion_count.restype = c_int64
    L.bamf_subregion_pixels.argtypes = [c_int, c_int, c_int, c_int, c_int]
    L.bamf_subregion_pixels.restype = c_int64
    L.bamf_make_subregions.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int] + [c_void_p] * 7 + \
        [c_int, c_void_p, c_void_p, c_void_p]
    L.bamf_synth_cells.argtypes = [c_int, c_int]
    L.bamf_synth_cells.restype = c_int64
    L.bamf_synth_frames.argtypes = [c_int, c_int, c_int] + [c_float] * 5 + [c_uint64, c_int, c_int] + \
        [c_void_p] * 4
    L.bamf_synth_roi_batch.argtypes = [c_int, c_int] + [c_float] * 7 + [c_uint64, c_int] + [c_void_p] * 6
    if hasattr(L, "bamf_rjmcmc_plan_create"):
        from . import rjmcmc_abi

        rjmcmc_abi.declare(L)
    if hasattr(L, "bamf_net_forward"):
        from . import net_abi

        net_abi.declare(L)


def header_functions() -> list[str]:
    """Names of every function declared in include/bamf_gpu.h."""
    import re

    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bamf_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))
