"""Host mirror of @RJ/findPSF.m:30-119 over bamf_find_psf_host: the 'Sigma'
fits and the Fourier shift-and-average on the device, the offset, peak
scaling and the 4 x 4 phase split in the C++ gateway.  The table it returns
is what Plan(..., sampled_psf=...) and cRJMCMC's SampledPSF take."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib


def find_psf(roi_stack: np.ndarray, gain: float = 1.0, offset: float = 0.0, psf_sigma: float = 1.3,
             box: int = 8, return_fits: bool = False):
    """roi_stack: (BoxSize, BoxSize, N) with MATLAB's meaning (rows = y), or
    (N, BoxSize, BoxSize) C-ordered boxes.  Returns SampledPSF, a
    (64, 64, 4, 4) Fortran-ordered float32 array (y, x, phase_y, phase_x);
    with return_fits also the positions the shift used."""
    a = np.asarray(roi_stack, np.float32)
    if a.ndim == 2:
        a = a[:, :, None]
    if a.ndim != 3:
        raise ValueError("ROIStack is BoxSize x BoxSize x N")
    if a.shape[0] == a.shape[1] == box:
        flat = np.asfortranarray(a).ravel(order="F")  # (box, box, N) MATLAB stack
        n = a.shape[2]
    else:
        n = a.shape[0]
        flat = np.ascontiguousarray(a.transpose(0, 2, 1)).ravel()  # each box column-major
    out = np.zeros(64 * 64 * 16, np.float32)
    size = np.zeros(4, np.int32)
    fx, fy = np.zeros(n, np.float32), np.zeros(n, np.float32)
    ptr = lambda v: v.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    check(lib().bamf_find_psf_host(ptr(flat), int(box), n, float(gain), float(offset), float(psf_sigma), ptr(out),
                                   ptr(size), ptr(fx), ptr(fy)), "bamf_find_psf_host")
    psf = np.asfortranarray(out.reshape(tuple(size), order="F"))
    return (psf, fx, fy) if return_fits else psf
