"""Host mirror of @RJ/findPriors.m:35-80 over bamf_find_priors_host: findROI,
gaussMLE, threhsSMA and the kernel density on the device, the scalar glue
(quantile, the second-mode cut, the Gamma fit) in the C++ gateway."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib


def find_priors(data: np.ndarray, psf_sigma: float, sma: dict, p_dp: float, bg_prior_range: float,
                bg_prior: str = "Yes", capacity: int = 1 << 20):
    """data: (Ysize, Xsize[, Z]) frames with MATLAB's meaning; sma holds MeanPhotons,
    MinPValue, MinPhotons and optionally MaxXY_SE (default 0.2, threhsSMA.m:2-4).
    Returns the RJStruct fields P_EmPhotons, P_BgEmPhotons, P_Offset and an info dict."""
    a = np.asarray(data, np.float32)
    if a.ndim == 2:
        a = a[:, :, None]
    ys, xs, nf = a.shape
    flat = np.asfortranarray(a).ravel(order="F")
    sig, bgp, off = (np.zeros(capacity, np.float64) for _ in range(3))
    n = ctypes.c_int64(0)
    info = np.zeros(8, np.float64)
    ptr = lambda v: v.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    check(lib().bamf_find_priors_host(ptr(flat), ys, xs, nf, float(psf_sigma), float(sma["MeanPhotons"]),
                                      float(sma.get("MaxXY_SE", 0.2)), float(sma["MinPValue"]),
                                      float(sma["MinPhotons"]), float(p_dp), float(bg_prior_range),
                                      1 if bg_prior == "Yes" else 0, ptr(sig), ptr(bgp), ptr(off), capacity,
                                      ctypes.byref(n), ptr(info)), "bamf_find_priors_host")
    m = min(n.value, capacity)
    return {"P_EmPhotons": sig[:m].copy(), "P_BgEmPhotons": bgp[:m].copy(), "P_Offset": off[:m].copy(),
            "info": {"rois": int(info[0]), "kept": int(info[1]), "photons": int(info[2]), "gamma_a": info[3],
                     "gamma_b": info[4], "refit": int(info[5]), "mean_photons": info[6], "minval": info[7]}}
