"""Host mirror of @RJ/gaussMLE.m for fit types 'Basic', 'Sigma', 'SigmaYX' and 'Z' on CCD and sCMOS cameras.

``gauss_mle(data, 'Basic', 'CCD', sigma)`` returns ``(Results, Statistics)``
with the fields gaussMLE.m:183-210 builds (Y, X, Photons, BG, *_SE, and Sigma,
Sigma_SE for 'Sigma', SigmaY, SigmaX and their SEs for 'SigmaYX', Z and Z_SE
for 'Z';
LogLikelihood; FitTime, FitPerSecond, NKernelsCalls, NFitsPerCall).  The fit
runs in one call of bamf_gauss_mle_xynb ('Basic', kernel _XYNB_),
bamf_gauss_mle_xynbs ('Sigma', _XYNBS_), bamf_gauss_mle_xynbsxsy
('SigmaYX', _XYNBSXSY_) or bamf_gauss_mle_xynbz ('Z', _XYNBZ_) on the
device; there is no host fallback.
"""
from __future__ import annotations

import time

import numpy as np

from . import _hip
from ._lib import BamfError, check, lib


_KERNELS = {"Basic": ("bamf_gauss_mle_xynb", 4), "Sigma": ("bamf_gauss_mle_xynbs", 5),
            "SigmaYX": ("bamf_gauss_mle_xynbsxsy", 6), "Z": ("bamf_gauss_mle_xynbz", 5)}
_ZFIELDS = ("Ax", "Ay", "Bx", "By", "Gamma", "D")


def gauss_mle_device(d_data, sigma, sz: int, iterations: int, nfits: int, d_params, d_crlb, d_ll,
                     stream=None, fit_type: str = "Basic", zstruct: dict | None = None, scmos=None) -> None:
    """Direct C-ABI call on device buffers (the kernel_MLEFit_* argument layouts).
    For 'Z', sigma = (SigmaX, SigmaY) and zstruct holds Ax, Ay, Bx, By, Gamma, D.
    scmos = (d_coords, d_gain_ratio, mapsz, d_z0 or None) selects the sCMOS kernels."""
    if fit_type not in _KERNELS:
        raise BamfError(f"fit type {fit_type!r}: 'Basic', 'Sigma', 'SigmaYX' and 'Z' are implemented")
    name = _KERNELS[fit_type][0]
    if scmos is not None:
        name = name.replace("bamf_gauss_mle_", "bamf_gauss_mle_scmos_")
    sv = stream.vp if stream else None
    fn = getattr(lib(), name)
    if fit_type == "Z":
        if zstruct is None:
            raise BamfError("fit type 'Z' needs the ZStruct calibration")
        sx, sy = (float(v) for v in np.ravel(sigma)[:2])
        cal = [float(zstruct[k]) for k in _ZFIELDS]
        if scmos is not None:
            dc, dg, mapsz, dz = scmos
            check(fn(d_data.vp, dc.vp, dg.vp, dz.vp, sx, *cal, sy, int(sz), int(mapsz), int(iterations), d_params.vp,
                     d_crlb.vp, d_ll.vp, int(nfits), sv), name)
        else:
            check(fn(d_data.vp, sx, *cal, sy, int(sz), int(iterations), d_params.vp, d_crlb.vp, d_ll.vp, int(nfits),
                     sv), name)
        return
    if scmos is not None:
        dc, dg, mapsz, _ = scmos
        check(fn(d_data.vp, dc.vp, dg.vp, float(np.mean(sigma)), int(sz), int(mapsz), int(iterations), d_params.vp,
                 d_crlb.vp, d_ll.vp, int(nfits), sv), name)
        return
    check(fn(d_data.vp, float(np.mean(sigma)), int(sz), int(iterations), d_params.vp, d_crlb.vp, d_ll.vp, int(nfits),
             sv), name)


def fit_boxes(boxes: np.ndarray, sigma, iterations: int = 20, fit_type: str = "Basic", zstruct: dict | None = None,
              coords: np.ndarray | None = None, gain_ratio: np.ndarray | None = None, z0: np.ndarray | None = None):
    """boxes: (N, sz, sz) float32 with box[i, jj, ii] = data[sz*jj + ii].
    Returns (params (NP, N), crlb (NP, N), ll (N,)) in the kernel's parameter-major
    layout, NP = 4 ('Basic'), 5 ('Sigma', 'Z') or 6 ('SigmaYX').  With coords
    ((N, 2) box corners, read as the kernel reads d_Coords) and gain_ratio
    (the square variance/gain^2 map, row r = coords[:, 1] + jj, column
    c = coords[:, 0] + ii), the sCMOS kernels."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n, sz = boxes.shape[0], boxes.shape[1]
    if boxes.shape[1] != boxes.shape[2]:
        raise BamfError("boxes must be square")
    if fit_type not in _KERNELS:
        raise BamfError(f"fit type {fit_type!r}: 'Basic', 'Sigma', 'SigmaYX' and 'Z' are implemented")
    npar = _KERNELS[fit_type][1]
    d_data = _hip.DeviceBuffer.from_array(boxes)
    d_p = _hip.DeviceBuffer(4 * npar * n)
    d_c = _hip.DeviceBuffer(4 * npar * n)
    d_l = _hip.DeviceBuffer(n * 4)
    scmos = None
    if coords is not None:
        g = np.ascontiguousarray(gain_ratio, np.float32)
        dz = _hip.DeviceBuffer.from_array(np.ascontiguousarray(z0 if z0 is not None else np.zeros(n), np.float32))
        scmos = (_hip.DeviceBuffer.from_array(np.ascontiguousarray(coords, np.float32)),
                 _hip.DeviceBuffer.from_array(g), g.shape[0], dz)
    gauss_mle_device(d_data, sigma, sz, iterations, n, d_p, d_c, d_l, fit_type=fit_type, zstruct=zstruct,
                     scmos=scmos)
    _hip.synchronize()
    return (d_p.download(np.float32).reshape(npar, n), d_c.download(np.float32).reshape(npar, n),
            d_l.download(np.float32))


def gauss_mle(data: np.ndarray, fit_type: str = "Basic", camera_type: str = "CCD", sigma=1.3, zstruct=None,
              iterations: int = 20, variance_image=None, box_corners=None):
    """data: MATLAB-layout stack (SZ, SZ, N) as a Fortran-ordered array, or (N, SZ, SZ) C-ordered boxes.
    CameraType 'SCMOS' takes variance_image (the square variance/gain^2 map) and
    box_corners ((N, 2), as the kernels read d_Coords) (gaussMLE.m:162-173)."""
    if fit_type not in _KERNELS or camera_type not in ("CCD", "SCMOS"):
        raise BamfError(f"fit type {fit_type!r} / camera {camera_type!r}: 'Basic', 'Sigma', 'SigmaYX' and 'Z' "
                        "on 'CCD' are implemented")
    data = np.asarray(data, dtype=np.float32)
    if data.ndim == 2:
        data = data[:, :, None] if data.flags.f_contiguous else data[None]
    if data.flags.f_contiguous and not data.flags.c_contiguous:
        boxes = np.ascontiguousarray(np.transpose(data, (2, 1, 0)))  # (N, col, row): [i, jj, ii] = data[ii, jj, i]
    else:
        boxes = np.ascontiguousarray(data)
    t0 = time.perf_counter()
    if camera_type == "SCMOS" and (variance_image is None or box_corners is None):
        raise BamfError("CameraType 'SCMOS' needs VarianceImage and BoxCorners")
    p, c, ll = fit_boxes(boxes, sigma, iterations, fit_type, zstruct,
                         box_corners if camera_type == "SCMOS" else None,
                         variance_image if camera_type == "SCMOS" else None)
    dt = time.perf_counter() - t0
    n = boxes.shape[0]
    results = {
        "Y": p[0], "X": p[1], "Photons": p[2], "BG": p[3],
        "Y_SE": np.sqrt(c[0]), "X_SE": np.sqrt(c[1]), "Photons_SE": np.sqrt(c[2]), "BG_SE": np.sqrt(c[3]),
        "LogLikelihood": ll,
    }
    if fit_type == "Sigma":  # gaussMLE.m:193-195
        results["Sigma"] = p[4]
        results["Sigma_SE"] = np.sqrt(c[4])
    if fit_type == "Z":  # gaussMLE.m:196-198
        results["Z"] = p[4]
        results["Z_SE"] = np.sqrt(c[4])
    if fit_type == "SigmaYX":  # gaussMLE.m:199-202
        results["SigmaY"], results["SigmaX"] = p[4], p[5]
        results["SigmaY_SE"], results["SigmaX_SE"] = np.sqrt(c[4]), np.sqrt(c[5])
    stats = {"FitTime": dt, "FitPerSecond": n / dt if dt > 0 else float("inf"), "NKernelsCalls": 1,
             "NFitsPerCall": n}
    return results, stats
