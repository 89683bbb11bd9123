"""Seeded synthetic inputs drawn from BAMF's forward model.

Model convention (LibRJMCMC.cpp:15-59): pixel (y, x) covers [x, x+1) x
[y, y+1); the expected count is bg + sum_n I_n * dPhi_x * dPhi_y with the
pixel-integrated Gaussian PSF; ROIs are column-major with y fastest
(data[y + x*Ysize]).  Counts are Poisson.  These generators make the
workloads the benchmarks and tests run on; they are not part of the hot path.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf, gammaln

SQRT2 = math.sqrt(2.0)


def gamma_pdf_table(k: float, theta: float, dx: float, n: int) -> np.ndarray:
    """Gamma(k, theta) pdf on 0, dx, ..., (n-1)dx (float32), the Stat table format."""
    x = np.arange(n, dtype=np.float64) * dx
    with np.errstate(divide="ignore"):
        lp = (k - 1) * np.log(x) - x / theta - gammaln(k) - k * math.log(theta)
    p = np.exp(lp)
    p[0] = 0.0 if k > 1 else p[0]
    return p.astype(np.float32)


def uniform_pdf_table(lo: float, hi: float, dx: float, n: int) -> np.ndarray:
    x = np.arange(n, dtype=np.float64) * dx
    p = np.where((x >= lo) & (x <= hi), 1.0 / (hi - lo), 0.0)
    return p.astype(np.float32)


def _axis_integrals(pos: np.ndarray, npx: int, sigma: float) -> np.ndarray:
    """(n_emitters, npx) pixel integrals of a unit Gaussian along one axis."""
    c = 1.0 / (sigma * SQRT2)
    edges = np.arange(npx + 1, dtype=np.float64)
    e = erf((edges[None, :] - pos[:, None]) * c)
    return 0.5 * (e[:, 1:] - e[:, :-1])


def roi_expected(I, X, Y, bg, xsize: int, ysize: int, sigma: float) -> np.ndarray:
    """Expected counts, column-major (data[y + x*ysize]) float64."""
    I = np.asarray(I, np.float64)
    gx = _axis_integrals(np.asarray(X, np.float64), xsize, sigma) * I[:, None]
    gy = _axis_integrals(np.asarray(Y, np.float64), ysize, sigma)
    m = bg + gx.T @ gy  # (xsize, ysize)
    return m.ravel()  # x-major, y fastest


def sampled_gauss_psf(size: int, oversampling: int, sigma_y: float, sigma_x: float | None = None,
                      size_x: int | None = None) -> np.ndarray:
    """A SampledPSF table (y, x, phase_y, phase_x) of the pixel-integrated
    Gaussian, in the layout NumModel reads (LibRJMCMC.cpp:86-169): entry
    (d0, d1, iy, ix) is the unit-intensity value of the pixel whose lower
    edges lie at d0 - ST0 - iy / OS and d1 - ST0 - ix / OS from the emitter,
    ST0 = size // 2 on both axes (the reference takes it from PSFsize(1))."""
    sigma_x = sigma_y if sigma_x is None else sigma_x
    size_x = size if size_x is None else size_x
    st0 = size // 2
    ph = np.arange(oversampling) / oversampling

    def axis(n, sig):
        e = np.arange(n)[:, None] - st0 - ph[None, :]  # (n, OS)
        c = 1.0 / (sig * math.sqrt(2.0))
        return 0.5 * (erf((e + 1.0) * c) - erf(e * c))

    gy, gx = axis(size, sigma_y), axis(size_x, sigma_x)
    t = gy[:, None, :, None] * gx[None, :, None, :]
    return np.asfortranarray(t.astype(np.float32))


def make_rois(n_roi: int, seed: int, sides=(16,), density=None, mean_n=3.0, n_cap=None, photons=(5.0, 200.0),
              bg=(10.0, 10.0), sigma=1.3, pixel_um=0.1):
    """Synthetic ROIs.  Returns a dict of per-ROI arrays and the flat data pool.

    sides       ROI side choices (uniform)
    density     (lo, hi) emitters per um^2 (uniform per ROI); None -> Poisson(mean_n)
    photons     Gamma (k, theta)
    bg          (lo, hi) uniform background per ROI
    """
    rng = np.random.default_rng(seed)
    side = rng.choice(np.asarray(sides), size=n_roi)
    if density is not None:
        dens = rng.uniform(density[0], density[1], size=n_roi)
        lam = dens * (side * pixel_um) ** 2
    else:
        dens = np.full(n_roi, mean_n / (16 * 16 * pixel_um ** 2))
        lam = np.full(n_roi, mean_n, dtype=np.float64)
    nem = rng.poisson(lam)
    if n_cap is not None:
        nem = np.minimum(nem, n_cap)
    bgs = rng.uniform(bg[0], bg[1], size=n_roi)
    offs = np.zeros(n_roi + 1, np.int64)
    offs[1:] = np.cumsum(side.astype(np.int64) ** 2)
    data = np.zeros(offs[-1], np.float32)
    truth = []
    for r in range(n_roi):
        s = int(side[r])
        k = int(nem[r])
        X = rng.uniform(0, s, size=k)
        Y = rng.uniform(0, s, size=k)
        I = rng.gamma(photons[0], photons[1], size=k)
        mu = roi_expected(I, X, Y, bgs[r], s, s, sigma)
        data[offs[r]:offs[r + 1]] = rng.poisson(mu).astype(np.float32)
        truth.append((I, X, Y))
    return {"side": side.astype(np.int32), "offset": offs, "data": data, "n_true": nem.astype(np.int32),
            "bg": bgs.astype(np.float32), "density": dens, "rho": (dens * pixel_um ** 2).astype(np.float32),
            "truth": truth}


def make_frames(nframes: int, size: int, seed: int, density_um2=0.3, pixel_um=0.1, photons=(5.0, 200.0), bg=10.0,
                sigma=1.3, chunk=64):
    """Frames (nframes, size, size) float32 of Poisson counts; frame f, row y, col x at [f, y, x]
    (x contiguous).  Emitter positions are uniform in the frame."""
    rng = np.random.default_rng(seed)
    out = np.empty((nframes, size, size), np.float32)
    area = (size * pixel_um) ** 2
    half = 6
    for f0 in range(0, nframes, chunk):
        f1 = min(nframes, f0 + chunk)
        mu = np.full((f1 - f0, size, size), bg, np.float64)
        for f in range(f0, f1):
            k = rng.poisson(density_um2 * area)
            X = rng.uniform(0, size, size=k)
            Y = rng.uniform(0, size, size=k)
            I = rng.gamma(photons[0], photons[1], size=k)
            c = 1.0 / (sigma * SQRT2)
            for x, y, i in zip(X, Y, I):
                x0, x1 = max(0, int(x) - half), min(size, int(x) + half + 1)
                y0, y1 = max(0, int(y) - half), min(size, int(y) + half + 1)
                ex = erf((np.arange(x0, x1 + 1) - x) * c)
                ey = erf((np.arange(y0, y1 + 1) - y) * c)
                gx = 0.5 * (ex[1:] - ex[:-1])
                gy = 0.5 * (ey[1:] - ey[:-1])
                mu[f - f0, y0:y1, x0:x1] += i * np.outer(gy, gx)
        out[f0:f1] = rng.poisson(mu)
    return out


def make_mle_boxes(n: int, sz: int, seed: int, sigma=1.3, photons=(5.0, 200.0), bg=(5.0, 30.0), noise=True):
    """Single-emitter boxes for the MLE: box[i, jj, ii] = data[sz*jj + ii] with the emitter at
    (x along ii, y along jj).  Returns (boxes float32 (n, sz, sz), truth dict)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(sz / 2 - 1, sz / 2 + 1, n) - 0.5
    y = rng.uniform(sz / 2 - 1, sz / 2 + 1, n) - 0.5
    I = rng.gamma(photons[0], photons[1], n)
    b = rng.uniform(bg[0], bg[1], n)
    c = 1.0 / (sigma * SQRT2)
    e = np.arange(sz + 1, dtype=np.float64)
    gx = 0.5 * np.diff(erf((e[None, :] - x[:, None]) * c), axis=1)  # (n, sz) along ii
    gy = 0.5 * np.diff(erf((e[None, :] - y[:, None]) * c), axis=1)  # along jj
    mu = b[:, None, None] + I[:, None, None] * gy[:, :, None] * gx[:, None, :]  # [i, jj, ii]
    boxes = rng.poisson(mu).astype(np.float32) if noise else mu.astype(np.float32)
    return boxes, {"x": x, "y": y, "I": I, "bg": b}
