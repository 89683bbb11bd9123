"""Oracle (test infrastructure only): numpy/scipy restatement of the prior
construction of @RJ/findPriors.m:41-80 after the fits, and of
@RJ/threhsSMA.m:1-32.

  threshold         threhsSMA.m:1-17 ('Basic': k = BoxSize^2 - 4), including
                    its whole-array isreal() tests (one negative X or Y CRLB,
                    whose square root is complex, rejects every fit)
  matlab_quantile   quantile(x, p) for a vector (sorted samples at (i - 0.5)/n)
  kde_pdf           pdf(fitdist(x, 'Kernel', 'BandWidth', h), grid), normal kernel
  priors_from_fits  findPriors.m:43-80 on the fits

Pinned by tests/test_oracle_priors.py against scipy and numpy routines that
compute the same quantities independently (chi2.sf, gaussian_kde,
gamma.fit, np.quantile(method='hazen'), np.gradient) and against hand-built
photon sets with and without a second mode.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import brentq
from scipy.special import digamma, gammainc, gammaln


def threshold(params, crlb, ll, box, max_xy_se, min_pvalue, min_photons):
    """threhsSMA.m:1-17 on gaussMLE outputs (parameter-major: Y, X, Photons, BG)."""
    y, x, ph, bg = (np.asarray(params[k], np.float32) for k in range(4))
    cy, cx = np.asarray(crlb[0], np.float32), np.asarray(crlb[1], np.float32)
    if np.any(cx < 0) or np.any(cy < 0):  # isreal(SMD.X_SE) / isreal(SMD.Y_SE) are false
        return np.zeros(len(ph), bool)
    with np.errstate(invalid="ignore"):
        x_se, y_se = np.sqrt(cx), np.sqrt(cy)
    k = box * box - 4
    x2 = -2.0 * np.asarray(ll, np.float64)
    pv = 1.0 - gammainc(k / 2.0, x2 / 2.0)  # gammainc(X2/2, k/2) in MATLAB's order
    return ((x_se < max_xy_se) & (y_se < max_xy_se) & (pv > min_pvalue) & (ph > min_photons) & (x > 0) & (y > 0)
            & (x < box) & (y < box) & (bg > 0))


def matlab_quantile(v, p):
    s = np.sort(np.asarray(v, np.float64))
    n = len(s)
    pos = n * p + 0.5
    if pos <= 1:
        return s[0]
    if pos >= n:
        return s[-1]
    f = int(np.floor(pos))
    return s[f - 1] + (pos - f) * (s[f] - s[f - 1])


def kde_pdf(x, h, grid):
    x = np.asarray(x, np.float64)
    out = np.zeros(len(grid))
    for g0 in range(0, len(grid), 2048):
        u = (np.asarray(grid[g0:g0 + 2048], np.float64)[:, None] - x[None, :]) / h
        out[g0:g0 + 2048] = np.exp(-0.5 * u * u).sum(1)
    return out / (len(x) * h * np.sqrt(2 * np.pi))


def grid_of(photon, dp):
    end = float(np.max(photon)) + 1000.0 * dp  # findPriors.m:47
    return np.arange(int(np.floor(end / dp * (1 + 4e-16))) + 1) * dp


def gamma_mle(x):
    x = np.asarray(x, np.float64)
    s = np.log(x.mean()) - np.log(x).mean()
    a = brentq(lambda a: np.log(a) - digamma(a) - s, 1e-6, 1e9, xtol=1e-14, rtol=1e-15)
    return a, x.mean() / a


def gampdf(x, a, b):
    x = np.asarray(x, np.float64)
    with np.errstate(divide="ignore"):
        out = np.exp((a - 1) * np.log(np.where(x > 0, x, 1.0)) - x / b - gammaln(a) - a * np.log(b))
    out = np.where(x > 0, out, 0.0)
    if a < 1:
        out = np.where(x == 0, np.inf, out)
    elif a == 1:
        out = np.where(x == 0, 1 / b, out)
    return out


def priors_from_fits(params, crlb, ll, psf_sigma, mean_photons, max_xy_se, min_pvalue, min_photons, p_dp,
                     bg_prior_range, bg_prior, box=8):
    """findPriors.m:41-80 given the gaussMLE outputs of its ROIs."""
    keep = threshold(params, crlb, ll, box, max_xy_se, min_pvalue, min_photons)
    photon = np.asarray(params[2], np.float32)[keep]
    bg = np.asarray(params[3], np.float32)[keep]
    nkept = len(photon)
    if len(photon) > 200:
        idv = np.float32(matlab_quantile(photon, 0.98))
        photon = photon[~(photon > idv)]
    grid = grid_of(photon, p_dp)
    pdf = kde_pdf(photon, 75.0, grid)
    gr = np.gradient(pdf)
    ind = np.nonzero(np.concatenate([[0], np.diff(np.sign(gr))]))[0] + 1  # 1-based
    refit = 0
    if len(ind):
        imax = np.nonzero(pdf == pdf[ind - 1].max())[0][0] + 1
        rem = ind[ind > imax]
        if len(rem) >= 2:
            photon = photon[~(photon.astype(np.float64) > rem[0] * p_dp)]
            grid = grid_of(photon, p_dp)
            pdf = kde_pdf(photon, 75.0, grid)
            refit = 1
    a, b = gamma_mle(bg)
    offset = gampdf(grid, a, b)
    mean_ph = float(np.mean(photon.astype(np.float64)))
    mb = mean_ph / bg_prior_range
    bgp = np.exp(-grid / mb) / mb if bg_prior else np.zeros(len(grid))
    return pdf, bgp, offset, {"kept": nkept, "photons": len(photon), "a": a, "b": b, "refit": refit,
                              "mean_photons": mean_ph}
