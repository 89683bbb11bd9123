"""Effective sample size of MCMC traces and the ROI-posterior criterion.

ESS uses Geyer's initial monotone sequence estimator (Geyer 1992, "Practical
Markov chain Monte Carlo", Statistical Science 7:473-483): autocovariances by
FFT, pairs Gamma_k = gamma_2k + gamma_2k+1 summed while positive and made
monotone non-increasing; tau = -1 + 2 sum Gamma_k; ESS = n / tau.

A trace with zero variance carries no autocorrelation information; it is
reported as ESS = n (the chain never left one value, e.g. a fixed emitter
count).
"""
from __future__ import annotations

import numpy as np


def ess_batch(x: np.ndarray) -> np.ndarray:
    """ESS of each row of x (n_chains, n_samples)."""
    x = np.asarray(x, np.float64)
    m, n = x.shape
    out = np.full(m, float(n))
    if n < 4:
        return out
    xc = x - x.mean(axis=1, keepdims=True)
    var = (xc * xc).mean(axis=1)
    live = var > 0
    if not live.any():
        return out
    xc = xc[live]
    nfft = 1 << (2 * n - 1).bit_length()
    f = np.fft.rfft(xc, nfft, axis=1)
    acov = np.fft.irfft(f * np.conj(f), nfft, axis=1)[:, :n] / n
    rho = acov / acov[:, :1]
    npair = n // 2
    pairs = rho[:, 0:2 * npair:2] + rho[:, 1:2 * npair:2]
    # initial positive sequence, then monotone
    pos = pairs > 0
    first_neg = np.where(pos.all(axis=1), npair, np.argmin(pos, axis=1))
    mask = np.arange(npair)[None, :] < first_neg[:, None]
    pairs = np.where(mask, pairs, 0.0)
    pairs = np.minimum.accumulate(np.where(mask, pairs, np.inf), axis=1)
    pairs = np.where(mask, pairs, 0.0)
    tau = -1.0 + 2.0 * pairs.sum(axis=1)
    tau = np.maximum(tau, 1.0 / n)
    out[live] = np.minimum(n / tau, float(n) * np.log10(max(n, 10)))
    return out


def roi_posteriors(trace_n: np.ndarray, trace_photons: np.ndarray, ess_target: float = 200.0):
    """Per-ROI ESS (minimum over the emitter-count and total-photon traces) and
    the mask of ROIs that reach the target."""
    e = np.minimum(ess_batch(trace_n.astype(np.float64)), ess_batch(trace_photons.astype(np.float64)))
    return e, e >= ess_target


def fixed_ess_rate(ess: np.ndarray, seconds: np.ndarray, burn_frac: float, target: float = 200.0) -> float:
    """ROI posteriors per second at a fixed ESS.

    ROI i's share of the measured run time, seconds[i], is rescaled to the
    time its chain needs to reach `target` ESS: the burn-in part is kept and
    the post-burn-in part scales by target / ESS_i (the ESS of a stationary
    chain grows linearly with its length).  The rate is the number of ROIs
    over the sum of those times, so ROIs that mix slowly weigh in as the
    longer runs they would need; ESS is floored at 1.
    """
    e = np.maximum(np.asarray(ess, np.float64), 1.0)
    t = np.asarray(seconds, np.float64) * (burn_frac + (1.0 - burn_frac) * target / e)
    return float(len(e) / t.sum())
