"""Pins the prior-construction oracle (oracle/priors.py) against independent
library routines and constructed cases:
  - the threshold's p-value 1 - gammainc(X2/2, k/2) equals chi2.sf(X2, k), and
    one negative CRLB rejects every fit (threhsSMA.m's whole-array isreal);
  - matlab_quantile equals numpy's 'hazen' quantile (MATLAB's definition);
  - kde_pdf equals scipy's gaussian_kde with the same bandwidth;
  - gamma_mle equals scipy.stats.gamma.fit with the location fixed at 0;
  - gampdf equals scipy.stats.gamma.pdf;
  - the second-mode cut fires for a photon set with a separated second mode
    above the main one and not for a unimodal set, and cuts at Ind*P_dP.
"""
import numpy as np
import pytest
from scipy import stats

from oracle import priors as op


def test_pvalue_is_chi_square_tail():
    rng = np.random.default_rng(1)
    n, box = 500, 8
    params = np.stack([rng.uniform(1, 7, n), rng.uniform(1, 7, n), rng.uniform(100, 3000, n), rng.uniform(1, 20, n)])
    crlb = np.stack([rng.uniform(0.001, 0.01, n)] * 2 + [np.ones(n)] * 2)
    ll = -rng.uniform(10, 120, n)
    keep = op.threshold(params, crlb, ll, box, 0.2, 0.01, 50)
    pv = stats.chi2.sf(-2 * ll, box * box - 4)
    assert np.array_equal(keep, pv > 0.01)
    crlb2 = crlb.copy()
    crlb2[1, 7] = -1e-6
    assert not op.threshold(params, crlb2, ll, box, 0.2, 0.01, 50).any()


def test_quantile_is_hazen():
    rng = np.random.default_rng(2)
    for n in (3, 10, 201, 1000):
        v = rng.gamma(3, 200, n)
        for p in (0.02, 0.5, 0.98, 0.999):
            assert abs(op.matlab_quantile(v, p) - np.quantile(v, p, method="hazen")) < 1e-9 * v.max()


def test_kde_is_gaussian_kde():
    rng = np.random.default_rng(3)
    x = rng.gamma(4, 300, 3000)
    grid = np.arange(0, 4000, 10.0)
    kde = stats.gaussian_kde(x, bw_method=75.0 / np.std(x, ddof=1))
    np.testing.assert_allclose(op.kde_pdf(x, 75.0, grid), kde(grid), rtol=1e-10, atol=1e-16)


def test_gamma_mle_is_scipy_fit():
    rng = np.random.default_rng(4)
    for a, b in ((2.5, 4.0), (8.0, 1.5), (0.8, 10.0)):
        x = rng.gamma(a, b, 4000)
        fa, _, fb = stats.gamma.fit(x, floc=0)
        ga, gb = op.gamma_mle(x)
        assert abs(ga - fa) < 1e-4 * fa and abs(gb - fb) < 1e-4 * fb
        grid = np.linspace(0, 60, 50)
        np.testing.assert_allclose(op.gampdf(grid[1:], ga, gb), stats.gamma.pdf(grid[1:], ga, scale=gb), rtol=1e-10)


def make_fits(photons, rng):
    n = len(photons)
    params = np.stack([rng.uniform(2, 6, n), rng.uniform(2, 6, n), photons, rng.gamma(6, 2, n)]).astype(np.float32)
    crlb = np.stack([np.full(n, 0.005)] * 2 + [np.ones(n)] * 2).astype(np.float32)
    ll = np.full(n, -25.0, np.float32)
    return params, crlb, ll


def test_second_mode_cut():
    rng = np.random.default_rng(5)
    main = rng.normal(1500, 200, 3000)
    second = rng.normal(4500, 150, 400)
    params, crlb, ll = make_fits(np.concatenate([main, second]), rng)
    pdf, _, _, info = op.priors_from_fits(params, crlb, ll, 1.3, 1000, 0.2, 0.01, 50, 5.0, 5.0, 1)
    assert info["refit"] == 1 and info["photons"] < 3300 and info["photons"] > 2800
    params, crlb, ll = make_fits(main, rng)
    pdf2, bgp, off, info2 = op.priors_from_fits(params, crlb, ll, 1.3, 1000, 0.2, 0.01, 50, 5.0, 5.0, 1)
    assert info2["refit"] == 0 and info2["photons"] == 3000 - int(np.sum(main > op.matlab_quantile(
        params[2][op.threshold(params, crlb, ll, 8, 0.2, 0.01, 50)], 0.98)))
    # the grid, the density's mass and the exponential background prior
    assert abs(pdf2.sum() * 5.0 - 1) < 1e-3
    assert abs(bgp[0] - 5.0 / info2["mean_photons"]) < 1e-12
