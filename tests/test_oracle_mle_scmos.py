"""Pins the oracle's sCMOS MLE restatement (orc_mle_scmos: cuda_gaussMLEv2.cu
kernel_MLEFit_SCMOS*) without the reference.  The sCMOS model treats d + g as
Poisson with mean mu + g, g the pixel's variance/gain^2 read from a map at the
box corner (Huang et al., Nat. Methods 2013).  Checks:
  - noiseless boxes (d = mu): the fit returns the generating parameters;
  - noisy boxes: the fit is the float64 maximiser of
    sum((d + g) log(mu + g) - (mu + g)) over (x, y, I, bg) with a random map,
    and reading the map transposed (a plausible slip) gives another answer;
  - CRLB and log-likelihood: the reference's CRLB/LL loop reuses the g of the
    last pixel of the last Newton iteration for every pixel (the gain index is
    only formed in the Newton loop), so both equal the closed forms with that
    one g: diag(inv(J^T J / (mu + g_last))) from finite differences, and
    sum((d + g) log(mu + g) - mu - (d + g) log(d + g) + d) with g = g_last.
"""
import numpy as np
import pytest
from scipy.optimize import minimize
from scipy.special import erf

SIG = 1.3
MAPSZ = 40


def model64(theta, sz, sigma=SIG):
    x, y, n, bg = theta[:4]
    e = np.arange(sz + 1, dtype=np.float64)
    c = 1 / (sigma * np.sqrt(2))
    return bg + n * np.outer(0.5 * np.diff(erf((e - y) * c)), 0.5 * np.diff(erf((e - x) * c)))  # [jj, ii]


def setup(n, sz, seed, noise):
    rng = np.random.default_rng(seed)
    gain = rng.uniform(0.5, 40.0, size=(MAPSZ, MAPSZ)).astype(np.float32)  # [row = y, col = x]
    coords = np.stack([rng.integers(0, MAPSZ - sz, n), rng.integers(0, MAPSZ - sz, n)], 1).astype(np.float32)
    th = np.stack([rng.uniform(sz / 2 - 1, sz / 2 + 1, n) - 0.5, rng.uniform(sz / 2 - 1, sz / 2 + 1, n) - 0.5,
                   rng.uniform(300, 2000, n), rng.uniform(5, 20, n)], 1)
    boxes = np.zeros((n, sz, sz), np.float32)
    g = np.zeros((n, sz, sz))
    for i in range(n):
        x0, y0 = int(coords[i, 0]), int(coords[i, 1])
        g[i] = gain[y0:y0 + sz, x0:x0 + sz]  # [jj, ii]
        mu = model64(th[i], sz)
        boxes[i] = (rng.poisson(mu + g[i]) - g[i]) if noise else mu
    return boxes, coords, gain, th, g


def test_noiseless_recovers_truth(oracle):
    boxes, coords, gain, th, _ = setup(24, 9, 61, noise=False)
    for ft in (0, 1):
        p, _, _ = oracle.mle_scmos(ft, boxes, coords, gain, SIG, 40)
        np.testing.assert_allclose(p[0], th[:, 0], atol=2e-3)
        np.testing.assert_allclose(p[1], th[:, 1], atol=2e-3)
        np.testing.assert_allclose(p[2], th[:, 2], rtol=2e-3)
        if ft == 1:
            np.testing.assert_allclose(p[4], SIG, rtol=2e-3)


@pytest.mark.parametrize("seed", [62, 63])
def test_noisy_fit_is_scmos_mle(oracle, seed):
    sz = 9
    boxes, coords, gain, th, g = setup(8, sz, seed, noise=True)
    p, _, _ = oracle.mle_scmos(0, boxes, coords, gain, SIG, 40)
    pt, _, _ = oracle.mle_scmos(0, boxes, coords, np.ascontiguousarray(gain.T), SIG, 40)
    moved = 0
    for i in range(8):
        d = boxes[i].astype(np.float64)

        def nll(t):
            mu = model64(t, sz)
            if np.any(mu + g[i] <= 0):
                return 1e30
            return float(np.sum((mu + g[i]) - (d + g[i]) * np.log(mu + g[i])))

        r = minimize(nll, th[i], method="Nelder-Mead",
                     options={"xatol": 1e-8, "fatol": 1e-11, "maxiter": 20000, "maxfev": 20000})
        assert abs(p[0, i] - r.x[0]) < 2e-3 and abs(p[1, i] - r.x[1]) < 2e-3
        assert abs(p[2, i] - r.x[2]) < 2e-3 * r.x[2]
        moved += abs(pt[2, i] - r.x[2]) > 2e-3 * r.x[2]
    assert moved >= 4  # the transposed map is a different likelihood


def test_crlb_and_ll_use_the_last_pixels_gain(oracle):
    sz = 9
    boxes, coords, gain, th, g = setup(8, sz, 64, noise=True)
    p, c, ll = oracle.mle_scmos(0, boxes, coords, gain, SIG, 40)
    for i in range(8):
        gl = g[i][sz - 1, sz - 1]  # pixel ii = jj = sz-1, the last of the Newton loop
        t = p[:, i].astype(np.float64)
        mu = model64(t, sz)
        J = np.zeros((sz * sz, 4))
        for k in range(4):
            h = 1e-4 * max(1.0, abs(t[k]))
            tp, tm = t.copy(), t.copy()
            tp[k] += h
            tm[k] -= h
            J[:, k] = ((model64(tp, sz) - model64(tm, sz)) / (2 * h)).ravel()
        F = J.T @ (J / (mu.ravel() + gl)[:, None])
        np.testing.assert_allclose(c[:, i], np.diag(np.linalg.inv(F)), rtol=1e-2)
        d = boxes[i].astype(np.float64)
        want = np.sum((d + gl) * np.log(mu + gl) - mu - (d + gl) * np.log(d + gl) + d)
        assert abs(ll[i] - want) < 1e-3 * max(1.0, abs(want))
