"""Pins the oracle's FindROI restatement (oracle/findroi.c) without the reference.

  - coefficients: the published Young-van Vliet formulas in float64, then single;
  - the recursive filter keeps constants, is linear, and its impulse response
    has the second moment of a Gaussian of the requested sigma;
  - maxX / maxY against brute-force window maxima written with numpy;
  - detections of isolated blobs land on the blob pixels, in MATLAB find()
    order, with box corners clamped as findROI.m:130-148 describes.
"""
import math

import numpy as np
import pytest


def yvv(sigma):
    sigma = float(sigma)  # MATLAB evaluates in double
    q = 0.98711 * sigma - 0.96330 if sigma > 2.5 else 3.97156 - 4.14554 * math.sqrt(1 - 0.26891 * sigma)
    qq = q * q
    qqq = qq * q  # findROI.m:175-176 forms q^3 as (q*q)*q
    b0 = 1.57825 + 2.44413 * q + 1.4281 * qq + 0.422205 * qqq
    b1 = 2.44413 * q + 2.85619 * qq + 1.26661 * qqq
    b2 = -1.4281 * qq - 1.26661 * qqq
    b3 = 0.422205 * qqq
    return [np.float32(v) for v in (b0, b1, b2, b3, 1 - (b1 + b2 + b3) / b0)]


@pytest.mark.parametrize("sigma", [0.8, 1.3, 2.6, 3.5])
def test_coeffs(oracle, sigma):
    assert oracle.gauss_coeffs(sigma) == [float(v) for v in yvv(np.float32(sigma))]


def test_constant_and_linearity(oracle):
    L = oracle.lib()
    c = oracle.gauss_coeffs(1.3)
    x = np.full((3, 16, 40), 7.0, np.float32)
    L.orc_gauss_x(x.ctypes.data, 40, 16, 3, *c)
    L.orc_gauss_y(x.ctypes.data, 40, 16, 3, *c)
    np.testing.assert_allclose(x, 7.0, rtol=2e-5)
    rng = np.random.default_rng(0)
    a = rng.uniform(0, 10, (1, 32, 32)).astype(np.float32)
    b = rng.uniform(0, 10, (1, 32, 32)).astype(np.float32)
    s = a + b
    for arr in (a, b, s):
        L.orc_gauss_x(arr.ctypes.data, 32, 32, 1, *c)
        L.orc_gauss_y(arr.ctypes.data, 32, 32, 1, *c)
    np.testing.assert_allclose(s, a + b, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("sigma", [1.3, 2.6])
def test_impulse_response_matches_lfilter(oracle, sigma):
    """The causal + anti-causal recursion equals scipy.signal.lfilter run forward
    and backward in float64 (zero state away from an isolated impulse)."""
    from scipy.signal import lfilter

    L = oracle.lib()
    b0, b1, b2, b3, B = oracle.gauss_coeffs(sigma)
    n = 257
    x = np.zeros((1, 1, n), np.float32)
    x[0, 0, n // 2] = 1.0
    L.orc_gauss_x(x.ctypes.data, n, 1, 1, b0, b1, b2, b3, B)
    h = x.ravel().astype(np.float64)
    imp = np.zeros(n)
    imp[n // 2] = 1.0
    a = [1.0, -b1 / b0, -b2 / b0, -b3 / b0]
    ref = lfilter([B], a, lfilter([B], a, imp)[::-1])[::-1]
    np.testing.assert_allclose(h, ref, atol=2e-7)
    pos = np.arange(n) - n // 2
    assert abs(h.sum() - 1) < 1e-4
    assert abs(np.sum(h * pos)) < 1e-4  # symmetric
    # Young-van Vliet (1995) overestimates small sigmas; the width is still of order sigma
    assert 0.8 * sigma < math.sqrt(np.sum(h * pos ** 2)) < 1.25 * sigma


def brute_maxx(a, k, minval):
    z, ys, xs = a.shape
    out = np.empty_like(a)
    for x in range(xs):
        w = a[:, :, max(0, x - k):min(xs, x + k + 1)]
        m = np.maximum(np.float32(minval), w.max(axis=2))
        v = a[:, :, x]
        out[:, :, x] = np.where(m > v, -m, np.where(m == v, m, np.float32(0)))
    return out


def brute_maxy(b, k, minval):
    z, ys, xs = b.shape
    out = np.empty_like(b)
    for y in range(ys):
        w = np.abs(b[:, max(0, y - k):min(ys, y + k + 1), :])
        m = np.maximum(np.float32(minval), w.max(axis=1))
        out[:, y, :] = (np.abs(m - b[:, y, :]).astype(np.float64) < 1e-6).astype(np.float32)
    return out


@pytest.mark.parametrize("k", [1, 3, 7])
def test_window_max_bruteforce(oracle, k):
    L = oracle.lib()
    rng = np.random.default_rng(k)
    a = rng.normal(0, 10, (2, 20, 24)).astype(np.float32)
    a[0, 3, 5] = a[0, 3, 6]  # ties
    b = np.zeros_like(a)
    L.orc_max_x(a.ctypes.data, b.ctypes.data, 24, 20, 2, k, np.float32(1.5))
    np.testing.assert_array_equal(b, brute_maxx(a, k, 1.5))
    r = np.zeros_like(a)
    L.orc_max_y(r.ctypes.data, b.ctypes.data, 24, 20, 2, k, np.float32(1.5))
    np.testing.assert_array_equal(r, brute_maxy(b, k, 1.5))


def test_isolated_blobs_order_and_corners(oracle):
    ys, xs, nz = 40, 40, 2
    img = np.full((nz, ys, xs), 10.0, np.float32)  # [frame, y(run), x(contiguous)]
    blobs = [(0, 20, 20), (0, 1, 30), (0, 25, 37), (1, 35, 2), (1, 10, 10)]
    yy, xx = np.mgrid[0:ys, 0:xs]
    for f, y, x in blobs:
        img[f] += 2000 * np.exp(-((yy - y) ** 2 + (xx - x) ** 2) / (2 * 1.3 ** 2)).astype(np.float32)
    res, lin, row, col, fr = oracle.find_roi(img.ravel(), xs, ys, nz, 1.3, 2.6, 7, 23.5)
    # flag pixels are the blob centres (kernel layout: lin = f*X*Y + y*X + x)
    want = sorted(f * xs * ys + y * xs + x for f, y, x in blobs)
    assert list(lin) == want and res.sum() == len(blobs)
    # ind2sub with MATLAB dims [Ysize Xsize Z]: row = lin % Ysize + 1, col = (lin // Ysize) % Xsize + 1
    half = math.floor(7 / 2 + 0.5)
    for l, r, c, f in zip(lin, row, col, fr):
        ii, jj, kk = l % ys + 1, (l // ys) % xs + 1, l // (xs * ys) + 1
        sr = min(max(ii - half + 1, 1), ys - 7 + 1)
        sc = min(max(jj - half + 1, 1), xs - 7 + 1)
        assert (r, c, f) == (sr, sc, kk)
