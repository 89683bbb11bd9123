"""Device forward-model generator (bamf_synth_frames, bamf_synth_roi_batch)
against independent references:
  - noiseless frames and ROIs equal the float64 pixel-integrated Gaussian of
    their own reported emitters (scipy erf), and the training targets equal a
    numpy restatement of the occupancy rule;
  - per-cell emitter counts are Poisson(density x cell area), positions are
    uniform in their cell, intensities Gamma(k, theta) (chi-square / KS);
  - pixel counts of flat frames are Poisson(mu) on both sides of the sampler
    switch at mu = 10 (chi-square, mean and variance);
  - the output is a deterministic function of the seed."""
import numpy as np
import pytest
from scipy import special, stats

from bamf_amd import _hip
from bamf_amd._lib import check, lib

pytestmark = pytest.mark.gpu
CAP = 128


def frames(nf, ys, xs, density, k=5.0, theta=200.0, bg=10.0, sigma=1.3, seed=7, noise=1):
    nc = lib().bamf_synth_cells(ys, xs)
    df = _hip.DeviceBuffer(nf * ys * xs * 4)
    dt = _hip.DeviceBuffer(nf * nc * CAP * 12)
    dc = _hip.DeviceBuffer(nf * nc * 4)
    check(lib().bamf_synth_frames(nf, ys, xs, density, k, theta, bg, sigma, seed, CAP, noise, df.vp, dt.vp, dc.vp,
                                  None), "bamf_synth_frames")
    _hip.synchronize()
    cnt = dc.download(np.int32).reshape(nf, nc)
    tru = dt.download(np.float32).reshape(nf, nc, CAP, 3)
    return df.download(np.float32).reshape(nf, ys, xs), tru, cnt


def model(X, Y, I, bg, ys, xs, sigma):
    c = 1.0 / (sigma * np.sqrt(2.0))
    ex = 0.5 * np.diff(special.erf((np.arange(xs + 1)[None, :] - X[:, None]) * c), axis=1)
    ey = 0.5 * np.diff(special.erf((np.arange(ys + 1)[None, :] - Y[:, None]) * c), axis=1)
    return bg + np.einsum("e,ey,ex->yx", I, ey, ex)


def test_noiseless_frames_equal_the_model_of_their_emitters():
    nf, ys, xs = 3, 70, 90  # ragged last cells in both directions
    img, tru, cnt = frames(nf, ys, xs, 0.04, noise=0)
    for f in range(nf):
        em = np.concatenate([tru[f, c, :cnt[f, c]] for c in range(cnt.shape[1])]).astype(np.float64)
        assert len(em) > 100
        assert em[:, 0].min() >= 0 and em[:, 0].max() < xs and em[:, 1].min() >= 0 and em[:, 1].max() < ys
        mu = model(em[:, 0], em[:, 1], em[:, 2], 10.0, ys, xs, 1.3)
        np.testing.assert_allclose(img[f], mu, rtol=2e-5, atol=1e-4)


def test_cells_are_a_poisson_process_with_gamma_intensities():
    nf, ys, xs, dens = 400, 64, 64, 0.03
    _, tru, cnt = frames(nf, ys, xs, dens, k=4.0, theta=125.0, noise=0, seed=11)
    lam = dens * 256
    c = cnt.ravel()
    assert abs(c.mean() / lam - 1) < 0.01 and abs(c.var() / lam - 1) < 0.05
    kmax = int(lam + 6 * np.sqrt(lam))
    obs = np.bincount(np.minimum(c, kmax), minlength=kmax + 1)
    p = stats.poisson.pmf(np.arange(kmax + 1), lam)
    p[-1] += stats.poisson.sf(kmax, lam)
    keep = p * c.size > 5
    chi = stats.chisquare(obs[keep], p[keep] / p[keep].sum() * obs[keep].sum())
    assert chi.pvalue > 1e-3, chi
    em = np.concatenate([tru[f, q, :cnt[f, q]] for f in range(nf) for q in range(cnt.shape[1])])
    assert stats.kstest(np.mod(em[:, 0], 16) / 16, "uniform").pvalue > 1e-3
    assert stats.kstest(np.mod(em[:, 1], 16) / 16, "uniform").pvalue > 1e-3
    assert stats.kstest(em[:, 2], stats.gamma(4.0, scale=125.0).cdf).pvalue > 1e-3
    # a Gamma with another shape is told apart
    assert stats.kstest(em[:, 2], stats.gamma(4.4, scale=125.0 * 4 / 4.4).cdf).pvalue < 1e-6


@pytest.mark.parametrize("mu", [0.3, 4.0, 9.99, 10.0, 37.0, 300.0])
def test_pixel_counts_are_poisson(mu):
    img, _, cnt = frames(16, 128, 128, 0.0, bg=mu, seed=int(mu * 100))
    assert cnt.sum() == 0
    v = img.ravel().astype(np.int64)
    assert np.all(img == np.round(img))
    n = v.size
    assert abs(v.mean() - mu) < 5 * np.sqrt(mu / n) and abs(v.var() / mu - 1) < 0.03
    lo, hi = int(stats.poisson.ppf(1e-4, mu)), int(stats.poisson.ppf(1 - 1e-4, mu))
    k = np.arange(lo, hi + 1)
    obs = np.array([(v == j).sum() for j in k])
    obs[0] += (v < lo).sum()
    obs[-1] += (v > hi).sum()
    p = stats.poisson.pmf(k, mu)
    p[0] += stats.poisson.cdf(lo - 1, mu)
    p[-1] += stats.poisson.sf(hi, mu)
    keep = p * n > 5
    chi = stats.chisquare(obs[keep], p[keep] / p[keep].sum() * obs[keep].sum())
    assert chi.pvalue > 1e-4, (mu, chi)


def test_frames_are_deterministic():
    a = frames(2, 50, 40, 0.05, seed=5)
    b = frames(2, 50, 40, 0.05, seed=5)
    c = frames(2, 50, 40, 0.05, seed=6)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])
    assert not np.array_equal(a[0], c[0])


def roi_batch(n, side, noise, seed=3, d=(0.005, 0.05), bg=(5.0, 30.0)):
    P = side * side
    img, occ, logi = (_hip.DeviceBuffer(n * P * 4) for _ in range(3))
    nem, tru = _hip.DeviceBuffer(n * 4), _hip.DeviceBuffer(n * (3 * CAP + 1) * 4)
    check(lib().bamf_synth_roi_batch(n, side, d[0], d[1], bg[0], bg[1], 5.0, 200.0, 1.3, seed, noise, img.vp, occ.vp,
                                     logi.vp, nem.vp, tru.vp, None), "bamf_synth_roi_batch")
    _hip.synchronize()
    return (img.download(np.float32).reshape(n, side, side), occ.download(np.float32).reshape(n, side, side),
            logi.download(np.float32).reshape(n, side, side), nem.download(np.int32),
            tru.download(np.float32).reshape(n, 3 * CAP + 1))


@pytest.mark.parametrize("side", [16, 13, 32])
def test_roi_batch_model_and_targets(side):
    n = 64
    img, occ, logi, nem, tru = roi_batch(n, side, noise=0)
    assert nem.sum() > 20
    for r in range(n):
        k = nem[r]
        em = tru[r, 1:1 + 3 * k].reshape(k, 3).astype(np.float64)
        # [x][y] layout: transpose the (y, x) model
        mu = model(em[:, 0], em[:, 1], em[:, 2], float(tru[r, 0]), side, side, 1.3).T
        np.testing.assert_allclose(img[r], mu, rtol=2e-5, atol=1e-4)
        isum = np.zeros((side, side))
        for x, y, i in em:
            isum[min(int(np.floor(x + 0.5)), side - 1), min(int(np.floor(y + 0.5)), side - 1)] += i
        assert np.array_equal(occ[r], (isum > 0).astype(np.float32))
        np.testing.assert_allclose(logi[r][isum > 0], np.log(isum[isum > 0]), rtol=1e-5)
        assert np.all(logi[r][isum == 0] == 0)
    noisy = roi_batch(n, side, noise=1)[0]
    assert np.all(noisy == np.round(noisy)) and not np.array_equal(noisy, img)
