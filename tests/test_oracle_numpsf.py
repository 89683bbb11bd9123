"""The oracle's sampled-PSF model (NumModel, LibRJMCMC.cpp:62-170) and the
oracle chain with OptModel = 1, pinned by what the interpolation fixes:

  - a table linear in the emitter-to-pixel offset is reproduced exactly at any
    position (bilinear interpolation is exact on linear functions), which
    fixes the table's axis order, the signs of the ST/IM offsets, the
    interpolation weights and the wrap of the last phase onto phase 0 of the
    next pixel;
  - a table sampled from the erf model reproduces Model (LibRJMCMC.cpp:15-59)
    at the sampled phases and converges to it quadratically in between;
  - a table holding each entry's own linear index shows which entry every
    pixel reads at a sampled phase, the out-of-table reads landing on the last
    entry (LibRJMCMC.cpp:131-158);
  - the chain on a finely sampled erf table follows the erf chain step for
    step under the same draws.
The reference is MATLAB-mex C++ (unbuildable here); these are its semantics
restated, not its outputs."""
import numpy as np
import pytest

from bamf_amd import synth
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, oracle_params, params


def lin_table(p0, p1, os_, a, b, c):
    st0 = p0 // 2
    d0 = np.arange(p0)[:, None, None, None]
    d1 = np.arange(p1)[None, :, None, None]
    iy = np.arange(os_)[None, None, :, None]
    ix = np.arange(os_)[None, None, None, :]
    return np.asfortranarray((a + b * (d0 - st0 - iy / os_) + c * (d1 - st0 - ix / os_)).astype(np.float32))


def img(m, xs, ys):
    return np.asarray(m).reshape(xs, ys).T  # (y, x)


def test_linear_table_is_reproduced_exactly(oracle):
    p0, p1, os_ = 40, 36, 5
    a, b, c = 3.0, 0.25, -0.125
    T = lin_table(p0, p1, os_, a, b, c)
    xs, ys = 11, 12
    rng = np.random.default_rng(3)
    # random phases, phases past the last sample (the wrap branch), exact samples
    X = np.concatenate([rng.uniform(0.5, 10.5, 6), [3.9, 7.95, 2.0, 4.6]]).astype(np.float32)
    Y = np.concatenate([rng.uniform(0.5, 11.5, 6), [5.85, 1.99, 6.4, 3.0]]).astype(np.float32)
    I = rng.integers(1, 5, len(X)).astype(np.float32)
    bg, abg, bbg = 2.5, 0.125, -0.0625
    m = img(oracle.num_model(I, X, Y, bg, abg, bbg, xs, ys, T), xs, ys)
    yy, xx = np.mgrid[0:ys, 0:xs]
    want = bg + abg * xx + bbg * yy
    for k in range(len(X)):
        want = want + I[k] * (a + b * (yy - float(Y[k])) + c * (xx - float(X[k])))
    assert np.abs(m - want).max() < 2e-5 * (np.abs(want).max() + I.sum())
    # one emitter at a time: each lies on the same plane
    for k in range(len(X)):
        m1 = img(oracle.num_model(I[k:k + 1], X[k:k + 1], Y[k:k + 1], 0, 0, 0, xs, ys, T), xs, ys)
        w1 = I[k] * (a + b * (yy - float(Y[k])) + c * (xx - float(X[k])))
        assert np.abs(m1 - w1).max() < 1e-5 * I[k] * 10, k


def test_erf_table_matches_model_at_samples_and_converges(oracle):
    xs = ys = 14
    sigma = 1.3
    rng = np.random.default_rng(8)
    errs = {}
    for os_ in (4, 16):
        T = synth.sampled_gauss_psf(40, os_, sigma)
        # sampled phases: the table's own values
        X = (rng.integers(2, 12, 4) + rng.integers(0, os_, 4) / os_).astype(np.float32)
        Y = (rng.integers(2, 12, 4) + rng.integers(0, os_, 4) / os_).astype(np.float32)
        I = np.full(4, 500.0, np.float32)
        mn = oracle.num_model(I, X, Y, 10, 0, 0, xs, ys, T)
        ma = oracle.model(I, X, Y, 10, 0, 0, xs, ys, sigma)
        assert np.abs(mn - ma).max() < 1e-5 * 500, os_
        # between samples: bilinear error of a smooth function, O(1 / OS^2)
        X = rng.uniform(2, 12, 50).astype(np.float32)
        Y = rng.uniform(2, 12, 50).astype(np.float32)
        e = 0.0
        for k in range(50):
            mn = oracle.num_model(I[:1], X[k:k + 1], Y[k:k + 1], 0, 0, 0, xs, ys, T)
            ma = oracle.model(I[:1], X[k:k + 1], Y[k:k + 1], 0, 0, 0, xs, ys, sigma)
            e = max(e, float(np.abs(mn - ma).max()) / 500)
        errs[os_] = e
    assert errs[4] < 0.01 and errs[16] < errs[4] / 8, errs


def test_index_table_reads_and_clamps(oracle):
    p0, p1, p2, os_ = 6, 7, 2, 2
    n = p0 * p1 * p2 * os_
    T = np.asfortranarray(np.arange(n, dtype=np.float32).reshape((p0, p1, p2, os_), order="F"))
    xs, ys = 10, 9
    st0 = p0 // 2
    seen = {"low": 0, "high": 0, "in": 0}
    for X, Y in [(4.5, 3.0), (1.0, 7.5), (8.5, 0.5), (0.0, 0.0), (9.0, 8.5), (9.0, 8.0), (0.5, 0.5)]:
        m = img(oracle.num_model([2.0], [X], [Y], 0, 0, 0, xs, ys, T), xs, ys)
        iy, ix = int(round((Y % 1) * os_)), int(round((X % 1) * os_))
        yy, xx = np.mgrid[0:ys, 0:xs]
        d0 = st0 - int(np.floor(Y)) + yy
        d1 = st0 - int(np.floor(X)) + xx
        lin = d0 + p0 * (d1 + p1 * (iy + p2 * ix))
        low, high = lin < 0, lin >= n
        seen["low"] += low.sum()
        seen["high"] += high.sum()
        seen["in"] += (~low & ~high).sum()
        want = 2.0 * np.where(low | high, n - 1, lin)
        assert np.array_equal(m, want.astype(np.float32)), (X, Y)
    assert min(seen.values()) > 0, seen


def test_sampled_erf_chain_follows_erf_chain(oracle):
    rois = synth.make_rois(6, seed=44, sides=(16,), mean_n=3.0, n_cap=8)
    p = params(nb=400, nt=400)
    op = oracle_params(oracle, p)
    T = synth.sampled_gauss_psf(40, 32, p.psf_sigma)
    matched = []
    for r in range(6):
        o = rois["offset"]
        d = rois["data"][o[r]:o[r + 1]]
        a = oracle.rjmcmc(d, 16, 16, op, OFF_PDF, SIG_PDF, BACK_PDF, 10.0, seed=21, roi=r)
        b = oracle.rjmcmc(d, 16, 16, op, OFF_PDF, SIG_PDF, BACK_PDF, 10.0, seed=21, roi=r, psf=T)
        assert a["status"] == 0 and b["status"] == 0
        sa, sb = a["steps"], b["steps"]
        diff = np.flatnonzero((sa["flags"] != sb["flags"]) | (sa["kind"] != sb["kind"]) | (sa["n"] != sb["n"]))
        k = int(diff[0]) if len(diff) else len(sa)
        matched.append(k)
        la, lb = sa["log_pr"][:k], sb["log_pr"][:k]
        fin = np.isfinite(la) & np.isfinite(lb)
        assert np.all(np.abs(la[fin] - lb[fin]) <= 0.02 * (1 + np.abs(la[fin]))), r
    assert np.median(matched) >= 300, matched


def test_chain_rejects_a_missing_table(oracle):
    p = params(nb=10, nt=10)
    op = oracle_params(oracle, p)
    with pytest.raises(ValueError):
        oracle.rjmcmc(np.zeros(64, np.float32), 8, 8, op, OFF_PDF, SIG_PDF, BACK_PDF, 10.0,
                      psf=np.zeros((4, 4, 2), np.float32))
