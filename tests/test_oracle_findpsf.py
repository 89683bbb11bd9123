"""The oracle's restatement of @RJ/findPSF.m (oracle/findpsf.py), pinned by:

  - the shift theorem: a ROI circularly shifted by one pixel along x (or y),
    with its fitted position moved by one, adds the identical spectrum;
  - the split into the 4 x 4 phase images: before the moves and the
    normalisation the sixteen images hold every pixel of the padded image
    exactly once, each on rows and columns 4 apart; the moves shift the middle
    third (rows / columns 21..43) by one; each image then sums to 1;
  - the round trip that gives the table its meaning: ROIs of an elliptical
    pixel-integrated Gaussian at known positions, through findPSF, give a
    table with which NumModel (LibRJMCMC.cpp:62-170) reproduces the same
    Gaussian at new positions (a transposed axis, a reversed phase or a
    misplaced shift fails by tens of percent).  The agreement holds with the
    fitted positions in 1-based pixel-centre coordinates (x + 0.5 for a pixel
    x covering [x, x + 1)), the convention the device pipeline feeds it.
"""
import numpy as np

from bamf_amd import synth
from oracle import findpsf as fp

B = 8


def elliptic_rois(oracle, n, rng, sy=1.1, sx=1.6):
    fine = synth.sampled_gauss_psf(24, 32, sy, sx)
    X = rng.uniform(3.0, 5.0, n)
    Y = rng.uniform(3.0, 5.0, n)
    R = np.zeros((n, B, B))
    for k in range(n):
        R[k] = oracle.num_model([1000.0], [X[k]], [Y[k]], 0, 0, 0, B, B, fine).reshape(B, B).T
    return R, X, Y, fine


def test_shift_theorem(oracle):
    rng = np.random.default_rng(4)
    R, X, Y, _ = elliptic_rois(oracle, 1, rng)
    base = fp.shift_average(R[:1], X[:1], Y[:1])
    for axis, dx, dy in ((1, 1.0, 0.0), (0, 0.0, 1.0)):
        R2 = np.roll(R[0], 1, axis=axis)
        two = fp.shift_average(np.stack([R[0], R2]), [X[0], X[0] + dx], [Y[0], Y[0] + dy])
        assert np.abs(two - 2 * base).max() < 1e-9 * np.abs(base).max(), axis
        wrong = fp.shift_average(np.stack([R[0], R2]), [X[0], X[0] - dx], [Y[0], Y[0] - dy])
        assert np.abs(wrong - 2 * base).max() > 0.1 * np.abs(base).max(), axis


def test_phase_split_is_a_normalised_permutation():
    rng = np.random.default_rng(5)
    img = rng.uniform(0.1, 1.0, (B + 24, B + 24))  # distinct values
    p = np.pad(img, fp.PAD_R)
    bare = fp.phase_images(img, moves=False, normalise=False)
    # the sixteen images hold every pixel of the padded image exactly once
    assert np.array_equal(np.sort(bare.ravel()), np.sort(p.ravel()))
    # and each is the sub-sampled grid of one phase pair: rows, columns 4 apart
    for jj in range(4):
        for ii in range(4):
            t = bare[:, :, jj, ii]
            r = np.nonzero(np.isin(p, t[t > 0]).any(axis=1))[0]
            c = np.nonzero(np.isin(p, t[t > 0]).any(axis=0))[0]
            assert np.all(np.diff(r) == 4) and np.all(np.diff(c) == 4), (jj, ii)
    # the moves: the middle third (rows / columns 21..43, 0-based) shifted by one for phases > 0
    moved = fp.phase_images(img, normalise=False)
    for jj in range(4):
        for ii in range(4):
            t = bare[:, :, jj, ii].copy()
            if jj > 0:
                t[21:44, :] = t[20:43, :].copy()
            if ii > 0:
                t[:, 21:44] = t[:, 20:43].copy()
            assert np.array_equal(moved[:, :, jj, ii], t), (jj, ii)
    out = fp.phase_images(img)
    assert np.allclose(out.sum(axis=(0, 1)), 1.0)
    assert np.allclose(out * moved.sum(axis=(0, 1)), moved)


def test_round_trip_through_numodel(oracle):
    rng = np.random.default_rng(6)
    R, X, Y, fine = elliptic_rois(oracle, 200, rng)
    errs = {}
    for name, dx in (("centre 1-based", 0.5), ("0-based", -0.5)):
        T = fp.find_psf(R, X + dx, Y + dx)
        assert T.shape == (64, 64, 4, 4) and np.allclose(T.sum(axis=(0, 1)), 1.0, atol=1e-5)
        e = 0.0
        for k in range(30):
            x0, y0 = rng.uniform(4, 12, 2)
            mn = oracle.num_model([1000.0], [x0], [y0], 0, 0, 0, 16, 16, T)
            mf = oracle.num_model([1000.0], [x0], [y0], 0, 0, 0, 16, 16, fine)
            e = max(e, float(np.abs(mn - mf).max() / mf.max()))
        errs[name] = e
    assert errs["centre 1-based"] < 0.08 and errs["0-based"] > 0.3, errs
    # the table transposed (x and y swapped) no longer reproduces the ellipse
    T = fp.find_psf(R, X + 0.5, Y + 0.5)
    Tt = np.asfortranarray(T.transpose(1, 0, 3, 2))
    x0, y0 = 8.3, 7.6
    mf = oracle.num_model([1000.0], [x0], [y0], 0, 0, 0, 16, 16, fine)
    mt = oracle.num_model([1000.0], [x0], [y0], 0, 0, 0, 16, 16, Tt)
    assert np.abs(mt - mf).max() / mf.max() > 0.1
