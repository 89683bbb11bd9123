"""findPSF on the device (bamf_find_psf_host: 'Sigma' fits, the Fourier
shift-and-average of psf.hip, the table split) against the oracle restatement
of @RJ/findPSF.m on the same positions, the fits against the oracle's 'Sigma'
fit, and the round trip that gives the table its meaning: ROIs of an
elliptical Gaussian, through the whole device pipeline, give a table with
which NumModel reproduces that Gaussian."""
import numpy as np
import pytest

from bamf_amd import synth
from bamf_amd._lib import BamfError
from bamf_amd.psf import find_psf
from bamf_amd.rjmcmc import Plan, RoiSet, run_batch
from oracle import findpsf as fp
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params

pytestmark = pytest.mark.gpu

B = 8


def stack(oracle, n, seed, noise=True, sy=1.1, sx=1.6, photons=2000.0, bg=5.0):
    rng = np.random.default_rng(seed)
    fine = synth.sampled_gauss_psf(24, 32, sy, sx)
    X, Y = rng.uniform(3.0, 5.0, n), rng.uniform(3.0, 5.0, n)
    R = np.zeros((B, B, n), np.float32)
    for k in range(n):
        mu = oracle.num_model([photons], [X[k]], [Y[k]], bg, 0, 0, B, B, fine).reshape(B, B).T
        R[:, :, k] = rng.poisson(mu) if noise else mu
    return R, X, Y, fine


@pytest.mark.parametrize("n,gain,offset", [(120, 1.0, 0.0), (1100, 2.5, 100.0)])
def test_find_psf_matches_oracle(oracle, n, gain, offset):
    R, *_ = stack(oracle, n, seed=n)
    raw = R * gain + offset
    print(f"findPSF: {n} ROIs", flush=True)
    psf, fx, fy = find_psf(raw, gain=gain, offset=offset, psf_sigma=1.3, return_fits=True)
    assert psf.shape == (64, 64, 4, 4) and psf.flags.f_contiguous
    r = ((raw - offset) / gain).astype(np.float32)
    rois = np.ascontiguousarray(r.transpose(2, 0, 1))
    # fits: the library's 'Sigma' fit (oracle restatement), as 1-based pixel centres (+ 0.5)
    p, _, _ = oracle.mle_xynbs(np.ascontiguousarray(r.transpose(2, 1, 0)), 1.3, 20)
    dy, dx = np.abs(fy - (p[0] + 0.5)), np.abs(fx - (p[1] + 0.5))
    assert np.quantile(dy, 0.99) < 1e-3 and np.quantile(dx, 0.99) < 1e-3
    want = fp.find_psf(rois, fx, fy)
    assert np.abs(psf - want).max() < 2e-5 * want.max(), float(np.abs(psf - want).max() / want.max())
    assert np.allclose(psf.sum(axis=(0, 1)), 1.0, atol=1e-5)


def test_find_psf_round_trip_and_chain(oracle):
    R, X, Y, fine = stack(oracle, 300, seed=9, noise=False)
    psf = find_psf(R, psf_sigma=1.3)
    rng = np.random.default_rng(2)
    err = 0.0
    for _ in range(20):
        x0, y0 = rng.uniform(4, 12, 2)
        mn = oracle.num_model([1000.0], [x0], [y0], 0, 0, 0, 16, 16, psf)
        mf = oracle.num_model([1000.0], [x0], [y0], 0, 0, 0, 16, 16, fine)
        err = max(err, float(np.abs(mn - mf).max() / mf.max()))
    assert err < 0.08, err
    # the table drives the sampled-PSF chains
    p = params(nb=300, nt=300)
    rois = synth.make_rois(8, seed=12, sides=(16,), mean_n=2.0)
    g = run_batch(Plan(p, OFF_PDF, SIG_PDF, BACK_PDF, sampled_psf=psf), RoiSet.from_synth(rois), seed=3)
    assert np.all(g.status == 0)


def test_find_psf_box_size_checked(oracle):
    with pytest.raises(BamfError):
        find_psf(np.ones((10, 10, 40), np.float32), box=10)
