"""Pins the oracle's gaussMLE restatement (oracle/mle.c) without the reference.

Checks are chosen so that a plausible slip (dropped term, wrong sign,
transposed axes) fails one of them:
  - the 1-D pixel integral against a float64 erf formula;
  - centre of mass and the smoothing filter against closed forms;
  - the LU partial inverse against numpy.linalg.inv;
  - noiseless boxes: the fit returns the generating (x, y, I, bg);
  - CRLB against the inverse of a Fisher matrix built from finite differences
    of the model (independent of the analytic derivatives);
  - noisy boxes: the fit equals the Poisson maximum-likelihood point found by
    scipy.optimize in float64.
"""
import numpy as np
import pytest
from scipy.optimize import minimize
from scipy.special import erf

from bamf_amd import synth

SIG = 1.3


def model64(theta, sz, sigma=SIG):
    x, y, n, bg = theta
    e = np.arange(sz + 1, dtype=np.float64)
    c = 1 / (sigma * np.sqrt(2))
    gx = 0.5 * np.diff(erf((e - x) * c))
    gy = 0.5 * np.diff(erf((e - y) * c))
    return bg + n * np.outer(gy, gx)  # [jj, ii]


def test_int_gauss_1d(oracle):
    L = oracle.lib()
    for x in [0.3, 2.5, 3.7, 6.1]:
        for ii in range(7):
            ref = 0.5 * (erf((ii - x + 1) / (SIG * np.sqrt(2))) - erf((ii - x) / (SIG * np.sqrt(2))))
            assert abs(L.orc_int_gauss_1d(ii, x, SIG) - ref) < 2e-7


def test_center_of_mass(oracle):
    import ctypes
    L = oracle.lib()
    box = np.zeros((7, 7), np.float32)
    box[2, 5] = 10.0  # data[sz*jj + ii] with jj=2, ii=5
    x, y = ctypes.c_float(), ctypes.c_float()
    L.orc_center_of_mass_2d(7, box.ctypes.data, ctypes.byref(x), ctypes.byref(y))
    assert (x.value, y.value) == (5.0, 2.0)
    box[:] = 3.0
    L.orc_center_of_mass_2d(7, box.ctypes.data, ctypes.byref(x), ctypes.byref(y))
    assert abs(x.value - 3.0) < 1e-6 and abs(y.value - 3.0) < 1e-6


def test_fmaxmin_filter(oracle):
    import ctypes
    L = oracle.lib()
    mx, mn = ctypes.c_float(), ctypes.c_float()
    box = np.full((7, 7), 4.25, np.float32)  # a normalised filter keeps a constant
    L.orc_gauss_fmaxmin_2d(7, SIG, box.ctypes.data, ctypes.byref(mx), ctypes.byref(mn))
    assert abs(mx.value - 4.25) < 1e-5 and abs(mn.value - 4.25) < 1e-5
    rng = np.random.default_rng(1)
    box = rng.uniform(0, 100, (7, 7)).astype(np.float32)
    L.orc_gauss_fmaxmin_2d(7, SIG, box.ctypes.data, ctypes.byref(mx), ctypes.byref(mn))
    # independent: normalised 2-D Gaussian weights as a 4-D tensor contraction in float64
    i = np.arange(7)
    g = np.exp(-((i[:, None] - i[None, :]) ** 2) / (2 * SIG * SIG))
    f = (g @ box.astype(np.float64) @ g.T) / (g.sum(1)[:, None] * g.sum(1)[None, :])
    assert abs(mx.value - f.max()) < 1e-4 * f.max()
    assert abs(mn.value - f.min()) < 1e-4 * f.max()


def test_mat_inv_n(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    for _ in range(20):
        a = rng.normal(size=(4, 4))
        m = (a @ a.T + 4 * np.eye(4)).astype(np.float32)
        minv = np.zeros(16, np.float32)
        diag = np.zeros(4, np.float32)
        mm = m.copy()
        L.orc_mat_inv_n(mm.ctypes.data, minv.ctypes.data, diag.ctypes.data, 4)
        ref = np.linalg.inv(m.astype(np.float64))
        np.testing.assert_allclose(diag, np.diag(ref), rtol=1e-4)
        np.testing.assert_allclose(minv.reshape(4, 4).T, ref, rtol=1e-3, atol=1e-6)


def test_noiseless_converges_to_truth(oracle):
    boxes, t = synth.make_mle_boxes(64, 7, seed=3, noise=False)
    p, c, ll = oracle.mle_xynb(boxes, SIG, 20)
    np.testing.assert_allclose(p[0], t["x"], atol=2e-3)
    np.testing.assert_allclose(p[1], t["y"], atol=2e-3)
    np.testing.assert_allclose(p[2], t["I"], rtol=2e-3)
    np.testing.assert_allclose(p[3], t["bg"], rtol=2e-3)
    # log-likelihood ratio to the saturated model is ~0 when model == data
    assert np.all(np.abs(ll) < 1e-3 * boxes.reshape(64, -1).sum(1))


def test_crlb_against_finite_difference_fisher(oracle):
    boxes, t = synth.make_mle_boxes(16, 7, seed=4, noise=False)
    p, c, _ = oracle.mle_xynb(boxes, SIG, 20)
    for i in range(16):
        th = np.array([t["x"][i], t["y"][i], t["I"][i], t["bg"][i]])
        mu = model64(th, 7)
        J = np.zeros((49, 4))
        for k in range(4):
            h = 1e-5 * max(1.0, abs(th[k]))
            tp, tm = th.copy(), th.copy()
            tp[k] += h
            tm[k] -= h
            J[:, k] = ((model64(tp, 7) - model64(tm, 7)) / (2 * h)).ravel()
        F = J.T @ (J / mu.ravel()[:, None])
        np.testing.assert_allclose(c[:, i], np.diag(np.linalg.inv(F)), rtol=5e-3)


@pytest.mark.parametrize("seed", [5, 6])
def test_noisy_fit_is_poisson_mle(oracle, seed):
    boxes, t = synth.make_mle_boxes(12, 7, seed=seed, noise=True)
    p, _, _ = oracle.mle_xynb(boxes, SIG, 20)
    for i in range(12):
        d = boxes[i].astype(np.float64)

        def nll(th):
            mu = model64(th, 7)
            if np.any(mu <= 0):
                return 1e30
            return float(np.sum(mu - d * np.log(mu)))

        th0 = np.array([t["x"][i], t["y"][i], t["I"][i], t["bg"][i]])
        r = minimize(nll, th0, method="Nelder-Mead", options={"xatol": 1e-7, "fatol": 1e-10, "maxiter": 20000})
        assert abs(p[0, i] - r.x[0]) < 2e-3 and abs(p[1, i] - r.x[1]) < 2e-3
        assert abs(p[2, i] - r.x[2]) < 2e-3 * r.x[2]
        assert abs(p[3, i] - r.x[3]) < 5e-3 * max(r.x[3], 1.0)
