"""ESS estimator against closed forms: iid draws and AR(1) processes, whose
integrated autocorrelation time is (1 + phi) / (1 - phi)."""
import numpy as np
import pytest

from bamf_amd.ess import ess_batch, roi_posteriors


def ar1(phi, n, m, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=(m, n))
    x = np.empty_like(e)
    x[:, 0] = e[:, 0] / np.sqrt(1 - phi ** 2)
    for t in range(1, n):
        x[:, t] = phi * x[:, t - 1] + e[:, t]
    return x


@pytest.mark.parametrize("phi", [0.0, 0.5, 0.9])
def test_ar1(phi):
    n = 4000
    e = ess_batch(ar1(phi, n, 64, 1))
    want = n * (1 - phi) / (1 + phi)
    assert abs(np.median(e) / want - 1) < 0.15


def test_constant_trace_and_criterion():
    x = np.zeros((2, 100))
    x[1] = np.arange(100) % 2  # alternating: antithetic, capped
    e = ess_batch(x)
    assert e[0] == 100
    tn = np.stack([np.full(500, 3), np.repeat([2, 3], 250)])
    tp = ar1(0.1, 500, 2, 3) + 1000
    ess, ok = roi_posteriors(tn, tp, 200)
    assert ok[0] and not ok[1]


def test_fixed_ess_rate():
    from bamf_amd.ess import fixed_ess_rate

    # two ROIs sharing 1 s equally, burn-in half: ESS 200 needs no extension
    assert abs(fixed_ess_rate(np.array([200.0, 200.0]), np.array([0.5, 0.5]), 0.5) - 2.0) < 1e-12
    # ESS 100 doubles the post-burn-in half: 0.5 * (0.5 + 1) each
    assert abs(fixed_ess_rate(np.array([100.0, 100.0]), np.array([0.5, 0.5]), 0.5) - 2 / 1.5) < 1e-12
    # a stuck chain (ESS floored at 1) dominates
    r = fixed_ess_rate(np.array([200.0, 0.1]), np.array([0.5, 0.5]), 0.0)
    assert abs(r - 2 / (0.5 + 100.0)) < 1e-12
