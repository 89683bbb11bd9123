"""GPU gaussMLE (bamf_gauss_mle_xynb) against the oracle: (x, y, I, bg) within
1e-4 relative (BASELINE tolerance), CRLB and log-likelihood within 1e-3."""
import numpy as np
import pytest

from bamf_amd import synth
from bamf_amd.gauss_mle import fit_boxes, gauss_mle

pytestmark = pytest.mark.gpu


def rel(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


@pytest.mark.parametrize("sz,n,seed", [(7, 3000, 11), (5, 500, 12), (9, 500, 13), (11, 200, 14)])
def test_mle_matches_oracle(oracle, sz, n, seed):
    boxes, _ = synth.make_mle_boxes(n, sz, seed=seed)
    pg, cg, lg = fit_boxes(boxes, 1.3, 20)
    po, co, lo = oracle.mle_xynb(boxes, 1.3, 20)
    print(f"sz={sz} n={n} seed={seed}", flush=True)
    for k, floor in zip(range(4), (1.0, 1.0, 1.0, 1.0)):
        r = rel(pg[k], po[k], floor)
        assert np.quantile(r, 0.999) < 1e-4 and r.max() < 1e-3, (k, r.max())
    assert np.quantile(rel(cg, co, 1e-6), 0.999) < 1e-3
    assert np.quantile(rel(lg, lo, 1.0), 0.999) < 1e-3


def test_mle_edge_cases(oracle):
    """Arrays are parameter-major: p[k, i] is parameter k of box i.
    Box 0 is all zeros (0/0 centre of mass), box 1 is constant (no emitter, so
    the position is unidentifiable: only bg and a floor-level I are compared),
    box 2 is a very bright emitter, boxes 3.. are ordinary; 67 boxes is not a
    multiple of the work-group size."""
    boxes = np.zeros((67, 7, 7), np.float32)
    boxes[1] = 5.0
    bright, _ = synth.make_mle_boxes(1, 7, seed=70, photons=(5.0, 2e5), noise=False)
    boxes[2] = bright[0]
    boxes[3:], _ = synth.make_mle_boxes(64, 7, seed=7)
    pg, cg, lg = fit_boxes(boxes, 1.3, 20)
    po, co, lo = oracle.mle_xynb(boxes, 1.3, 20)
    assert np.array_equal(np.isfinite(po), np.isfinite(pg))
    assert np.array_equal(np.isnan(pg[:, 0]), np.isnan(po[:, 0]))
    assert abs(pg[3, 1] - po[3, 1]) < 1e-3 and abs(po[3, 1] - 5.0) < 0.05
    assert pg[2, 1] < 2.0 and po[2, 1] < 2.0
    assert np.all(rel(pg[:, 2], po[:, 2], 1.0) < 1e-4)
    assert np.all(rel(pg[:, 3:], po[:, 3:], 1.0) < 1e-4)


def test_gauss_mle_matlab_mirror():
    boxes, t = synth.make_mle_boxes(256, 7, seed=8, noise=False)
    stack = np.asfortranarray(np.transpose(boxes, (2, 1, 0)))  # MATLAB (SZ, SZ, N)
    res, stats = gauss_mle(stack, "Basic", "CCD", 1.3)
    # gaussMLE.m:183-186: column 1 -> Y; the kernel's x runs along the box's first MATLAB index
    np.testing.assert_allclose(res["Y"], t["x"], atol=2e-3)
    np.testing.assert_allclose(res["X"], t["y"], atol=2e-3)
    np.testing.assert_allclose(res["Photons"], t["I"], rtol=2e-3)
    assert stats["NFitsPerCall"] == 256


@pytest.mark.parametrize("sz,n,seed,true_sigma", [(7, 2000, 31, 1.3), (9, 500, 32, 1.6), (11, 200, 33, 1.1)])
def test_mle_sigma_matches_oracle(oracle, sz, n, seed, true_sigma):
    """bamf_gauss_mle_xynbs (kernel_MLEFit_XYNBS_) against the oracle: the
    same bar as the 'Basic' fit for (x, y, I, bg, sigma)."""
    boxes, _ = synth.make_mle_boxes(n, sz, seed=seed, sigma=true_sigma)
    pg, cg, lg = fit_boxes(boxes, 1.3, 20, "Sigma")
    po, co, lo = oracle.mle_xynbs(boxes, 1.3, 20)
    print(f"sigma fit sz={sz} n={n} seed={seed}", flush=True)
    for k in range(5):
        r = rel(pg[k], po[k], 1.0)
        assert np.quantile(r, 0.999) < 1e-4 and r.max() < 1e-3, (k, np.quantile(r, 0.999), r.max())
    assert np.quantile(rel(cg, co, 1e-6), 0.999) < 1e-3
    assert np.quantile(rel(lg, lo, 1.0), 0.999) < 1e-3
    assert np.all(pg[4] >= 0.5) and np.all(pg[4] <= sz / 2)


def test_gauss_mle_sigma_mirror():
    boxes, t = synth.make_mle_boxes(400, 9, seed=34, noise=False, sigma=1.5, photons=(20.0, 100.0))
    res, stats = gauss_mle(boxes, "Sigma", "CCD", 1.3, iterations=100)
    np.testing.assert_allclose(res["Sigma"], 1.5, rtol=2e-3)
    np.testing.assert_allclose(res["Photons"], t["I"], rtol=3e-3)
    assert res["Sigma_SE"].shape == (400,) and np.all(np.isfinite(res["Sigma_SE"]))
    assert stats["NFitsPerCall"] == 400


def elliptic_boxes(n, sz, seed, sx, sy):
    from scipy.special import erf
    rng = np.random.default_rng(seed)
    x = rng.uniform(sz / 2 - 1, sz / 2 + 1, n) - 0.5
    y = rng.uniform(sz / 2 - 1, sz / 2 + 1, n) - 0.5
    I = rng.gamma(5.0, 200.0, n)
    b = rng.uniform(5, 30, n)
    e = np.arange(sz + 1, dtype=np.float64)
    gx = 0.5 * np.diff(erf((e[None, :] - x[:, None]) / (sx * np.sqrt(2))), axis=1)
    gy = 0.5 * np.diff(erf((e[None, :] - y[:, None]) / (sy * np.sqrt(2))), axis=1)
    mu = b[:, None, None] + I[:, None, None] * gy[:, :, None] * gx[:, None, :]
    return rng.poisson(mu).astype(np.float32)


@pytest.mark.parametrize("sz,n,seed", [(7, 2000, 35), (9, 500, 36), (11, 200, 37)])
def test_mle_sigmayx_matches_oracle(oracle, sz, n, seed):
    """bamf_gauss_mle_xynbsxsy (kernel_MLEFit_XYNBSXSY_) against the oracle."""
    boxes = elliptic_boxes(n, sz, seed, 1.2, 1.6)
    pg, cg, lg = fit_boxes(boxes, 1.3, 20, "SigmaYX")
    po, co, lo = oracle.mle_xynbsxsy(boxes, 1.3, 20)
    print(f"sigma-yx fit sz={sz} n={n} seed={seed}", flush=True)
    for k in range(6):
        r = rel(pg[k], po[k], 1.0)
        assert np.quantile(r, 0.999) < 1e-4 and r.max() < 1e-3, (k, np.quantile(r, 0.999), r.max())
    # a fit whose emitter leaves the box with both widths clamped has a singular
    # Fisher matrix: NaN in the reference's arithmetic, so non-finite on both sides
    fin = np.isfinite(co).all(0)
    assert np.array_equal(fin, np.isfinite(cg).all(0)) and fin.mean() > 0.99
    assert np.quantile(rel(cg[:, fin], co[:, fin], 1e-6), 0.999) < 1e-3
    assert np.quantile(rel(lg, lo, 1.0), 0.999) < 1e-3
    res, _ = gauss_mle(boxes, "SigmaYX", "CCD", 1.3)
    assert np.array_equal(res["SigmaY"], pg[4]) and np.array_equal(res["SigmaX"], pg[5])


ZSTRUCT = {"Ax": 0.05, "Ay": -0.05, "Bx": 0.02, "By": 0.03, "Gamma": 0.35, "D": 0.6}
ZSIG = (1.1, 1.25)


def z_boxes(n, sz, seed):
    from scipy.special import erf
    rng = np.random.default_rng(seed)
    x = rng.uniform(sz / 2 - 1, sz / 2 + 1, n) - 0.5
    y = rng.uniform(sz / 2 - 1, sz / 2 + 1, n) - 0.5
    I = rng.gamma(5.0, 300.0, n)
    b = rng.uniform(5, 20, n)
    z = rng.uniform(-0.3, 0.3, n)
    d = ZSTRUCT["D"]

    def alpha(t, a, bb):
        u = t / d
        return 1 + u ** 2 + a * u ** 3 + bb * u ** 4
    sx = ZSIG[0] * np.sqrt(alpha(z - ZSTRUCT["Gamma"], ZSTRUCT["Ax"], ZSTRUCT["Bx"]))
    sy = ZSIG[1] * np.sqrt(alpha(z + ZSTRUCT["Gamma"], ZSTRUCT["Ay"], ZSTRUCT["By"]))
    e = np.arange(sz + 1, dtype=np.float64)
    gx = 0.5 * np.diff(erf((e[None, :] - x[:, None]) / (sx[:, None] * np.sqrt(2))), axis=1)
    gy = 0.5 * np.diff(erf((e[None, :] - y[:, None]) / (sy[:, None] * np.sqrt(2))), axis=1)
    mu = b[:, None, None] + I[:, None, None] * gy[:, :, None] * gx[:, None, :]
    return rng.poisson(mu).astype(np.float32)


@pytest.mark.parametrize("sz,n,seed", [(7, 2000, 38), (9, 500, 39), (11, 200, 40)])
def test_mle_z_matches_oracle(oracle, sz, n, seed):
    """bamf_gauss_mle_xynbz (kernel_MLEFit_XYNBZ_) against the oracle."""
    boxes = z_boxes(n, sz, seed)
    pg, cg, lg = fit_boxes(boxes, ZSIG, 20, "Z", ZSTRUCT)
    cal = (ZSIG[0], ZSTRUCT["Ax"], ZSTRUCT["Ay"], ZSTRUCT["Bx"], ZSTRUCT["By"], ZSTRUCT["Gamma"], ZSTRUCT["D"],
           ZSIG[1])
    po, co, lo = oracle.mle_xynbz(boxes, cal, 20)
    print(f"z fit sz={sz} n={n} seed={seed}", flush=True)
    for k in range(5):
        r = rel(pg[k], po[k], 1.0)
        assert np.quantile(r, 0.999) < 1e-4 and r.max() < 1e-3, (k, np.quantile(r, 0.999), r.max())
    fin = np.isfinite(co).all(0)
    assert np.array_equal(fin, np.isfinite(cg).all(0)) and fin.mean() > 0.99
    assert np.quantile(rel(cg[:, fin], co[:, fin], 1e-6), 0.999) < 1e-3
    assert np.quantile(rel(lg, lo, 1.0), 0.999) < 1e-3
    res, _ = gauss_mle(boxes, "Z", "CCD", ZSIG, ZSTRUCT)
    assert np.array_equal(res["Z"], pg[4]) and res["Z_SE"].shape == (n,)


def scmos_setup(n, sz, seed, mapsz=48):
    rng = np.random.default_rng(seed)
    gain = rng.uniform(0.5, 40.0, size=(mapsz, mapsz)).astype(np.float32)
    coords = np.stack([rng.integers(0, mapsz - sz, n), rng.integers(0, mapsz - sz, n)], 1).astype(np.float32)
    boxes, _ = synth.make_mle_boxes(n, sz, seed=seed, noise=False)
    for i in range(n):
        x0, y0 = int(coords[i, 0]), int(coords[i, 1])
        g = gain[y0:y0 + sz, x0:x0 + sz]
        boxes[i] = (rng.poisson(boxes[i] + g) - g).astype(np.float32)
    return boxes, coords, gain


@pytest.mark.parametrize("fit_type,ft", [("Basic", 0), ("Sigma", 1), ("SigmaYX", 2), ("Z", 3)])
@pytest.mark.parametrize("sz,n,seed", [(7, 1500, 44), (11, 150, 45)])
def test_mle_scmos_matches_oracle(oracle, fit_type, ft, sz, n, seed):
    """The sCMOS kernels (kernel_MLEFit_SCMOS*) against the oracle, every fit type,
    a specialised and the generic box size."""
    boxes, coords, gain = scmos_setup(n, sz, seed)
    z0 = np.linspace(-0.1, 0.1, n).astype(np.float32)
    sig = ZSIG if fit_type == "Z" else 1.3
    pg, cg, lg = fit_boxes(boxes, sig, 20, fit_type, ZSTRUCT if fit_type == "Z" else None, coords, gain, z0)
    cal = (ZSIG[0], ZSTRUCT["Ax"], ZSTRUCT["Ay"], ZSTRUCT["Bx"], ZSTRUCT["By"], ZSTRUCT["Gamma"], ZSTRUCT["D"],
           ZSIG[1])
    po, co, lo = oracle.mle_scmos(ft, boxes, coords, gain, ZSIG[0] if fit_type == "Z" else 1.3, 20,
                                  zcal=cal if fit_type == "Z" else None, z0=z0)
    print(f"scmos {fit_type} sz={sz} n={n}", flush=True)
    for k in range(pg.shape[0]):
        r = rel(pg[k], po[k], 1.0)
        assert np.quantile(r, 0.999) < 1e-4 and r.max() < 1e-3, (k, np.quantile(r, 0.999), r.max())
    # near-singular Fisher matrices (an emitter leaving the box, clamped widths),
    # inverted without pivoting as the reference does, turn last-ulp differences
    # into inf on one side and a huge finite value on the other: allow a few
    fo, fg = np.isfinite(co).all(0), np.isfinite(cg).all(0)
    assert (fo != fg).mean() <= 2e-3 and fo.mean() > 0.99
    fin = fo & fg
    assert np.quantile(rel(cg[:, fin], co[:, fin], 1e-6), 0.999) < 1e-3
    assert np.quantile(rel(lg, lo, 1.0), 0.999) < 1e-3
