"""The tile pipeline of analyzeROIs.m / mcmc.m on the device
(bamf_amd.refine.analyze_tiles): two 40 x 40 frames with bright emitters, some
inside the tiles' overlap bands, cut into 16 px tiles with 3 px overlap;
RJMCMC, findMap, the moves-only refinement from the signal SMD and the
ClustInfo assembly.  Every emitter must come out exactly once, in frame
coordinates, in its own frame; the refinement must start where findMap put
the emitters; and ClustInfo must be what the host assembly makes of the
device outputs."""
import numpy as np
import pytest

from bamf_amd import _hip, refine, synth
from bamf_amd.tiles import Tiling

pytestmark = pytest.mark.gpu

DX = 10.0
EMITTERS = {  # frame -> (x, y) in 0-based frame coordinates; 15.5 / 17.2 / 29.4 sit in overlap bands
    1: [(5.3, 6.1), (15.5, 8.2), (24.7, 25.3), (8.4, 33.6), (36.0, 12.5)],
    2: [(17.2, 17.8), (29.4, 5.5), (4.6, 22.2), (33.3, 34.1)],
}


def frames():
    ys = xs = 40
    data = np.zeros((ys, xs, 2), np.float32)
    rng = np.random.default_rng(12)
    for f, em in EMITTERS.items():
        X = np.array([e[0] for e in em])
        Y = np.array([e[1] for e in em])
        mu = synth.roi_expected(np.full(len(em), 4000.0), X, Y, 10.0, xs, ys, 1.3).reshape((ys, xs), order="F")
        data[:, :, f - 1] = rng.poisson(mu)
    return data


def rjstruct():
    return {"PSF_Sigma": 1.3, "SRZoom": 4, "N_Burnin": 2000, "N_Trials": 2000, "N_BurninMCMC": 500,
            "N_TrialsMCMC": 1000, "P_Offset": synth.gamma_pdf_table(10, 1, DX, 1001),
            "P_EmPhotons": synth.gamma_pdf_table(8, 500, DX, 1001), "P_BgEmPhotons": np.zeros(1001, np.float32),
            "P_dP": DX, "IstdFg": 10.0, "XstdFg": 0.1, "P_Burnin": [0.4, 0.15, 0.15, 0.15, 0.15, 0, 0, 0],
            "P_Trials": [0.6, 0.1, 0.1, 0.1, 0.1, 0, 0, 0], "MinPhoton": 50.0, "Rho": 0.01, "BGstd": 1.0,
            "IsBackg": 0, "XDrift": [], "YDrift": []}


def test_analyze_tiles_finds_every_emitter_once():
    data = frames()
    t = Tiling(40, 40, 2, 16, 3).run(_hip.DeviceBuffer.from_array(np.asfortranarray(data).ravel(order="F")))
    r = rjstruct()
    print(f"analyze_tiles: {t.n} tiles", flush=True)
    ci = refine.analyze_tiles(t, r, seed=5, debug=True)
    dev = ci.pop("_device")
    assert ci["Bg"].size == t.n and np.all(np.abs(ci["Bg"] - 10.0) < 2.0)
    n_ref = int(np.sum(dev["mf"] != 0))
    assert ci["JumpMC"].size == n_ref and np.all((ci["JumpMC"] > 0) & (ci["JumpMC"] <= 1))
    # the refinement started from findMap's emitters and stayed with them
    for k in range(t.n):
        if dev["mf"][k]:
            d = np.abs(dev["found"][k]["X"] - dev["smd"][k]["X"]) + np.abs(dev["found"][k]["Y"] - dev["smd"][k]["Y"])
            assert d.size == dev["mf"][k] and np.all(d < 0.5), (k, d)
    # each emitter exactly once, in frame coordinates, in its frame
    matched = np.zeros(ci["X"].size, bool)
    for f, em in EMITTERS.items():
        for x, y in em:
            dist = np.hypot(ci["X"] - x, ci["Y"] - y)
            hits = np.flatnonzero((dist < 0.3) & (ci["FrameNum"] == f))
            assert hits.size == 1, (f, x, y, dist[ci["FrameNum"] == f].min() if np.any(ci["FrameNum"] == f) else None)
            matched[hits] = True
    assert matched.sum() >= 0.9 * ci["X"].size, (matched.sum(), ci["X"].size)
    assert np.all(ci["X_SE"][matched] < 0.1) and np.all(np.abs(ci["I"][matched] / 4000.0 - 1) < 0.15)
    assert np.all(ci["ROInum"] == 1) and ci["ROIcornerX"].size == ci["X"].size
    # ClustInfo is the host assembly of the device outputs
    again = refine.assemble_clust_info(dev, t.struct(), r)
    for k in refine.CLUST_FIELDS:
        assert np.array_equal(again[k], ci[k], equal_nan=True), k
