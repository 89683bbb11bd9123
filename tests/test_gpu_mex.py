"""The mex wrappers (bamf_amd/csrc/mex), built against the project's mx
runtime, run end to end on the device with MATLAB's argument layout and
compared with the C-ABI host gateways they sit on:
  - cRJMCMC_gpu (cRJMCMC.cpp:52-305's arguments and outputs) against
    bamf_crjmcmc on the same ROI and seed: PGrid, PBack, every Chain and
    ProposedChain entry, Stat;
  - findROI_gpu against the host mirror (find_roi): SMD, ROIStack, Res;
  - gaussMLE_gpu ('Basic' and 'Sigma') against the device fit entry;
  - findPriors_gpu against find_priors; findPSF_gpu against find_psf."""
import numpy as np
import pytest

from bamf_amd import mexsim, synth
from bamf_amd._lib import BamfError
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params
from test_gpu_gateway import crjmcmc

pytestmark = pytest.mark.gpu


def mcmc_struct(p, psf_size=None):
    s = {"PSFsigma": p.psf_sigma, "N_Trials": p.n_trials, "N_Burnin": p.n_burnin, "I_stdFg": p.i_std_fg,
         "I_stdBg": p.i_std_bg, "X_stdFg": p.x_std_fg, "X_stdBg": p.x_std_bg,
         "P_Burnin": np.array(list(p.p_burnin), np.float32), "P_Trials": np.array(list(p.p_trials), np.float32),
         "Rho": p.rho, "Bnd_out": p.bnd_out, "Grid_Zoom": p.grid_zoom, "Split_std": p.split_std,
         "Icutoff": p.icutoff, "BG_std": p.bg_std, "ABBG_std": p.abbg_std, "DX": p.dx, "DriftX": p.drift_x,
         "DriftY": p.drift_y, "IsBackg": p.is_backg}
    if psf_size is not None:
        s["PSFsize"] = np.asarray(psf_size, np.int32)
    return s


@pytest.mark.parametrize("side,n0", [(12, 0), (16, 2)])
def test_crjmcmc_mex_matches_gateway(side, n0):
    rois = synth.make_rois(1, seed=40 + side, sides=(side,), mean_n=3.0)
    p = params(nb=600, nt=600)
    data = rois["data"].astype(np.float32)
    i0 = np.array([900.0, 600.0][:n0], np.float32)
    x0 = np.array([4.4, 9.1][:n0], np.float32)
    y0 = np.array([6.2, 3.3][:n0], np.float32)
    bg0, seed = 11.0, 4321
    print(f"cRJMCMC_gpu {side} x {side}, {n0} initial emitters", flush=True)
    rc, pg, pb, acc, prop, ae, pe, counts, stat = crjmcmc(p, data, side, side, bg0, seed, n0=n0,
                                                          i0=i0 if n0 else None, x0=x0 if n0 else None,
                                                          y0=y0 if n0 else None, rec_cap=1201, em_cap=1201 * 16 + 1024)
    assert rc == 0
    e_active = {"N": float(n0), "I": i0, "X": x0, "Y": y0, "BG": bg0}
    out = mexsim.call("cRJMCMC_gpu", 5, data.reshape((side, side), order="F"), e_active, mcmc_struct(p), side, side,
                      OFF_PDF, SIG_PDF, BACK_PDF, np.zeros((0, 0), np.float32), np.zeros(1, np.float32), 0, seed)
    PGrid, PBack, Chain, Proposed, Stat = out
    z = p.grid_zoom
    assert PGrid.dtype == np.float32 and PGrid.shape == (side * z, side * z)
    assert np.array_equal(PGrid.ravel(order="F").view(np.uint32), pg.view(np.uint32))
    assert np.array_equal(PBack.ravel(order="F").view(np.uint32), pb.view(np.uint32))
    nrec = int(counts[0])
    assert len(Chain) == len(Proposed) == nrec > 100
    for k in range(nrec - 1):
        for ch, rec, em in ((Chain[k], acc[k], ae), (Proposed[k], prop[k], pe)):
            nn = max(int(rec["n"]), 1)
            e = em[rec["em_index"]:rec["em_index"] + nn]
            assert ch["N"][0, 0] == (0.0 if e[0, 2] == 0 else rec["n"]), k
            for f, c in (("X", 0), ("Y", 1), ("Photons", 2), ("Signal", 3)):
                assert np.array_equal(ch[f].ravel(), e[:, c].astype(np.float64)), (k, f)
            for f in ("bg", "abg", "bbg"):
                assert ch[f.upper()][0, 0] == np.float64(rec[f])
        # LLR, PR, JumpType and Accepted of both outputs are the accepted chain's
        for ch in (Chain[k], Proposed[k]):
            assert ch["LLR"][0, 0] == np.float64(acc[k]["llr"]) or (np.isnan(ch["LLR"][0, 0]) and
                                                                    np.isnan(acc[k]["llr"]))
            assert ch["JumpType"][0, 0] == acc[k]["jump_type"] and ch["Accepted"][0, 0] == acc[k]["accepted"]
    # the last entry is allocated and left empty, as cRJMCMC.cpp:168-229 leaves it
    assert Chain[-1]["N"] is None
    names = ["NTrials", "Accept_Jump", "Accept_Split", "Accept_Merge", "Accept_Birth", "Accept_Death",
             "Accept_JumpFg", "Accept_SplitFg", "Accept_MergeFg", "Accept_BirthFg", "Accept_DeathFg",
             "Accept_JumpBg", "Accept_SplitBg", "Accept_MergeBg", "Accept_BirthBg", "Accept_DeathBg"]
    got = np.array([Stat[n][0, 0] for n in names])
    assert np.array_equal(np.isnan(got), np.isnan(stat)) and np.array_equal(got[~np.isnan(got)],
                                                                            stat[~np.isnan(stat)])
    # an ROI beyond 32 px is refused with the library's message
    with pytest.raises(BamfError, match="1..32"):
        mexsim.call("cRJMCMC_gpu", 5, np.zeros((33, 33), np.float32), {"N": 0.0, "I": i0[:0], "X": x0[:0],
                                                                      "Y": y0[:0], "BG": 1.0},
                    mcmc_struct(p), 33, 33, OFF_PDF, SIG_PDF, BACK_PDF, np.zeros((0, 0), np.float32),
                    np.zeros(1, np.float32), 0)


def test_find_roi_mex_matches_host_mirror():
    from bamf_amd.find_roi import default_minval, find_roi

    frames = synth.make_frames(5, 64, seed=29, density_um2=1.0)
    img = np.asfortranarray(np.transpose(frames, (1, 2, 0)).astype(np.float32))  # (Y, X, Z)
    print("findROI_gpu 5 x 64 x 64", flush=True)
    smd, stack, res = mexsim.call("findROI_gpu", 3, img, 1.3, 2.6, 7.0, float(default_minval()))
    smd_h, stack_h, res_h = find_roi(img)
    n = len(smd_h["XBoxCorner"])
    assert n > 3 and smd["XBoxCorner"].shape == (n, 1)
    for f in ("XBoxCorner", "YBoxCorner", "FrameNum"):
        assert np.array_equal(smd[f].ravel(), np.asarray(smd_h[f], np.float64).ravel()), f
    assert stack.dtype == np.float64 and stack.shape == (7, 7, n)
    assert np.array_equal(stack, np.asarray(stack_h, np.float64))
    assert np.array_equal(res, np.asarray(res_h, np.float64))
    # the defaults of findROI.m:60-75 when only the image is given
    smd_d = mexsim.call("findROI_gpu", 1, img)
    assert np.array_equal(smd_d["XBoxCorner"], smd["XBoxCorner"])


@pytest.mark.parametrize("fit", ["Basic", "Sigma"])
def test_gauss_mle_mex_matches_device_entry(fit):
    from bamf_amd.gauss_mle import fit_boxes

    boxes = np.ascontiguousarray(synth.make_mle_boxes(2000, 7, seed=31)[0], np.float32)  # (n, j, i)
    n = boxes.shape[0]
    stack = np.asfortranarray(boxes.transpose(2, 1, 0))  # MATLAB (SZ, SZ, N)
    print(f"gaussMLE_gpu {fit}: {n} boxes", flush=True)
    res, st = mexsim.call("gaussMLE_gpu", 2, stack, fit, "CCD", 1.3)
    p, c, ll = fit_boxes(boxes, 1.3, 20, fit)
    names = ["Y", "X", "Photons", "BG"] + (["Sigma"] if fit == "Sigma" else [])
    for k, nm in enumerate(names):
        assert np.array_equal(res[nm].ravel().view(np.uint32), p[k].view(np.uint32)), nm
        assert np.array_equal(res[nm + "_SE"].ravel(), np.sqrt(c[k])), nm
    assert np.array_equal(res["LogLikelihood"].ravel().view(np.uint32), ll.view(np.uint32))
    assert st["NFitsPerCall"][0, 0] == n and st["FitTime"][0, 0] > 0


def test_find_priors_mex_matches_host_mirror():
    from bamf_amd.priors import find_priors

    frames = synth.make_frames(20, 96, seed=33, density_um2=0.6, photons=(8.0, 250.0))
    img = np.asfortranarray(np.transpose(frames, (1, 2, 0)).astype(np.float32))
    sma = {"MeanPhotons": 1500.0, "MinPValue": 0.001, "MinPhotons": 50.0}
    print("findPriors_gpu 20 x 96 x 96", flush=True)
    sig, bg, off = mexsim.call("findPriors_gpu", 3, img, 1.3, sma, 5.0, 5.0, "Yes")
    want = find_priors(img, 1.3, sma, 5.0, 5.0, "Yes")
    assert sig.shape == (1, len(want["P_EmPhotons"])) and sig.size > 10
    assert np.array_equal(sig.ravel(), want["P_EmPhotons"])
    assert np.array_equal(bg.ravel(), want["P_BgEmPhotons"])
    assert np.array_equal(off.ravel(), want["P_Offset"])


def test_find_psf_mex_matches_host_mirror():
    from bamf_amd.psf import find_psf

    rng = np.random.default_rng(35)
    n, B = 200, 8
    R = np.zeros((B, B, n), np.float32)
    for k in range(n):
        X, Y = rng.uniform(3, 5, 2)
        mu = synth.roi_expected(np.array([1500.0]), np.array([X]), np.array([Y]), 4.0, B, B, 1.3).reshape(B, B)
        R[:, :, k] = rng.poisson(mu)
    raw = np.asfortranarray(R * 2.0 + 50.0)
    print("findPSF_gpu 200 ROIs", flush=True)
    psf = mexsim.call("findPSF_gpu", 1, raw, 2.0, 50.0, 1.3, 8.0)
    want = find_psf(raw, gain=2.0, offset=50.0, psf_sigma=1.3)
    assert psf.shape == (64, 64, 4, 4)
    assert np.array_equal(psf.view(np.uint32), np.asarray(want).view(np.uint32))
