"""The @RJ/rjmcmc.m mirror (bamf_amd.rjmcmc.rjmcmc) against cRJMCMC_gpu, the
mex wrapper, called with the MCMC struct, E_active and defaults that
rjmcmc.m:40-102 build (restated here from the .m file, independently of the
mirror):
  - an ROI cut from a larger Data array: PImage and PBackg are SRZoom times
    Data's size with the ROI's block filled, equal to the mex's PGrid and
    PBack; Chain, ProposedChain and Stat equal;
  - SampledPSF with OptModel 1 reaches the chain (equal to the mex run with
    the table and MCMC.PSFsize = size(SampledPSF), different from the erf
    run); OptModel 0, an empty OptModel or a scalar SampledPSF leave the erf
    model (rjmcmc.m:82-85,98-101)."""
import numpy as np
import pytest

from bamf_amd import mexsim, synth
from bamf_amd.rjmcmc import rjmcmc
from test_gpu_mex import mcmc_struct

pytestmark = pytest.mark.gpu

DX = 10.0


def rjstruct():
    return {"PSF_Sigma": 1.3, "SRZoom": 10, "N_Trials": 500, "N_Burnin": 500,
            "P_Offset": synth.gamma_pdf_table(10, 1, DX, 1001), "P_EmPhotons": synth.gamma_pdf_table(5, 200, DX, 1001),
            "P_BgEmPhotons": np.zeros(1001, np.float32), "P_dP": DX, "IstdFg": 10.0, "XstdFg": 0.1,
            "P_Burnin": [0.4, 0.15, 0.15, 0.15, 0.15, 0, 0, 0], "P_Trials": [0.6, 0.1, 0.1, 0.1, 0.1, 0, 0, 0],
            "MinPhoton": 50.0, "Rho": 0.02, "BGstd": 1.0, "IsBackg": 0}


def mex_as_rjmcmc_m(data, r, ysize, xsize, psf=None, opt=0, seed=0):
    """cRJMCMC_gpu with the arguments rjmcmc.m:43-102 prepares."""
    from bamf_amd.rjmcmc import make_params

    p_off = np.asarray(r["P_Offset"])
    bg = np.float32(r["P_dP"] * (np.flatnonzero(p_off == p_off.max())[0]))  # find(max == P_Offset) - 1
    p = make_params(psf_sigma=r["PSF_Sigma"], n_burnin=r["N_Burnin"], n_trials=r["N_Trials"], p_burnin=r["P_Burnin"],
                    p_trials=r["P_Trials"], rho=r["Rho"], x_std_fg=r["XstdFg"], i_std_fg=r["IstdFg"],
                    bg_std=r["BGstd"], dx=r["P_dP"], grid_zoom=r["SRZoom"], icutoff=r["MinPhoton"],
                    is_backg=r["IsBackg"])
    assert p.i_std_bg == np.float32(r["IstdFg"]) / 2 and p.x_std_bg == np.float32(r["XstdFg"]) / 2
    assert p.split_std == 0.5 and p.bnd_out == 2 and np.float32(p.abbg_std) == np.float32(0.02)
    sampled = np.zeros((10, 10, 2, 2), np.float32) if psf is None else np.asarray(psf, np.float32)
    e_active = {"N": np.int32(0), "I": np.zeros(0, np.float32), "X": np.zeros(0, np.float32),
                "Y": np.zeros(0, np.float32), "BG": bg}
    roi = np.asarray(data, np.float32)[:ysize, :xsize]
    return mexsim.call("cRJMCMC_gpu", 5, np.asfortranarray(roi), e_active,
                       mcmc_struct(p, psf_size=sampled.shape), np.int32(xsize), np.int32(ysize),
                       np.asarray(r["P_Offset"], np.float32), np.asarray(r["P_EmPhotons"], np.float32),
                       np.asarray(r["P_BgEmPhotons"], np.float32), np.zeros((ysize, xsize), np.float32),
                       np.asfortranarray(sampled), np.int32(opt), seed)


def same_chain(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for k, v in x.items():
            w = y[k]
            if v is None:
                assert w is None
                continue
            assert np.array_equal(np.asarray(v, np.float64).ravel(), np.asarray(w, np.float64).ravel(),
                                  equal_nan=True), k


def roi_data(seed, ys=20, xs=18):
    rng = np.random.default_rng(seed)
    k = 4
    X, Y = rng.uniform(2, 10, k), rng.uniform(2, 12, k)
    I = rng.gamma(5, 200, k)
    mu = synth.roi_expected(I, X, Y, 10.0, xs, ys, 1.3).reshape(xs, ys).T  # [y, x]
    return rng.poisson(mu).astype(np.float32)


def test_mirror_equals_mex_on_a_cut_roi():
    data = roi_data(3)
    r = rjstruct()
    print("rjmcmc mirror: 14 x 12 cut of 20 x 18", flush=True)
    chain, stat, pimg, pbg, prop = rjmcmc(data, r, 14, 12, seed=77)
    PG, PB, CH, PR, ST = mex_as_rjmcmc_m(data, r, 14, 12, seed=77)
    z = r["SRZoom"]
    assert pimg.shape == (z * 20, z * 18) and pbg.shape == pimg.shape
    assert np.array_equal(pimg[:z * 14, :z * 12], PG) and np.array_equal(pbg[:z * 14, :z * 12], PB)
    assert not pimg[z * 14:].any() and not pimg[:, z * 12:].any()
    assert PG.sum() > 0
    same_chain(chain, CH)
    same_chain(prop, PR)
    assert len(chain) > 100
    for k, v in ST.items():
        assert np.array_equal(np.float64(stat[k]), v.ravel()[0], equal_nan=True), k
    # a scalar sCMOSVar is the default (zeros): the same run
    chain2, _, pimg2, _, _ = rjmcmc(data, r, 14, 12, scmos_var=0.0, seed=77)
    assert np.array_equal(pimg2, pimg)
    with pytest.raises(Exception, match="single ROI"):
        rjmcmc(np.zeros((8, 8, 2), np.float32), r, 8, 8)


def test_mirror_passes_sampled_psf():
    psf = synth.sampled_gauss_psf(40, 4, 1.2, 1.6)
    data = roi_data(5, 16, 16)
    r = rjstruct()
    print("rjmcmc mirror: SampledPSF 40 x 40 x 4 x 4", flush=True)
    chain, _, pimg, _, _ = rjmcmc(data, r, 16, 16, sampled_psf_in=psf, opt_model_in=1, seed=9)
    PG, _, CH, _, _ = mex_as_rjmcmc_m(data, r, 16, 16, psf=psf, opt=1, seed=9)
    assert np.array_equal(pimg, PG)
    same_chain(chain, CH)
    erf_run = rjmcmc(data, r, 16, 16, seed=9)[2]
    assert not np.array_equal(erf_run, pimg)
    # OptModel 0, an empty OptModel or a scalar table: the erf model
    for kw in ({"sampled_psf_in": psf, "opt_model_in": 0}, {"sampled_psf_in": psf, "opt_model_in": []},
               {"sampled_psf_in": 0.0, "opt_model_in": 1}):
        assert np.array_equal(rjmcmc(data, r, 16, 16, seed=9, **kw)[2], erf_run), kw
    PG0 = mex_as_rjmcmc_m(data, r, 16, 16, psf=psf, opt=0, seed=9)[0]
    assert np.array_equal(PG0, erf_run)
