"""Shared setup for the chain tests: config-1 style parameters and priors."""
import numpy as np

from bamf_amd import synth
from bamf_amd.rjmcmc import make_params

DX = 10.0
SIG_PDF = synth.gamma_pdf_table(5, 200, DX, 1001)     # Gamma(k=5, theta=200) photons
OFF_PDF = synth.gamma_pdf_table(10, 1, DX, 1001)      # offset prior, mean 10
BACK_PDF = np.zeros(1001, np.float32)


def params(nb=3000, nt=3000, **kw):
    return make_params(n_burnin=nb, n_trials=nt, **kw)


def oracle_params(oracle, p):
    return oracle.McmcParams.from_buffer_copy(p)


def run_oracle_chains(oracle, rois, p, seed, src=None, bg0=10.0, record=True, grids=False, back=None,
                      sig=None, off=None, scmos=None, n0=None, psf=None, chain=False):
    """One oracle chain per ROI.  Returns (results, tapes)."""
    src = oracle.SRC_PHILOX if src is None else src
    op = oracle_params(oracle, p)
    out, tapes = [], []
    for r in range(len(rois["side"]) if "side" in rois else len(rois["xsize"])):
        o = rois["offset"]
        xs = int(rois["xsize"][r]) if "xsize" in rois else int(rois["side"][r])
        ys = int(rois["ysize"][r]) if "ysize" in rois else int(rois["side"][r])
        kw = {}
        if n0 is not None:
            kw = dict(n0=int(n0[r][0].size), i0=n0[r][0], x0=n0[r][1], y0=n0[r][2])
        res = oracle.rjmcmc(rois["data"][o[r]:o[r + 1]], xs, ys, op, OFF_PDF if off is None else off,
                            SIG_PDF if sig is None else sig, BACK_PDF if back is None else back,
                            bg0=float(bg0[r]) if np.ndim(bg0) else bg0, src_kind=src, seed=seed, roi=r,
                            record=record, grids=grids, scmos=None if scmos is None else scmos[o[r]:o[r + 1]],
                            psf=psf, chain=chain, **kw)
        out.append(res)
        tapes.append(res.get("tape"))
    return out, tapes
