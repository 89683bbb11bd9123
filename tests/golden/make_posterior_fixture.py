"""Writes tests/golden/posterior_batched_uneven.npz: long oracle (CPU
restatement) reference chains for the posterior-agreement test.

ROIs: ten of the batched uneven workload (20,000 ROIs, sides {8, 12, 16, 24,
32}, density U[0.5, 5] /um^2, Gamma(5, 200) photons, background U[5, 30],
seeded as the benchmark seeds it), two per side, among them dense 32 x 32
ROIs of 30 and more emitters.  Per ROI, `chains` independent oracle chains of
n_burnin + n_trials = 1,000,000 jumps with the benchmark's parameters
(P_Burnin [.4 .15 .15 .15 .15 0 0 0], P_Trials [.6 .1 .1 .1 .1 0 0 0], rho
0.0275 per pixel, offset prior Gamma(4, 5), signal prior Gamma(5, 200));
each chain's final state is one posterior sample (independent samples: the
emitter count of a dense ROI decorrelates over hundreds of thousands of
jumps, so samples thinned from one chain would not be; at 100,000 jumps the
30-emitter ROI's chains have not yet forgotten their empty start).  Stored: the ROIs and their
truth, the parameters and prior tables, every sample's N, the per-ROI
localization errors of the samples and the MAP-N localizations
(bamf_amd/posterior.py).

    python tests/golden/make_posterior_fixture.py [--procs 8]
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "posterior_batched_uneven.npz")
WORKLOAD_SEED = 20261203 + 2   # the benchmark's batched uneven workload
CHAIN_SEED = 777
DX = 10.0
RHO = 2.75 * 0.01
P_BURNIN = (0.4, 0.15, 0.15, 0.15, 0.15, 0, 0, 0)
P_TRIALS = (0.6, 0.1, 0.1, 0.1, 0.1, 0, 0, 0)
N_BURNIN, N_TRIALS = 20000, 980000
CHAINS = 128

_G = {}


def pick_rois(rois):
    side = rois["side"]
    n_true = rois["n_true"]
    picks = []
    for s in (8, 12, 16, 24):
        idx = np.where((side == s) & (n_true >= 1))[0]
        med = np.median(n_true[idx])
        order = idx[np.argsort(np.abs(n_true[idx] - med), kind="stable")]
        picks += [int(order[0]), int(order[1])]
    idx32 = np.where(side == 32)[0]
    picks.append(int(idx32[np.argmin(np.abs(n_true[idx32] - 30))]))
    picks.append(int(idx32[np.argmin(np.abs(n_true[idx32] - 43))]))
    return picks


def _chain(job):
    from oracle import pyoracle as po

    r, c = job
    g = _G
    p = g["params"]
    op = po.McmcParams.from_buffer_copy(p)
    o = g["offset"]
    s = int(g["side"][r])
    res = po.rjmcmc(g["data"][o[r]:o[r + 1]], s, s, op, g["off"], g["sig"], g["back"], bg0=g["bg0"],
                    seed=CHAIN_SEED, roi=int(r) * 1000 + c)
    assert res["status"] == 0
    return r, c, res["n"], np.stack([res["X"], res["Y"]], 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=os.cpu_count())
    ap.add_argument("--chains", type=int, default=CHAINS)
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--raw", default=None, help="also write every chain's final state to this .npz")
    args = ap.parse_args()
    from bamf_amd import synth
    from bamf_amd.rjmcmc import make_params
    from oracle import pyoracle as po
    from bamf_amd.posterior import localization_rmse_nm, map_localizations, posterior_rmse_nm, sample_errors

    po.build()
    rois = synth.make_rois(20000, seed=WORKLOAD_SEED, sides=(8, 12, 16, 24, 32), density=(0.5, 5.0),
                           bg=(5.0, 30.0))
    picks = pick_rois(rois)
    sig = synth.gamma_pdf_table(5, 200, DX, 1001)
    off = synth.gamma_pdf_table(4, 5, DX, 1001)
    back = np.zeros(1001, np.float32)
    bg0 = float(DX * int(np.argmax(off)))  # @RJ/rjmcmc.m:47
    params = make_params(n_burnin=N_BURNIN, n_trials=N_TRIALS, rho=RHO, p_burnin=P_BURNIN, p_trials=P_TRIALS)
    sides = rois["side"][picks]
    offs = np.zeros(len(picks) + 1, np.int64)
    offs[1:] = np.cumsum(sides.astype(np.int64) ** 2)
    data = np.concatenate([rois["data"][rois["offset"][r]:rois["offset"][r + 1]] for r in picks])
    _G.update(params=params, offset=offs, side=sides, data=data, off=off, sig=sig, back=back, bg0=bg0)
    jobs = [(r, c) for r in range(len(picks)) for c in range(args.chains)]
    # longest first
    jobs.sort(key=lambda j: -int(sides[j[0]]) ** 2 * (1 + rois["n_true"][picks[j[0]]]))
    t0 = time.time()
    samples = {r: [None] * args.chains for r in range(len(picks))}
    with Pool(args.procs) as pool:
        for k, (r, c, n, xy) in enumerate(pool.imap_unordered(_chain, jobs)):
            samples[r][c] = (n, xy)
            if k % 50 == 0:
                print(f"{k}/{len(jobs)} chains, {time.time() - t0:.0f} s", flush=True)
    ns_all, locs, truths, errs = [], [], [], []
    for r in range(len(picks)):
        ns = np.array([s[0] for s in samples[r]])
        pos = [s[1] for s in samples[r]]
        ns_all.append(ns)
        locs.append(map_localizations(ns, pos))
        I, X, Y = rois["truth"][picks[r]]
        truths.append(np.stack([X, Y], 1))
        errs.append(sample_errors(ns, pos, truths[-1]))
    if args.raw:
        raw = {}
        for r in range(len(picks)):
            for c in range(args.chains):
                raw[f"n_{r}_{c}"], raw[f"em_{r}_{c}"] = samples[r][c]
        np.savez_compressed(args.raw, **raw)
    rmse, pairs = localization_rmse_nm(locs, truths)
    prmse = posterior_rmse_nm(errs)
    print(f"oracle posterior localization RMSE {prmse:.3f} nm; MAP-N localization RMSE {rmse:.2f} nm over {pairs} "
          f"pairs; {time.time() - t0:.0f} s")
    flat = lambda arrs: (np.concatenate(arrs) if arrs else np.zeros(0), np.array([len(a) for a in arrs]))  # noqa
    ns_flat, ns_len = flat(ns_all)
    loc_flat, loc_len = flat(locs)
    tr_flat, tr_len = flat(truths)
    tI = np.concatenate([np.asarray(rois["truth"][r][0], np.float64) for r in picks])
    np.savez_compressed(
        args.out, roi_index=np.array(picks), side=sides, offset=offs, data=data, bg_true=rois["bg"][picks],
        truth_xy=tr_flat.astype(np.float32), truth_i=tI.astype(np.float32), truth_len=tr_len,
        sig_pdf=sig, off_pdf=off, bg0=np.float32(bg0), rho=np.float32(RHO), p_burnin=np.array(P_BURNIN, np.float32),
        p_trials=np.array(P_TRIALS, np.float32), n_burnin=N_BURNIN, n_trials=N_TRIALS,
        chains=args.chains, chain_seed=CHAIN_SEED, oracle_err=np.array(errs, np.float64),
        oracle_post_rmse_nm=np.float64(prmse),
        oracle_n=ns_flat.astype(np.int16), oracle_n_len=ns_len, oracle_loc=loc_flat.astype(np.float32),
        oracle_loc_len=loc_len, oracle_rmse_nm=np.float64(rmse))
    print("wrote", args.out)


if __name__ == "__main__":
    main()
