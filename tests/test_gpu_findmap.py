"""findMap and the mcmc.m refinement summary on the device (bamf_find_map,
bamf_chain_summary) against the oracle restatement on the same Chain
records of real runs."""
import numpy as np
import pytest

from bamf_amd import synth
from bamf_amd.rjmcmc import PLAIN, Plan, RoiSet, alloc_outputs, chain_summary, fetch, find_map, run_batch
from oracle import findmap as fm
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params

pytestmark = pytest.mark.gpu

CAPS = (2200, 40000)


def run_with_records(p, rois, n0=None):
    n = len(rois["side"])
    rs = RoiSet.from_synth(rois, bg0=np.full(n, 10.0, np.float32)) if n0 is None else n0
    total = p.n_burnin + p.n_trials
    outs = alloc_outputs(n, total, 64, 128, 0, False, False, rs, p.grid_zoom, CAPS)
    run_batch(Plan(p, OFF_PDF, SIG_PDF, BACK_PDF), rs, seed=31, outs=outs, sync=False)
    res = fetch(outs, n, total, 64, 128, 0)
    assert np.all(res.status == 0)
    return outs, res


def test_find_map_matches_oracle():
    rois = synth.make_rois(48, seed=71, sides=(8, 12, 16), mean_n=4.0)
    p = params(nb=1000, nt=1000)
    outs, res = run_with_records(p, rois)
    print("find_map: 48 ROIs", flush=True)
    g = find_map(outs, 48, fb=1)
    ch = res.chain
    nonzero = 0
    for r in range(48):
        nloop = int(ch["count"][r, 0])
        mf, bg, cl = fm.find_map(ch["acc_rec"][r], ch["acc_em"][r], nloop, fb=1)
        assert g["status"][r] == 0 and g["n"][r] == mf, r
        np.testing.assert_allclose(g["bg"][r], bg, rtol=2e-6, atol=1e-6)
        k = len(cl)
        nonzero += k > 0
        np.testing.assert_allclose(g["em"][r, :k, :3], cl[:, :3], rtol=2e-6, atol=1e-5)
        np.testing.assert_allclose(g["em"][r, :k, 3:], cl[:, 3:], rtol=1e-4, atol=1e-4)
    assert nonzero > 30


def test_chain_summary_matches_oracle():
    rois = synth.make_rois(32, seed=72, sides=(12, 16), mean_n=3.0)
    n = 32
    inits = [t for t in rois["truth"]]
    n0 = np.array([len(t[0]) for t in inits], np.int32)
    eoff = np.zeros(n, np.int64)
    eoff[1:] = np.cumsum(n0)[:-1]
    em = (np.concatenate([t[0] for t in inits]).astype(np.float32),
          np.concatenate([t[1] for t in inits]).astype(np.float32),
          np.concatenate([t[2] for t in inits]).astype(np.float32), eoff)
    rs = RoiSet(rois["data"], rois["offset"], rois["side"], rois["side"], bg0=np.full(n, 10.0, np.float32),
                n0=n0, em0=em)
    # mcmc.m:103: moves only
    p = params(nb=1000, nt=1000, p_burnin=(1, 0, 0, 0, 0, 0, 0, 0), p_trials=(1, 0, 0, 0, 0, 0, 0, 0))
    outs, res = run_with_records(p, rois, n0=rs)
    print("chain_summary: 32 ROIs", flush=True)
    g = chain_summary(outs, n)
    ch = res.chain
    for r in range(n):
        k, out = fm.chain_summary(ch["acc_rec"][r], ch["acc_em"][r], int(ch["count"][r, 0]))
        if k is None:  # no emitters: moves end before any record (LibRJMCMC.cpp:769)
            assert int(ch["count"][r, 0]) < 2 and g["n"][r] == 0, r
            continue
        assert g["status"][r] == 0 and g["n"][r] == k, r
        np.testing.assert_allclose(g["em"][r, :k, :3], out[:, :3], rtol=2e-6, atol=1e-5)
        np.testing.assert_allclose(g["em"][r, :k, 3:], out[:, 3:], rtol=1e-4, atol=1e-5)
