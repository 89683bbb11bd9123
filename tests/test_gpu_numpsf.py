"""GPU chain with the sampled-PSF model (SampledPSF, OptModel = 1: NumModel,
LibRJMCMC.cpp:62-170) against the oracle chain with the same model: replay
parity to the standard of the erf-model replay tests (discrete outcomes exact,
log PR within 1e-5 relative plus the single-precision floor), the same run
through the host gateway, free-running agreement under the same Philox
stream, and the argument checks."""
import numpy as np
import pytest

from bamf_amd import synth
from bamf_amd._lib import BamfError, lib
from bamf_amd.rjmcmc import PLAIN, REPLAY, Plan, RoiSet, make_replay, run_batch
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params, run_oracle_chains
from test_gpu_gateway import crjmcmc
from test_gpu_rjmcmc import check_replay

pytestmark = pytest.mark.gpu


def tables():
    """An elliptical erf table large enough for 32-pixel ROIs, and a small one
    whose far pixels read past the table (its last entry made visible)."""
    big = synth.sampled_gauss_psf(72, 4, 1.2, 1.5)
    small = np.array(synth.sampled_gauss_psf(12, 3, 1.4), order="F")
    small[-1, -1, -1, -1] = 1e-3
    return {"big": big, "small": small}


def uneven_rois(n, seed):
    rng = np.random.default_rng(seed)
    xs = rng.choice([8, 12, 16, 24, 32], n)
    ys = rng.choice([8, 12, 20, 32], n)
    offs = np.zeros(n + 1, np.int64)
    offs[1:] = np.cumsum(xs * ys)
    data = np.zeros(offs[-1], np.float32)
    for r in range(n):
        k = rng.poisson(0.05 * xs[r] * ys[r] / 4)
        X, Y = rng.uniform(0, xs[r], k), rng.uniform(0, ys[r], k)
        I = rng.gamma(5, 200, k)
        data[offs[r]:offs[r + 1]] = rng.poisson(synth.roi_expected(I, X, Y, 12.0, int(xs[r]), int(ys[r]), 1.3))
    return {"data": data, "offset": offs, "xsize": xs, "ysize": ys}


@pytest.mark.parametrize("which", ["big", "small"])
def test_replay_sampled_psf(oracle, which):
    psf = tables()[which]
    rois = uneven_rois(16, 31 if which == "big" else 32)
    p = params(nb=1200, nt=1200, is_backg=1, p_burnin=(0.3, 0.15, 0.15, 0.15, 0.15, 0, 0, 0.1),
               p_trials=(0.5, 0.1, 0.1, 0.1, 0.1, 0, 0, 0.1))
    back = synth.gamma_pdf_table(2, 40, 10.0, 1001)
    ores, tapes = run_oracle_chains(oracle, rois, p, seed=13, bg0=12.0, back=back, grids=True, psf=psf)
    plan = Plan(p, OFF_PDF, SIG_PDF, back, sampled_psf=psf)
    n = len(rois["xsize"])
    rs = RoiSet(rois["data"], rois["offset"], rois["xsize"], rois["ysize"], bg0=np.full(n, 12.0, np.float32))
    rp = make_replay(tapes, np.stack([o["steps"] for o in ores]))
    print(f"replay sampled PSF ({which}, dims {psf.shape}): {n} ROIs", flush=True)
    g = run_batch(plan, rs, mode=REPLAY, replay=rp, record_steps=True, grids=True, hist_len=129)
    d, offs = rois["data"], rois["offset"]
    check_replay(rois, g, ores, tapes, p, lambda r: d[offs[r]:offs[r + 1]].sum())
    off = 0
    for r in range(n):
        sz = int(rois["xsize"][r] * rois["ysize"][r] * p.grid_zoom ** 2)
        assert np.array_equal(g.pgrid[off:off + sz], ores[r]["pgrid"]), r
        assert np.array_equal(g.pback[off:off + sz], ores[r]["pback"]), r
        off += sz


def test_free_running_sampled_psf_matches_oracle(oracle):
    psf = tables()["big"]
    rois = synth.make_rois(24, seed=303, sides=(16,), mean_n=3.0, n_cap=10)
    p = params(nb=1500, nt=1500)
    ores, _ = run_oracle_chains(oracle, rois, p, seed=17, record=False, psf=psf)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF, sampled_psf=psf)
    rs = RoiSet.from_synth(rois, bg0=np.full(24, 10.0, np.float32))
    print("free-running sampled PSF: 24 ROIs", flush=True)
    g = run_batch(plan, rs, seed=17, mode=PLAIN, record_steps=True, hist_len=129)
    assert np.all(g.status == 0)
    full = 0
    for r in range(24):
        o, gs = ores[r]["steps"], g.steps[r]
        same = (gs["n"] == o["n"]) & (gs["draws"] == o["draws"]) & (gs["kind"] == o["kind"])
        if same.all():
            full += 1
            assert np.array_equal(g.counters[r], ores[r]["counters"])
        else:
            s = int(np.argmin(same))
            assert (gs["flags"][s] & 1) != (o["flags"][s] & 1) or gs["pix"][s] != o["pix"][s], (r, s)
    assert full >= 18, full


def test_gateway_sampled_psf_matches_batched_entry():
    psf = tables()["big"]
    rois = synth.make_rois(1, seed=77, sides=(14,), mean_n=3.0)
    p = params(nb=600, nt=600)
    data = rois["data"].astype(np.float32)
    rc, pg, *_ = crjmcmc(p, data, 14, 14, 10.0, 5, psf=psf)
    assert rc == 0, lib().bamf_last_error()
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF, sampled_psf=psf)
    g = run_batch(plan, RoiSet(data, np.array([0, data.size]), [14], [14], bg0=[10.0]), seed=5, grids=True)
    assert np.array_equal(pg.view(np.uint32), g.pgrid.view(np.uint32))
    # the same plan back on the erf model runs the erf chain
    plan.set_sampled_psf(None)
    ge = run_batch(plan, RoiSet(data, np.array([0, data.size]), [14], [14], bg0=[10.0]), seed=5, grids=True)
    pe = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    ge2 = run_batch(pe, RoiSet(data, np.array([0, data.size]), [14], [14], bg0=[10.0]), seed=5, grids=True)
    assert np.array_equal(ge.pgrid.view(np.uint32), ge2.pgrid.view(np.uint32))


def test_sampled_psf_argument_checks():
    p = params(nb=10, nt=10)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    with pytest.raises(BamfError):
        plan.set_sampled_psf(np.zeros((4, 4, 2), np.float32))
    rc, *_ = crjmcmc(p, np.zeros(64, np.float32), 8, 8, 1.0, 1, psf=np.zeros((0, 4, 2, 2), np.float32))
    assert rc != 0 and b"PSFsize" in lib().bamf_last_error()


def test_sampled_psf_segments_and_graph_equal_one_launch():
    """The sampled-PSF kernel carries the segment feature: a chain split into
    launches (its model image kept in HBM between them), and the sequence
    replayed from a HIP graph with new seeds, equal one-launch runs bit for bit."""
    from bamf_amd import _hip
    from bamf_amd.rjmcmc import alloc_outputs, fetch
    from test_gpu_segments import KW, assert_same, rois_uneven

    psf = tables()["big"]
    rois = rois_uneven(96, seed=62)
    n = len(rois["side"])
    p = params(nb=500, nt=500, rho=0.0275)
    rs = RoiSet.from_synth(rois, bg0=np.full(n, 20.0, np.float32))
    ref = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF, sampled_psf=psf)
    one = run_batch(ref, rs, seed=9, **KW)
    assert np.all(one.status == 0)
    seg_plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF, sampled_psf=psf)
    seg_plan.set_segments(4)
    print("sampled PSF: 4 segments", flush=True)
    assert_same(one, run_batch(seg_plan, rs, seed=9, **KW), n, 100)  # 500 trials thinned by 5
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF, sampled_psf=psf)
    plan.set_segments(4, graph=True)
    total = p.n_burnin + p.n_trials
    outs = alloc_outputs(n, total, KW["hist_len"], KW["final_cap"], KW["trace_len"], True, True, rs, p.grid_zoom,
                         KW["chain_caps"])
    for it, seed in enumerate((9, 10, 9)):
        for k in ("pgrid", "pback"):
            outs[k].upload(np.zeros(outs[k].nbytes // 4, np.float32))
        print(f"sampled PSF graph replay {it}, seed {seed}", flush=True)
        run_batch(plan, rs, seed=seed, outs=outs, trace_thin=KW["trace_thin"], sync=False)
        _hip.synchronize()
        got = fetch(outs, n, total, KW["hist_len"], KW["final_cap"], KW["trace_len"])
        want = one if seed == 9 else run_batch(ref, rs, seed=seed, **KW)
        assert_same(want, got, n, 100)
