"""Segmented chains (bamf_rjmcmc_set_segments): a chain split into several
launches, its state kept in device memory between them, and the launch
sequence replayed from a HIP graph, must give exactly the one-launch run's
outputs (states, counters, histograms, traces, step records, grids, chain
records, log posteriors, work counts)."""
import numpy as np
import pytest

from bamf_amd import _hip, synth
from bamf_amd.rjmcmc import PLAIN, REPLAY, Plan, RoiSet, alloc_outputs, fetch, make_replay, run_batch
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params, run_oracle_chains

pytestmark = pytest.mark.gpu

KW = dict(hist_len=129, final_cap=128, trace_len=120, trace_thin=5, record_steps=True, grids=True,
          chain_caps=(1300, 12000))


def rois_uneven(n=160, seed=61):
    return synth.make_rois(n, seed=seed, sides=(8, 12, 16, 24, 32), density=(0.5, 5.0), bg=(5.0, 30.0))


def assert_same(a, b, n, n_trace):
    """Bitwise equality of every written output (buffers are not cleared, so
    final-state slots beyond N and trace slots beyond the chain are skipped)."""
    for f in ("status", "counters", "hist", "logpost", "alg_flops", "pgrid", "pback"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.tobytes() == y.tobytes(), f
    assert a.trace_n[:, :n_trace].tobytes() == b.trace_n[:, :n_trace].tobytes()
    assert a.trace_photons[:, :n_trace].tobytes() == b.trace_photons[:, :n_trace].tobytes()
    cap = (a.final.shape[1] - 4) // 4
    for r in range(n):
        k = int(a.final[r, 0])
        assert a.final[r, 0:4].tobytes() == b.final[r, 0:4].tobytes(), r  # N, BG, ABG, BBG
        for j in range(4):  # I, X, Y, Signal of the first N slots
            sl = slice(4 + j * cap, 4 + j * cap + k)
            assert a.final[r, sl].tobytes() == b.final[r, sl].tobytes(), (r, j)
    assert a.steps.tobytes() == b.steps.tobytes()
    ca, cb = a.chain, b.chain
    assert np.array_equal(ca["count"], cb["count"])
    for r in range(n):
        k, e1, e2 = (int(v) for v in ca["count"][r])
        k = min(k, 1300)
        for f in ("acc_rec", "prop_rec"):
            assert ca[f][r, :k].tobytes() == cb[f][r, :k].tobytes(), (r, f)
        assert ca["acc_em"][r, :min(e1, 12000)].tobytes() == cb["acc_em"][r, :min(e1, 12000)].tobytes(), r
        assert ca["prop_em"][r, :min(e2, 12000)].tobytes() == cb["prop_em"][r, :min(e2, 12000)].tobytes(), r


@pytest.mark.parametrize("nseg,launch", [(2, 2), (7, 2), (3, 1)])
def test_segments_equal_one_launch(nseg, launch):
    """launch 2: each segment one launch over both size classes (the default);
    1: a launch per class and segment, the classes side by side."""
    rois = rois_uneven()
    n = len(rois["side"])
    p = params(nb=600, nt=600, rho=0.0275)
    rs = RoiSet.from_synth(rois, bg0=np.full(n, 20.0, np.float32))
    one = run_batch(Plan(p, OFF_PDF, SIG_PDF, BACK_PDF), rs, seed=5, **KW)
    assert np.all(one.status == 0)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    plan.set_tuning((4, 4), launch)
    plan.set_segments(nseg)
    print(f"segments: {nseg}, launch setting {launch}", flush=True)
    seg = run_batch(plan, rs, seed=5, **KW)
    assert_same(one, seg, n, 120)


def test_logpost_output_optional():
    """Without the log-PR total (d_logpost NULL) the chains and every other
    output are the same, bit for bit."""
    rois = rois_uneven(64, seed=65)
    n = len(rois["side"])
    p = params(nb=400, nt=400, rho=0.0275)
    rs = RoiSet.from_synth(rois, bg0=np.full(n, 20.0, np.float32))
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    kw = dict(hist_len=129, final_cap=128, trace_len=80, trace_thin=5)
    base = run_batch(plan, rs, seed=13, **kw)
    assert np.all(base.status == 0) and base.logpost is not None
    total = p.n_burnin + p.n_trials
    outs = alloc_outputs(n, total, 129, 128, 80, False, False, rs, p.grid_zoom, logpost=False)
    print("run without the log-PR total", flush=True)
    run_batch(plan, rs, seed=13, outs=outs, trace_thin=5, sync=False)
    _hip.synchronize()
    got = fetch(outs, n, total, 129, 128, 80)
    assert got.logpost is None
    for f in ("status", "counters", "hist", "alg_flops", "n_final", "trace_n", "trace_photons"):
        assert getattr(base, f).tobytes() == getattr(got, f).tobytes(), f
    for r in range(n):
        k = int(base.final[r, 0])
        assert base.final[r, 0:4].tobytes() == got.final[r, 0:4].tobytes(), r  # N, BG, ABG, BBG
        for j in range(4):
            sl = slice(4 + j * 128, 4 + j * 128 + k)
            assert base.final[r, sl].tobytes() == got.final[r, sl].tobytes(), (r, j)


def test_graph_replays_with_new_seeds():
    rois = rois_uneven(96, seed=62)
    n = len(rois["side"])
    p = params(nb=500, nt=500, rho=0.0275)
    rs = RoiSet.from_synth(rois, bg0=np.full(n, 20.0, np.float32))
    ref = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    plan.set_segments(4, graph=True)
    total = p.n_burnin + p.n_trials
    outs = alloc_outputs(n, total, KW["hist_len"], KW["final_cap"], KW["trace_len"], True, True, rs, p.grid_zoom,
                         KW["chain_caps"])
    for it, seed in enumerate((11, 12, 11)):
        # the grids accumulate: clear them between replays, as for any run
        for k in ("pgrid", "pback"):
            outs[k].upload(np.zeros(outs[k].nbytes // 4, np.float32))
        print(f"graph replay {it}, seed {seed}", flush=True)
        run_batch(plan, rs, seed=seed, outs=outs, trace_thin=KW["trace_thin"], sync=False)
        _hip.synchronize()
        got = fetch(outs, n, total, KW["hist_len"], KW["final_cap"], KW["trace_len"])
        want = run_batch(ref, rs, seed=seed, **KW)
        assert np.all(want.status == 0)
        assert_same(want, got, n, 100)


def test_segments_in_replay_mode(oracle):
    rois = synth.make_rois(12, seed=63, sides=(8, 16, 24), mean_n=3.0)
    n = 12
    p = params(nb=400, nt=400)
    ores, tapes = run_oracle_chains(oracle, rois, p, seed=3)
    rp = make_replay(tapes, np.stack([o["steps"] for o in ores]))
    rs = RoiSet.from_synth(rois, bg0=np.full(n, 10.0, np.float32))
    one = run_batch(Plan(p, OFF_PDF, SIG_PDF, BACK_PDF), rs, mode=REPLAY, replay=rp, **KW)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    plan.set_segments(5)
    print("replay segments: 5", flush=True)
    seg = run_batch(plan, rs, mode=REPLAY, replay=rp, **KW)
    assert np.all(one.status == 0)
    assert_same(one, seg, n, 80)


def test_tuning_changes_no_result():
    """Every launch setting (one launch over both size classes, a launch per
    class side by side or one after the other), every register budget and
    every ROI tile bound (the ROIs of at most that many pixel slots run in the
    4-slot class) give the same chains, bit for bit; the autotuner returns one
    of the compiled settings."""
    rois = rois_uneven(120, seed=64)
    n = len(rois["side"])
    p = params(nb=400, nt=400, rho=0.0275)
    rs = RoiSet.from_synth(rois, bg0=np.full(n, 20.0, np.float32))
    kw = dict(KW, trace_len=80)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    assert plan.tuning() == {"chains_per_simd": [4, 4], "concurrent": 2, "small_max_slots": 4}
    base = run_batch(plan, rs, seed=9, **kw)
    assert np.all(base.status == 0)
    for cps, conc, sm in (((4, 4), 1, 4), ((4, 4), 0, 4), ((3, 3), 1, 4), ((5, 5), 1, 4), ((3, 5), 0, 4),
                          ((5, 3), 1, 4), ((4, 4), 2, 1), ((4, 4), 2, 2), ((3, 5), 1, 1), ((4, 4), 2, 4)):
        plan.set_tuning(cps, conc, sm)
        assert plan.tuning()["small_max_slots"] == sm
        print(f"tuning {cps} concurrent={conc} small_max_slots={sm}", flush=True)
        assert_same(base, run_batch(plan, rs, seed=9, **kw), n, 80)
    best = plan.autotune(rs, sample_jumps=200)
    print("autotune:", best, flush=True)
    assert best["ms"] > 0 and all(c in (3, 4, 5) for c in best["chains_per_simd"]) and best["concurrent"] in (0, 1, 2)
    assert best["small_max_slots"] in (1, 2, 4)
    assert plan.tuning()["chains_per_simd"] == best["chains_per_simd"]
    assert_same(base, run_batch(plan, rs, seed=9, **kw), n, 80)
    with pytest.raises(Exception):
        plan.set_tuning((2, 4), 1)
    with pytest.raises(Exception):
        plan.set_tuning((3, 3), 2)  # the one launch over both classes has one register budget
    with pytest.raises(Exception):
        plan.set_tuning((4, 4), 3)
    with pytest.raises(Exception):
        plan.set_tuning((4, 4), 2, 5)  # the small class holds at most 4 slots per lane
