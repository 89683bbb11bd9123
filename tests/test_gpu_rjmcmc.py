"""GPU chain (bamf_rjmcmc_run) against the oracle chain.

Replay: the GPU chain consumes the oracle's recorded draws and follows its
recorded decisions and birth pixels.  The proposals are then computed from
identical states, so every step must match the oracle exactly in what it
proposes (kind, JumpType, draws, bounds-test outcome, N, BG, photons, final
emitters, grids, counters).  The GPU's own accept decisions must equal the
oracle's except at near-ties.  Its log PR is held to the oracle's own
single-precision error, per step: the likelihood-ratio sum
S = sum(M - M' + D log(M'/M)) is evaluated in float64 from the step's two
states (S64); the oracle's float32 sum S32 (its float32 models, terms and
left-to-right sum) misses it by e_o = |S32 - S64|; the GPU's sum (S32 plus the
difference of the two log PRs, whose other factors are the same floats) must
miss it by at most 2 max(e_o, r), r the float32 rounding scale of the step's
sum: 2^-24 times the quadrature sum of the roundings of every changed model
value (M and M', each entering its term with weight 1 + D/M), every term
and every partial sum.

Free-running: both use the same Philox stream, so chains agree step for step
until a near-tie decision differs.
"""
import math

import numpy as np
import pytest
from scipy import special

from bamf_amd import synth
from bamf_amd.rjmcmc import PLAIN, REPLAY, Plan, RoiSet, make_replay, run_batch
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params, run_oracle_chains

pytestmark = pytest.mark.gpu


def accept_uniform(tape, steps):
    """The accept uniform of each step (the last draw) as float."""
    ends = np.cumsum(steps["draws"])
    u = np.full(len(steps), np.nan)
    ok = steps["draws"] > 0
    u[ok] = tape[ends[ok] - 1].astype(np.float64) / 2147483648.0
    return u


def _model64(I, X, Y, bg, abg, bbg, xs, ys, sigma):
    """The pixel-integrated Gaussian model (LibRJMCMC.cpp:42-59) in float64, layout y + x*ys."""
    c = 1.0 / (sigma * math.sqrt(2.0))
    x = np.arange(xs, dtype=np.float64)
    y = np.arange(ys, dtype=np.float64)
    m = bg + abg * x[:, None] + bbg * y[None, :]
    if len(I):
        I, X, Y = (np.asarray(v, np.float64) for v in (I, X, Y))
        gx = 0.5 * (special.erf((x[None, :] + 1 - X[:, None]) * c) - special.erf((x[None, :] - X[:, None]) * c))
        gy = 0.5 * (special.erf((y[None, :] + 1 - Y[:, None]) * c) - special.erf((y[None, :] - Y[:, None]) * c))
        m = m + np.einsum("e,ex,ey->xy", I, gx, gy)
    return m.ravel()


def _states(ch, which):
    head, val, em = ch[which + "_head"], ch[which + "_val"], ch[which + "_em"]
    out = []
    for k in range(len(head)):
        n, at = int(head[k, 0]), int(head[k, 3])
        e = em[at:at + n]
        out.append((e[:, 2], e[:, 0], e[:, 1], float(val[k, 2]), float(val[k, 3]), float(val[k, 4])))
    return out


def llr_sum_errors(oracle, ores_r, data, xs, ys, sigma, bg0, init=None):
    """Per step with a likelihood evaluation: (step, S64, e_o = |S32 - S64|, rounding scale)."""
    st, ch = ores_r["steps"], ores_r["chain"]
    acc, prop = _states(ch, "acc"), _states(ch, "prop")
    d64 = data.astype(np.float64)
    P = xs * ys
    before = init if init is not None else (np.zeros(0), np.zeros(0), np.zeros(0), bg0, 0.0, 0.0)
    out, k = [], 0
    for s in np.flatnonzero((st["flags"] & 2) != 0):
        a, t = before, prop[k]
        if not (st["flags"][s] & 8):
            m32 = oracle.model(a[0], a[1], a[2], a[3], a[4], a[5], xs, ys, sigma)
            mt32 = oracle.model(t[0], t[1], t[2], t[3], t[4], t[5], xs, ys, sigma)
            # logf as the oracle's libm evaluates it: the float64 log rounded once to float32
            # (numpy's own float32 log is off by a few ulps)
            lq = np.log((mt32 / m32).astype(np.float64)).astype(np.float32)
            terms32 = (m32 - mt32) + data * lq
            s32 = float(np.add.accumulate(terms32.astype(np.float32), dtype=np.float32)[-1])
            m64 = _model64(*a, xs, ys, sigma)
            mt64 = _model64(*t, xs, ys, sigma)
            terms64 = (m64 - mt64) + d64 * np.log(mt64 / m64)
            s64 = float(terms64.sum())
            # float32 rounding scale of this step's sum, independent roundings in quadrature:
            # each changed model value (weight 1 + D/M in its term), each term, each partial
            # sum.  The device forms M' incrementally (M plus or minus stencils), so M' is
            # rounded at the larger of the two magnitudes: removing a bright emitter leaves
            # the rounding of the brighter M in the dimmer M'.
            ch = np.abs(mt64 - m64) > 2.0 ** -24 * m64
            big = np.maximum(m64, mt64)[ch]
            w = (m64 + d64)[ch] ** 2 + (big * (1.0 + d64[ch] / mt64[ch])) ** 2
            scale = 2.0 ** -24 * math.sqrt(float(w.sum() + (terms64 ** 2).sum() + (np.cumsum(terms64) ** 2).sum()))
            out.append((s, s64, s32, scale))
        before = acc[k]
        k += 1
    return out


def check_llr_against_float64(oracle, gres, ores, rois_data, offs, xsz, ysz, sigma, bg0, inits=None):
    """GPU log-likelihood-ratio sums within twice the oracle's own float32 error of the float64 sum."""
    ratios, excess, n_steps = [], [], 0
    for r in range(len(ores)):
        o, gs = ores[r]["steps"], gres.steps[r]
        data = rois_data[offs[r]:offs[r + 1]]
        init = None
        if inits is not None:
            I, X, Y = inits[r]
            init = (I, X, Y, bg0, 0.0, 0.0)
        for s, s64, s32, scale in llr_sum_errors(oracle, ores[r], data, int(xsz[r]), int(ysz[r]), sigma, bg0, init):
            if not (np.isfinite(o["log_pr"][s]) and np.isfinite(gs["log_pr"][s])):
                continue
            sg = s32 + (float(gs["log_pr"][s]) - float(o["log_pr"][s]))
            e_o, e_g = abs(s32 - s64), abs(sg - s64)
            ratios.append(e_g / max(e_o, scale))
            excess.append(((e_g - 2.0 * e_o) / scale, r, int(s), e_g, e_o, scale))
            n_steps += 1
    ratios = np.array(ratios)
    q = np.quantile(ratios, [0.5, 0.99, 0.999, 1.0])
    worst = max(excess)
    print(f"float64 re-evaluation over {n_steps} steps: GPU error / max(oracle error, rounding scale) median "
          f"{q[0]:.2f}, p99 {q[1]:.2f}, p99.9 {q[2]:.2f}, max {q[3]:.2f}; worst (e_g - 2 e_o) / scale "
          f"{worst[0]:.2f} (roi, step, e_g, e_o, scale = {worst[1:]})", flush=True)
    # every step: within twice the oracle's error plus 3 sigma of the step's own rounding
    assert worst[0] <= 3.0, worst
    # and 99.9 % of steps within twice the larger of the two
    assert np.mean(ratios <= 2.0) >= 0.999, np.mean(ratios <= 2.0)


def check_replay(rois, gres, ores, tapes, p, data_sum, stats=None):
    g = gres.steps
    n_roi = g.shape[0]
    nd_total = 0
    rel_errs = []
    for r in range(n_roi):
        o = ores[r]["steps"]
        gs = g[r]
        assert gres.status[r] == 0, (r, gres.status[r])
        for f in ("kind", "jump_type", "n", "draws", "n_prop"):
            assert np.array_equal(gs[f], o[f]), (r, f, np.flatnonzero(gs[f] != o[f])[:5])
        assert np.array_equal(gs["bg"].view(np.uint32), o["bg"].view(np.uint32)), r
        assert np.array_equal(gs["photons"].view(np.uint32), o["photons"].view(np.uint32)), r
        assert np.array_equal(gs["flags"] & 0b1110, o["flags"] & 0b1110), r
        births = o["kind"] == 3
        if births.any():
            assert np.mean(gs["pix"][births] == o["pix"][births]) >= 0.99
        lp_o, lp_g = o["log_pr"], gs["log_pr"]
        # near-tie band for the decision check: 1e-5 relative plus 16 single-precision
        # ulps of sum D (the log PR itself is held to the float64 re-evaluation)
        tol = 1e-5 * np.abs(lp_o) + 16 * 2.0 ** -23 * max(float(data_sum(r)), 1.0)
        fin = np.isfinite(lp_o) & (o["pr"] > 1e-30) & (o["pr"] < 1e30) & np.isfinite(o["pr"])
        assert np.array_equal(np.isfinite(lp_g[fin]), np.ones(fin.sum(), bool)), r
        err = np.abs(lp_g[fin] - lp_o[fin])
        rel_errs.append(err / np.maximum(np.abs(lp_o[fin]), 1e-300))
        # own decision vs the oracle's: differences only at near-ties
        dec = (gs["flags"] & 1) != (o["flags"] & 1)
        u = accept_uniform(tapes[r], o)
        for s in np.flatnonzero(dec):
            nd_total += 1
            gap = abs(np.log(u[s]) - lp_o[s])
            assert gap <= tol[s] + 1e-6, (r, s, gap, lp_o[s], lp_g[s])
        assert np.array_equal(gres.counters[r], ores[r]["counters"]), r
        n = ores[r]["n"]
        assert gres.n_final[r] == n
        cap = (gres.final.shape[1] - 4) // 4
        fi = gres.final[r]
        assert np.array_equal(fi[4:4 + n].view(np.uint32), ores[r]["I"].view(np.uint32))
        assert np.array_equal(fi[4 + cap:4 + cap + n].view(np.uint32), ores[r]["X"].view(np.uint32))
        assert np.array_equal(fi[4 + 2 * cap:4 + 2 * cap + n].view(np.uint32), ores[r]["Y"].view(np.uint32))
    e = np.concatenate(rel_errs) if rel_errs else np.zeros(1)
    print(f"log PR |err|/|log PR|: median {np.median(e):.2e} p99 {np.quantile(e, 0.99):.2e} "
          f"max {e.max():.2e}; own-decision differences (near-ties): {nd_total}", flush=True)
    return nd_total


def test_replay_config1(oracle):
    """Config-1 distribution (16x16, N ~ Poisson(3) capped at 10, bg 10), 3000 + 3000 jumps."""
    rois = synth.make_rois(48, seed=101, sides=(16,), mean_n=3.0, n_cap=10)
    p = params()
    ores, tapes = run_oracle_chains(oracle, rois, p, seed=7, chain=True)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    rs = RoiSet.from_synth(rois, bg0=np.full(48, 10.0, np.float32))
    rp = make_replay(tapes, np.stack([o["steps"] for o in ores]))
    print("replay config1: 48 ROIs", flush=True)
    g = run_batch(plan, rs, mode=REPLAY, replay=rp, record_steps=True, hist_len=129)
    check_replay(rois, g, ores, tapes, p, lambda r: rois["data"][rois["offset"][r]:rois["offset"][r + 1]].sum())
    check_llr_against_float64(oracle, g, ores, rois["data"], rois["offset"], rois["side"], rois["side"], 1.3, 10.0)


def test_replay_uneven_sizes_and_background(oracle):
    """Sides 8..32, non-square ROIs, denser ROIs, initial emitters and the signal/background
    classification (IsBackg = 1, so backFore and the background prior branches run)."""
    rng = np.random.default_rng(5)
    n = 20
    xs = rng.choice([8, 12, 16, 24, 32], n)
    ys = rng.choice([8, 12, 20, 32], n)
    offs = np.zeros(n + 1, np.int64)
    offs[1:] = np.cumsum(xs * ys)
    data = np.zeros(offs[-1], np.float32)
    inits = []
    for r in range(n):
        k = rng.poisson(0.05 * xs[r] * ys[r] / 4)
        X, Y = rng.uniform(0, xs[r], k), rng.uniform(0, ys[r], k)
        I = rng.gamma(5, 200, k)
        mu = synth.roi_expected(I, X, Y, 15.0, int(xs[r]), int(ys[r]), 1.3)
        data[offs[r]:offs[r + 1]] = rng.poisson(mu)
        m = min(k, 2)
        inits.append((I[:m].astype(np.float32), X[:m].astype(np.float32) + 0.3, Y[:m].astype(np.float32) - 0.2))
    rois = {"data": data, "offset": offs, "xsize": xs, "ysize": ys}
    back = synth.gamma_pdf_table(2, 40, 10.0, 1001)
    p = params(nb=1500, nt=1500, is_backg=1, p_burnin=(0.3, 0.15, 0.15, 0.15, 0.15, 0, 0, 0.1),
               p_trials=(0.5, 0.1, 0.1, 0.1, 0.1, 0, 0, 0.1))
    ores, tapes = run_oracle_chains(oracle, rois, p, seed=11, bg0=15.0, back=back, n0=inits, grids=True, chain=True)
    plan = Plan(p, OFF_PDF, SIG_PDF, back)
    eoff = np.zeros(n, np.int64)
    n0 = np.array([len(t[0]) for t in inits], np.int32)
    eoff[1:] = np.cumsum(n0)[:-1]
    em = (np.concatenate([t[0] for t in inits]), np.concatenate([t[1] for t in inits]),
          np.concatenate([t[2] for t in inits]), eoff)
    rs = RoiSet(data, offs, xs, ys, bg0=np.full(n, 15.0, np.float32), n0=n0, em0=em)
    rp = make_replay(tapes, np.stack([o["steps"] for o in ores]))
    print("replay uneven: 20 ROIs", flush=True)
    g = run_batch(plan, rs, mode=REPLAY, replay=rp, record_steps=True, grids=True, hist_len=129)
    check_replay(rois, g, ores, tapes, p, lambda r: data[offs[r]:offs[r + 1]].sum())
    check_llr_against_float64(oracle, g, ores, data, offs, xs, ys, 1.3, 15.0, inits=inits)
    off = 0
    for r in range(n):
        sz = int(xs[r] * ys[r] * 400)
        assert np.array_equal(g.pgrid[off:off + sz], ores[r]["pgrid"]), r
        assert np.array_equal(g.pback[off:off + sz], ores[r]["pback"]), r
        off += sz


@pytest.mark.parametrize("sigma", [1.3, 2.5])
def test_replay_wide_rois_both_stencil_forms(oracle, sigma):
    """ROIs with X + Y + 2 > 64 (32 x 32, 31 x 32, 32 x 31): their stencils do not
    fit one edge per lane.  At sigma 1.3 (15 C >= 5) the device evaluates one erf
    per lane over 32-edge windows and fills the edges beyond with the saturated
    +-1; at sigma 2.5 (15 C < 5) it evaluates both axes' edges on every lane.
    Both must replay the oracle's chains exactly."""
    rng = np.random.default_rng(17)
    shapes = [(32, 32), (32, 32), (31, 32), (32, 31), (32, 32), (31, 32)]
    xs = np.array([s[0] for s in shapes])
    ys = np.array([s[1] for s in shapes])
    n = len(shapes)
    offs = np.zeros(n + 1, np.int64)
    offs[1:] = np.cumsum(xs * ys)
    data = np.zeros(offs[-1], np.float32)
    for r in range(n):
        k = rng.poisson(0.03 * xs[r] * ys[r] / 4)
        X, Y = rng.uniform(0, xs[r], k), rng.uniform(0, ys[r], k)
        mu = synth.roi_expected(rng.gamma(5, 200, k), X, Y, 12.0, int(xs[r]), int(ys[r]), sigma)
        data[offs[r]:offs[r + 1]] = rng.poisson(mu)
    rois = {"data": data, "offset": offs, "xsize": xs, "ysize": ys}
    p = params(nb=700, nt=700, psf_sigma=sigma)
    ores, tapes = run_oracle_chains(oracle, rois, p, seed=23, bg0=12.0, chain=True)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    rs = RoiSet(data, offs, xs, ys, bg0=np.full(n, 12.0, np.float32))
    rp = make_replay(tapes, np.stack([o["steps"] for o in ores]))
    print(f"replay wide ROIs, sigma {sigma}: {n} ROIs", flush=True)
    g = run_batch(plan, rs, mode=REPLAY, replay=rp, record_steps=True, hist_len=129)
    check_replay(rois, g, ores, tapes, p, lambda r: data[offs[r]:offs[r + 1]].sum())
    check_llr_against_float64(oracle, g, ores, data, offs, xs, ys, sigma, 12.0)


def test_replay_group_jumps(oracle):
    """gSplit / gMerge (P_Burnin entries 5 and 6, also reached in the trials
    phase through P_Burnin's cumulative vector) on clustered ROIs."""
    from test_oracle_gsplit import GROUP_P, clustered_rois

    rng = np.random.default_rng(8)
    rois, inits = clustered_rois(rng, 24)
    n = 24
    p = params(nb=1500, nt=1500, **GROUP_P)
    ores, tapes = run_oracle_chains(oracle, rois, p, seed=17, n0=inits)
    kinds = np.concatenate([o["steps"]["kind"][(o["steps"]["flags"] & 2) != 0] for o in ores])
    assert np.sum(kinds == 5) > 200 and np.sum(kinds == 6) > 200
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    n0 = np.array([len(t[0]) for t in inits], np.int32)
    eoff = np.zeros(n, np.int64)
    eoff[1:] = np.cumsum(n0)[:-1]
    em = (np.concatenate([t[0] for t in inits]), np.concatenate([t[1] for t in inits]),
          np.concatenate([t[2] for t in inits]), eoff)
    rs = RoiSet(rois["data"], rois["offset"], rois["xsize"], rois["ysize"], bg0=np.full(n, 10.0, np.float32),
                n0=n0, em0=em)
    rp = make_replay(tapes, np.stack([o["steps"] for o in ores]))
    print("replay group jumps: 24 ROIs", flush=True)
    g = run_batch(plan, rs, mode=REPLAY, replay=rp, record_steps=True, hist_len=129)
    data = rois["data"]
    offs = rois["offset"]
    check_replay(rois, g, ores, tapes, p, lambda r: data[offs[r]:offs[r + 1]].sum())


def test_free_running_matches_oracle_philox(oracle):
    """Same Philox stream on both sides: identical trajectories up to the first
    near-tie; most chains never meet one."""
    rois = synth.make_rois(32, seed=202, sides=(16,), mean_n=3.0, n_cap=10)
    p = params(nb=2000, nt=2000)
    ores, _ = run_oracle_chains(oracle, rois, p, seed=99, record=False)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    rs = RoiSet.from_synth(rois, bg0=np.full(32, 10.0, np.float32))
    print("free-running: 32 ROIs", flush=True)
    g = run_batch(plan, rs, seed=99, mode=PLAIN, record_steps=True, hist_len=129)
    full = 0
    for r in range(32):
        o, gs = ores[r]["steps"], g.steps[r]
        same = (gs["n"] == o["n"]) & (gs["draws"] == o["draws"]) & (gs["kind"] == o["kind"])
        if same.all():
            full += 1
            assert np.array_equal(g.counters[r], ores[r]["counters"])
        else:
            s = int(np.argmin(same))
            # the first difference is a near-tie accept decision or a birth-pixel tie
            assert (gs["flags"][s] & 1) != (o["flags"][s] & 1) or gs["pix"][s] != o["pix"][s], (r, s)
    assert full >= 24, full


def test_status_and_counters_shape():
    rois = synth.make_rois(3, seed=3, sides=(8,), mean_n=1.0)
    p = params(nb=50, nt=50)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    g = run_batch(plan, RoiSet.from_synth(rois), seed=1, hist_len=16, trace_len=50)
    assert np.all(g.status == 0)
    assert np.all(g.hist.sum(1) == 50)
    assert np.all(g.counters[:, :5] >= 1)
    assert g.trace_n.shape == (3, 50)


def test_branch_free_erf_accuracy():
    """The stencil's erf against float64 erf over the range the stencils use."""
    from scipy.special import erf

    from bamf_amd import _hip
    from bamf_amd._lib import check, lib

    x = np.concatenate([np.linspace(-30, 30, 200001), [0.0, -0.0, 1e-30, -1e-30]]).astype(np.float32)
    dx = _hip.DeviceBuffer.from_array(x)
    dy = _hip.DeviceBuffer(x.nbytes)
    check(lib().bamf_math_erf(dx.vp, dy.vp, len(x), None), "bamf_math_erf")
    _hip.synchronize()
    y = dy.download(np.float32)
    err = np.abs(y.astype(np.float64) - erf(x.astype(np.float64)))
    assert err.max() < 1.5e-7, err.max()
    # relative precision where erf is small (the series' range)
    small = np.abs(x) < 0.5
    rel = err[small] / np.maximum(np.abs(erf(x[small].astype(np.float64))), 1e-30)
    assert rel.max() < 5e-7, rel.max()
    # saturation: from |x| = 5 on the value is exactly -1 or +1 (the 32 x 32 stencil
    # fills the edges 15 px or more from an emitter with these values unevaluated
    # when 15 C >= 5); checked on every float32 in [5, 40]
    xs = np.arange(np.float32(5.0).view(np.int32), np.float32(40.0).view(np.int32) + 1, dtype=np.int32)
    xs = xs.view(np.float32)
    xs = np.concatenate([xs, -xs])
    ds = _hip.DeviceBuffer.from_array(xs)
    dt = _hip.DeviceBuffer(xs.nbytes)
    check(lib().bamf_math_erf(ds.vp, dt.vp, len(xs), None), "bamf_math_erf")
    _hip.synchronize()
    ys = dt.download(np.float32)
    assert np.array_equal(ys, np.sign(xs)), xs[ys != np.sign(xs)][:5]


def test_birth_death_stationary_law_gpu():
    """The GPU chains' N under an exactly flat likelihood follows the
    birth-death law fixed by the reference's birth and death ratios
    (derivation in test_oracle_rjmcmc.py)."""
    from test_oracle_rjmcmc import BD_BG, BD_SIDE, birth_death_setup, check_birth_death

    n_roi, nt = 64, 20000
    p, data, sig, off, back, law = birth_death_setup(n_trials=nt)
    plan = Plan(p, off, sig, back)
    rs = RoiSet(np.tile(data, n_roi), np.arange(n_roi + 1) * data.size, [BD_SIDE] * n_roi, [BD_SIDE] * n_roi,
                bg0=np.full(n_roi, BD_BG, np.float32))
    g = run_batch(plan, rs, seed=11, trace_len=nt, trace_thin=1, hist_len=129)
    assert np.all(g.status == 0)
    check_birth_death(g.trace_n.astype(np.int64).ravel(), law)


def test_learned_birth_keeps_birth_death_law():
    """Learned births with a non-uniform pixel map and an intensity mixture:
    the birth ratio's PixProb = P q(pixel) and prior(I)/q(I) factors restore
    the flat-likelihood N law of the reference's birth/death pair."""
    from bamf_amd import _hip
    from bamf_amd.rjmcmc import LEARNED
    from test_oracle_rjmcmc import BD_BG, BD_SIDE, birth_death_setup, check_birth_death

    n_roi, nt = 64, 20000
    p, data, sig, off, back, law = birth_death_setup(n_trials=nt)
    plan = Plan(p, off, sig, back)
    P = BD_SIDE * BD_SIDE
    rng = np.random.default_rng(5)
    q = rng.uniform(0.8, 1.25, (n_roi, P))
    q = (q / q.sum(axis=1, keepdims=True)).astype(np.float32)
    mu = np.full((n_roi, P), np.log(5e-7), np.float32)  # log-uniform part inside the prior's support
    dq, dmu = _hip.DeviceBuffer.from_array(q.ravel()), _hip.DeviceBuffer.from_array(mu.ravel())
    plan.set_proposal(dq, dmu, alpha=0.8, halfwidth=0.5, beta=0.0)
    rs = RoiSet(np.tile(data, n_roi), np.arange(n_roi + 1) * data.size, [BD_SIDE] * n_roi, [BD_SIDE] * n_roi,
                bg0=np.full(n_roi, BD_BG, np.float32))
    g = run_batch(plan, rs, seed=13, mode=LEARNED, trace_len=nt, trace_thin=1, hist_len=129)
    assert np.all(g.status == 0)
    check_birth_death(g.trace_n.astype(np.int64).ravel(), law)


def test_split_merge_law_depends_on_the_split_proposal():
    """The reference's split/merge pair is not reversible in the textbook
    sense (a merge may pick either child as jj, and the ratios use the offsets
    seen from that child), so its flat-likelihood stationary N law depends on
    the split proposal's density: here on Split_std itself, and equally on the
    learned split's map-drawn offsets (bamf_rjmcmc_set_split_proposal).  The
    learned split therefore cannot keep the reference's stationary law and is
    off by default.  Two runs with the same proposal agree (the control)."""
    from bamf_amd import _hip
    from bamf_amd.rjmcmc import LEARNED, make_params
    from test_oracle_rjmcmc import BD_BG, BD_BND, BD_MEANN, BD_SIDE, birth_death_setup

    n_roi, nt = 128, 20000
    _, data, sig, off, back, _ = birth_death_setup(n_trials=nt)
    mix = (0, 0.25, 0.25, 0.25, 0.25, 0, 0, 0)
    P = BD_SIDE * BD_SIDE
    rng = np.random.default_rng(21)
    q = rng.uniform(0.2, 3.0, (n_roi, P)).astype(np.float32)
    mu = np.full((n_roi, P), np.log(5e-7), np.float32)
    dq, dmu = _hip.DeviceBuffer.from_array(q.ravel()), _hip.DeviceBuffer.from_array(mu.ravel())
    rs = RoiSet(np.tile(data, n_roi), np.arange(n_roi + 1) * data.size, [BD_SIDE] * n_roi, [BD_SIDE] * n_roi,
                bg0=np.full(n_roi, BD_BG, np.float32))

    def run(split_std, seed, share=0.0):
        p = make_params(n_burnin=2000, n_trials=nt, rho=BD_MEANN / BD_SIDE ** 2, dx=1e-7, bnd_out=BD_BND,
                        p_burnin=mix, p_trials=mix, split_std=split_std)
        plan = Plan(p, off, sig, back)
        mode = PLAIN
        if share > 0:
            plan.set_proposal(dq, dmu, alpha=1.0, halfwidth=0.5, beta=0.0)
            plan.set_split_proposal(share, 2.6)
            mode = LEARNED
        g = run_batch(plan, rs, seed=seed, mode=mode, trace_len=nt, trace_thin=1, hist_len=129)
        assert np.all(g.status == 0)
        return g.trace_n.astype(np.int64)

    def compare(a, b):
        dm = b.mean(axis=1) - a.mean(axis=1)  # independent chains per ROI
        ha = np.bincount(a.ravel(), minlength=129)[:129] / a.size
        hb = np.bincount(b.ravel(), minlength=129)[:129] / b.size
        return dm.mean(), dm.std(ddof=1) / np.sqrt(n_roi), 0.5 * np.abs(ha - hb).sum()

    base = run(0.5, 31)
    same = compare(base, run(0.5, 32))
    wide = compare(base, run(1.2, 33))
    learned = compare(base, run(0.5, 34, share=0.9))
    print(f"split/merge N law vs Split_std 0.5 (mean N {base.mean():.3f}): same proposal {same}, Split_std 1.2 {wide}, "
          f"learned split {learned}", flush=True)
    assert abs(same[0]) < 4 * same[1] + 0.02 and same[2] < 0.03
    assert wide[0] > 6 * wide[1] and learned[0] > 6 * learned[1]


def test_trained_proposal_gain():
    """What the trained network gains: on the same ROIs, the photon trace of
    learned chains (independence intensity proposals from the network's
    log-intensity map) mixes faster than the reference's random walk (about
    1.5-2x the median ESS with 600 training steps)."""
    from bamf_amd import proposal
    from bamf_amd.ess import ess_batch
    from bamf_amd.rjmcmc import LEARNED

    prm, _ = proposal.train(600, batch=256, lr=3e-3, seed=2)
    rois = synth.make_rois(256, seed=31, sides=(16,), density=(0.5, 5.0), bg=(5.0, 30.0))
    p = params(nb=2000, nt=4000, rho=0.0275)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    rs = RoiSet.from_synth(rois, bg0=np.full(256, 10.0, np.float32))
    dq, dmu = proposal.proposal_maps(prm, rs)
    plan.set_proposal(dq, dmu, alpha=0.2, halfwidth=0.35, beta=1.0)
    res = {}
    for name, mode in (("plain", PLAIN), ("learned", LEARNED)):
        g = run_batch(plan, rs, seed=17, mode=mode, trace_len=800, trace_thin=5)
        assert np.all(g.status == 0)
        c = g.counters.sum(axis=0)
        en = ess_batch(g.trace_n.astype(np.float64))
        ep = ess_batch(g.trace_photons.astype(np.float64))
        res[name] = dict(birth_acc=c[5 + 3] / c[3], move_acc=c[5] / c[0], ess_n=float(np.median(en)),
                         ess_ph=float(np.median(ep)), ess_min=float(np.median(np.minimum(en, ep))))
    print("plain vs learned:", res, flush=True)
    # (training uses fp32 atomics, so the network and this ratio vary a little between runs)
    assert res["learned"]["ess_ph"] > 1.3 * res["plain"]["ess_ph"]


def test_learned_intensity_kernel_keeps_prior():
    """The independence intensity kernel of learned moves under an exactly flat
    likelihood: an emitter's intensity must be distributed as the prior
    (uniform on (0, 1e-6)) restricted to the log-uniform window it moves in.
    Dropping the kernel's Jacobian I'/I would give the log-uniform law instead."""
    from scipy import stats

    from bamf_amd import _hip
    from bamf_amd.rjmcmc import LEARNED, make_params
    from test_oracle_rjmcmc import BD_SIDE

    # background 1e-5 (inside the offset table, so the background factor of
    # the move ratio is 1) and intensities near 5e-14, far below ulp(1e-5) / 2
    n_roi, side, bg = 1024, BD_SIDE, 1e-5
    p = make_params(n_burnin=1000, n_trials=20000, rho=1.0 / side ** 2, dx=1e-7,
                    p_burnin=(1, 0, 0, 0, 0, 0, 0, 0), p_trials=(1, 0, 0, 0, 0, 0, 0, 0))
    plan = Plan(p, np.ones(201, np.float32), np.ones(11, np.float32), np.zeros(11, np.float32))
    P = side * side
    mu0, w = np.log(5e-14), 0.5
    dq = _hip.DeviceBuffer.from_array(np.full(n_roi * P, 1.0 / P, np.float32))
    dmu = _hip.DeviceBuffer.from_array(np.full(n_roi * P, mu0, np.float32))
    plan.set_proposal(dq, dmu, alpha=1.0, halfwidth=w, beta=0.5)
    data = np.full(n_roi * P, bg, np.float32)
    em = (np.full(n_roi, 5e-14, np.float32), np.full(n_roi, 4.0, np.float32), np.full(n_roi, 4.0, np.float32),
          np.arange(n_roi, dtype=np.int64))
    rs = RoiSet(data, np.arange(n_roi + 1) * P, [side] * n_roi, [side] * n_roi,
                bg0=np.full(n_roi, bg, np.float32), n0=np.ones(n_roi, np.int32), em0=em)
    g = run_batch(plan, rs, seed=21, mode=LEARNED, hist_len=8)
    assert np.all(g.status == 0) and np.all(g.n_final == 1)
    I = g.final[:, 4].astype(np.float64)
    lo, hi = np.exp(mu0 - w), np.exp(mu0 + w)
    assert np.all((I >= lo * 0.999) & (I <= hi * 1.001))
    ks = stats.kstest(I, stats.uniform(loc=lo, scale=hi - lo).cdf)
    ks_log = stats.kstest(np.log(I), stats.uniform(loc=mu0 - w, scale=2 * w).cdf)
    print("KS uniform p", ks.pvalue, "KS log-uniform p", ks_log.pvalue, flush=True)
    assert ks.pvalue > 0.01
    assert ks_log.pvalue < 1e-4


def test_chain_records_match_oracle(oracle):
    """Chain / ProposedChain records (RJMCMCChain::addEmitters) of replayed
    chains: heads, background terms and emitter entries equal the oracle's;
    LLR and PR agree to single-precision rounding."""
    rois = synth.make_rois(10, seed=41, sides=(8, 16, 24), mean_n=3.0)
    p = params(nb=600, nt=600, drift_x=0.25, drift_y=-0.5)
    op = oracle.McmcParams.from_buffer_copy(p)
    n = len(rois["side"])
    ores, tapes = [], []
    for r in range(n):
        o = rois["offset"]
        s = int(rois["side"][r])
        res = oracle.rjmcmc(rois["data"][o[r]:o[r + 1]], s, s, op, OFF_PDF, SIG_PDF, BACK_PDF, bg0=10.0, seed=7,
                            roi=r, record=True, chain=True, rec_cap=1300, em_cap=20000)
        assert res["status"] == 0
        ores.append(res)
        tapes.append(res["tape"])
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    rs = RoiSet.from_synth(rois, bg0=np.full(n, 10.0, np.float32))
    rp = make_replay(tapes, np.stack([o["steps"] for o in ores]))
    g = run_batch(plan, rs, mode=REPLAY, replay=rp, record_steps=True, chain_caps=(1300, 20000))
    assert np.all(g.status == 0)
    ch = g.chain
    for r in range(n):
        oc = ores[r]["chain"]
        nrec = len(oc["acc_head"])
        assert ch["count"][r, 0] == nrec
        assert ch["count"][r, 1] == len(oc["acc_em"]) and ch["count"][r, 2] == len(oc["prop_em"])
        for side, head, val, em in (("acc", "acc_head", "acc_val", "acc_em"), ("prop", "prop_head", "prop_val", "prop_em")):
            gr = ch[f"{side}_rec"][r, :nrec]
            oh, ov = oc[head], oc[val]
            assert np.array_equal(gr["n"], oh[:, 0]), (r, side)
            assert np.array_equal(gr["jump_type"], oh[:, 1]), (r, side)
            assert np.array_equal(gr["accepted"], oh[:, 2]), (r, side)
            assert np.array_equal(gr["em_index"], oh[:, 3]), (r, side)
            for k, f in ((2, "bg"), (3, "abg"), (4, "bbg")):
                assert np.array_equal(gr[f].view(np.uint32), ov[:, k].view(np.uint32)), (r, side, f)
            for k, f in ((0, "llr"), (1, "pr")):
                a, b = gr[f].astype(np.float64), ov[:, k].astype(np.float64)
                fin = np.isfinite(b) & (np.abs(b) < 1e30) & (np.abs(b) > 1e-30)
                assert np.all(np.abs(a[fin] - b[fin]) <= 1e-3 * np.abs(b[fin])), (r, side, f)
            ne = len(oc[em])
            assert np.array_equal(ch[f"{side}_em"][r, :ne].view(np.uint32), oc[em].view(np.uint32)), (r, side)


def test_conditional_fit_kernel_keeps_posterior():
    """The conditional-fit intensity proposal (learned mode, every single-emitter
    move, half of all moves): with moves only, 1024 independent chains on the
    same one-emitter ROI end in posterior draws; their intensities and
    backgrounds must have the plain random-walk chains' distribution (two-sample
    KS).  With one emitter the fit kernel is the only move that changes the
    intensity; without the reverse fit density in its ratio it would sample
    the product of the posterior and its proposal (sd 1.3 times the fit's),
    an intensity sd about 21 % smaller, beyond the 15 % bound below."""
    from scipy import stats

    from bamf_amd import _hip
    from bamf_amd.rjmcmc import LEARNED

    n_roi, side = 1024, 12
    rng = np.random.default_rng(77)
    mu_img = synth.roi_expected(np.array([1500.0]), np.array([6.3]), np.array([5.7]), 10.0, side, side, 1.3)
    data = rng.poisson(mu_img).astype(np.float32).ravel()
    p = params(nb=3000, nt=3000, p_burnin=(1, 0, 0, 0, 0, 0, 0, 0), p_trials=(1, 0, 0, 0, 0, 0, 0, 0))
    P = side * side
    em = (np.full(n_roi, 1500.0, np.float32), np.full(n_roi, 6.3, np.float32), np.full(n_roi, 5.7, np.float32),
          np.arange(n_roi, dtype=np.int64))
    rs = RoiSet(np.tile(data, n_roi), np.arange(n_roi + 1) * P, [side] * n_roi, [side] * n_roi,
                bg0=np.full(n_roi, 10.0, np.float32), n0=np.ones(n_roi, np.int32), em0=em)
    finals = {}
    for name in ("plain", "fit"):
        plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
        mode = PLAIN
        if name == "fit":
            dq = _hip.DeviceBuffer.from_array(np.full(n_roi * P, 1.0 / P, np.float32))
            dmu = _hip.DeviceBuffer.from_array(np.full(n_roi * P, np.log(1500.0), np.float32))
            plan.set_proposal(dq, dmu, alpha=0.2, halfwidth=0.35, beta=1.0)
            plan.set_intensity_proposal("fit", 1.3, 0.5)
            mode = LEARNED
        g = run_batch(plan, rs, seed=5, mode=mode, hist_len=8)
        assert np.all(g.status == 0) and np.all(g.n_final == 1)
        finals[name] = (g.final[:, 4].astype(np.float64), g.final[:, 1].astype(np.float64))
    (ip, bp), (i_f, bf) = finals["plain"], finals["fit"]
    ks_i = stats.ks_2samp(ip, i_f)
    ks_b = stats.ks_2samp(bp, bf)
    print(f"fit kernel: I mean {ip.mean():.1f} / {i_f.mean():.1f}, sd {ip.std():.1f} / {i_f.std():.1f}, "
          f"KS p {ks_i.pvalue:.3f}; BG KS p {ks_b.pvalue:.3f}", flush=True)
    assert ks_i.pvalue > 0.01 and ks_b.pvalue > 0.01
    assert abs(i_f.std() / ip.std() - 1) < 0.15
