"""Pins the oracle chain (oracle/rjmcmc.c) without the reference binary.

  - urand's mapping of a raw 31-bit integer, including U == 1.0 at the top;
  - Stat tables against the closed-form CDF of a uniform pdf, findPDF at grid
    points and in between, genRand as the inverse CDF;
  - the model: a one-pixel ROI in closed form, sum over a large ROI equals the
    photon count, the N = 0 model is the background plane plus sCMOS;
  - the draw counts of every jump kind (each randn takes 12 uniforms) and the
    accept rule U < PR checked against the recorded uniforms;
  - counters against the step records; replay of a recorded tape (Philox and
    libc rand() sources) reproduces the chain exactly;
  - fixed-dimension sampling (moves only) of one bright emitter: the posterior
    mean and spread of x, y agree with the Poisson MLE and its CRLB.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

from bamf_amd import synth
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params, run_oracle_chains


def test_urand_mapping(oracle):
    L = oracle.lib()
    assert L.orc_urand_from_raw(0) == 0.0
    assert L.orc_urand_from_raw(1 << 30) == 0.5
    assert L.orc_urand_from_raw((1 << 31) - 1) == 1.0  # rounds to 2^31 in single precision
    assert L.orc_urand_from_raw((1 << 31) - 65) < 1.0


def test_chain_raw_is_philox_word(oracle):
    for step, j in [(0, 0), (5, 3), (7, 6), (123, 1023)]:
        w = oracle.philox([step, j >> 2, 0, 0], [9, 42])
        assert oracle.lib().orc_chain_raw(42, 9, step, j) == w[j & 3] >> 1


def test_stat_tables_uniform_pdf(oracle):
    dx, lo, hi = 10.0, 100.0, 300.0
    pdf = synth.uniform_pdf_table(lo, hi, dx, 60)
    x, cdf, ux, uc = oracle.stat_tables(pdf, dx)
    np.testing.assert_array_equal(x, np.arange(60, dtype=np.float32) * np.float32(dx))
    # trapezoid rule of a step function: exact except the two ramps
    want = np.clip((x - lo) / (hi - lo), 0, 1)
    assert np.max(np.abs(cdf - want)) <= 0.5 * dx / (hi - lo) + 1e-6
    assert uc[0] == 0 and uc[-1] == 1 and np.all(np.diff(uc) > 0)
    L = oracle.lib()
    for v in [150.0, 200.0, 255.0]:
        assert abs(L.orc_stat_findpdf(pdf.ctypes.data, 60, dx, v) - 1 / (hi - lo)) < 1e-9
    assert L.orc_stat_findpdf(pdf.ctypes.data, 60, dx, 0.0) == 0.0
    assert L.orc_stat_findpdf(pdf.ctypes.data, 60, dx, 590.0) == 0.0
    assert L.orc_stat_findpdf(pdf.ctypes.data, 60, dx, float("nan")) == 0.0
    # genRand inverts the CDF (linear pieces)
    for u in [0.1, 0.37, 0.5, 0.9]:
        v = L.orc_stat_genrand(ux.ctypes.data, uc.ctypes.data, len(ux), u)
        assert abs(v - (lo + u * (hi - lo))) < dx
    assert L.orc_stat_genrand(ux.ctypes.data, uc.ctypes.data, len(ux), 1.0) == ux[-1]
    assert L.orc_stat_genrand(ux.ctypes.data, uc.ctypes.data, len(ux), 0.0) == 0.0


def test_model_closed_forms(oracle):
    s = 1.3
    c = 1 / (s * math.sqrt(2))
    m = oracle.model([1000.0], [0.4], [0.7], 5.0, 0.0, 0.0, 1, 1, s)
    want = 5.0 + 1000.0 * 0.5 * (erf((1 - 0.4) * c) - erf(-0.4 * c)) * 0.5 * (erf((1 - 0.7) * c) - erf(-0.7 * c))
    assert abs(m[0] - want) < 1e-3
    m = oracle.model([1234.0], [20.3], [19.6], 0.0, 0.0, 0.0, 40, 40, s)
    assert abs(m.sum() - 1234.0) < 0.05
    sc = np.arange(12, dtype=np.float32)
    m = oracle.model([], [], [], 3.0, 0.5, 0.25, 3, 4, s, scmos=sc)
    xx, yy = np.meshgrid(np.arange(3), np.arange(4), indexing="ij")
    np.testing.assert_allclose(m, (3.0 + 0.5 * xx + 0.25 * yy).ravel() + sc, rtol=1e-6)


# draws of each jump kind when it reaches the accept test (IsBackg = 0):
#   move single / group walk of one / background: dispatch, Ua, Ub, 3 randn, accept
#   split: dispatch, Ua, u1, 2 randn, accept;  merge: dispatch, Ua, closest, accept
#   birth: dispatch, pixel, y, x, genRand, accept;  death: dispatch, Ua, accept
def test_draw_counts_and_accept_rule(oracle):
    rois = synth.make_rois(12, seed=21, sides=(16,), mean_n=3.0, n_cap=10)
    p = params(nb=1500, nt=1500)
    res, tapes = run_oracle_chains(oracle, rois, p, seed=5)
    for r, t in zip(res, tapes):
        st = r["steps"]
        assert r["status"] == 0
        assert st["draws"].sum() == len(t)
        rec = (st["flags"] & 2) != 0
        failed = (st["flags"] & 8) != 0
        dec = rec & ~failed
        jt, k, d = st["jump_type"], st["kind"], st["draws"]
        assert np.all(d[dec & (k == 4)] == 3)
        assert np.all(d[dec & (k == 2)] == 4)
        assert np.all(d[dec & (k == 1)] == 28)
        assert np.all(d[dec & (k == 3)] == 6)
        assert np.all(d[dec & ((jt == 1) | (jt == 2))] == 40)
        assert np.all(d[failed] == d[failed] // 1)  # present
        assert np.all(d[failed & (k == 1)] == 27)  # no accept draw after a failed bounds test
        # accept rule: U < PR with U the last draw of the step
        ends = np.cumsum(d)
        u = t[ends[dec] - 1].astype(np.float64)
        uf = (u / 2147483648.0).astype(np.float32)
        acc = (st["flags"][dec] & 1) != 0
        assert np.array_equal(acc, uf < st["pr"][dec])
        # counters agree with the records (they start at 1 / 0)
        c = r["counters"]
        for kind in range(5):
            assert c[kind] == 1 + np.sum(k == kind)
            assert c[5 + kind] == np.sum((k == kind) & ((st["flags"] & 1) != 0))
        # state bookkeeping: N changes only on accepted split/merge/birth/death
        dn = np.diff(np.concatenate([[0], st["n"]]))
        assert np.all(dn[(k == 0) | ~((st["flags"] & 1) != 0)] == 0)


@pytest.mark.parametrize("src", ["philox", "libc"])
def test_tape_replay_reproduces_chain(oracle, src):
    rois = synth.make_rois(3, seed=23, sides=(12,), mean_n=2.0)
    p = params(nb=500, nt=500)
    kind = oracle.SRC_PHILOX if src == "philox" else oracle.SRC_LIBC
    a, tapes = run_oracle_chains(oracle, rois, p, seed=77, src=kind)
    b, _ = [], None
    op = oracle.McmcParams.from_buffer_copy(p)
    for r in range(3):
        o = rois["offset"]
        rr = oracle.rjmcmc(rois["data"][o[r]:o[r + 1]], 12, 12, op, OFF_PDF, SIG_PDF, BACK_PDF, bg0=10.0,
                           src_kind=oracle.SRC_TAPE, tape=tapes[r])
        assert rr["status"] == 0
        assert rr["steps"].tobytes() == a[r]["steps"].tobytes()  # NaN log_pr entries compare bitwise


def test_fixed_dimension_posterior_matches_mle(oracle):
    """Moves only (P = [1 0 0 0 0 0 0 0], as @RJ/mcmc.m:103 uses), one bright emitter
    started at the truth: the chain's posterior of (x, y) centres on the Poisson MLE
    with the spread its CRLB predicts."""
    rng = np.random.default_rng(8)
    side, x0, y0, I0, bg = 9, 4.3, 4.6, 3000.0, 10.0
    mu = synth.roi_expected([I0], [x0], [y0], bg, side, side, 1.3)
    data = rng.poisson(mu).astype(np.float32)
    p = params(nb=2000, nt=30000, p_burnin=(1, 0, 0, 0, 0, 0, 0, 0), p_trials=(1, 0, 0, 0, 0, 0, 0, 0))
    op = oracle.McmcParams.from_buffer_copy(p)
    sig = synth.gamma_pdf_table(5, 800, 10.0, 2001)
    off = synth.uniform_pdf_table(0.0, 100.0, 10.0, 2001)
    r = oracle.rjmcmc(data, side, side, op, off, sig, np.zeros(2001, np.float32), bg0=bg, seed=3, n0=1,
                      i0=np.array([I0], np.float32), x0=np.array([x0], np.float32), y0=np.array([y0], np.float32),
                      chain=True, rec_cap=32000, em_cap=64000)
    ch = r["chain"]
    em = ch["acc_em"]
    heads = ch["acc_head"]
    post = heads[3000:]  # past the burn-in records
    xs = em[post[:, 3], 0]
    ys = em[post[:, 3], 1]
    # MLE and CRLB of the same box in the MLE oracle's convention (x along the contiguous index)
    box = data.reshape(side, side)  # [x, y] -> the MLE box wants [jj, ii] with ii contiguous = y here
    p_mle, c_mle, _ = oracle.mle_xynb(box[None], 1.3, 40)
    mle_y, mle_x = p_mle[0, 0], p_mle[1, 0]  # theta[0] runs along ROI y (contiguous), theta[1] along x
    se_y, se_x = math.sqrt(c_mle[0, 0]), math.sqrt(c_mle[1, 0])
    ess_guess = 30
    assert abs(xs.mean() - (mle_x + 0.0)) < 4 * se_x / math.sqrt(ess_guess) + 0.02
    assert abs(ys.mean() - (mle_y + 0.0)) < 4 * se_y / math.sqrt(ess_guess) + 0.02
    assert 0.6 < xs.std() / se_x < 1.5 and 0.6 < ys.std() / se_y < 1.5
