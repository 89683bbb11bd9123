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
        w = oracle.philox([step >> 2, j, 0, 0], [9, 42])
        assert oracle.lib().orc_chain_raw(42, 9, step, j) == w[step & 3] >> 1


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


def test_philox_known_answers(oracle):
    """Philox4x32-10 against the known-answer vectors published with Random123
    (Salmon et al., SC'11; kat_vectors, philox4x32 R=10)."""
    kat = [([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
           ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
           ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1])]
    for ctr, key, want in kat:
        assert [int(v) for v in oracle.philox(ctr, key)] == want


# ---- birth/death balance -------------------------------------------------
# With emitter intensities far below one ulp of the background, M' == M
# exactly, so LLR = 1 and N performs a birth-death walk whose stationary law
# follows from the jump ratios alone (LibRJMCMC.cpp:1302 and :1360):
#   birth from N:   PR_b = P3 / (P4 * (X+Bnd)(Y+Bnd) * PixProb) * MeanN / (N + 1)
#   death from N+1: PR_d = (P4 / P3) * (N + 1) / MeanN
# with P = PV_Trials (the cumulative vector) after burn-in.  Detailed balance
# of the walk, with each kind proposed with probability 1/2, gives
#   pi(N+1) / pi(N) = min(1, PR_b(N)) / min(1, PR_d(N+1)).
BD_SIDE, BD_BG, BD_MEANN, BD_BND = 8, 100.0, 100.0, 2.0
BD_AREA = (BD_SIDE + BD_BND) ** 2


def birth_death_law(mean_n, area, p3, p4, nmax=128, n_plus=1.0):
    pi = np.zeros(nmax + 1)
    pi[0] = 1.0
    for n in range(nmax):
        prb = p3 / (p4 * area) * mean_n / (n + n_plus)
        prd = (p4 / p3) * (n + 1) / mean_n
        pi[n + 1] = pi[n] * min(1.0, prb) / min(1.0, prd)
    return pi / pi.sum()


def birth_death_setup(n_trials=300000):
    from bamf_amd.rjmcmc import make_params

    dx = 1e-7  # intensities in (0, 1e-6): far below ulp(100) / 2
    sig = np.full(11, 1e6, np.float32)  # the density of U(0, 1e-6) (its table integrates to 1)
    off = np.ones(11, np.float32)
    back = np.zeros(11, np.float32)
    p = make_params(n_burnin=2000, n_trials=n_trials, rho=BD_MEANN / BD_SIDE ** 2, dx=dx, bnd_out=BD_BND,
                    p_burnin=(0, 0, 0, 0.5, 0.5, 0, 0, 0), p_trials=(0, 0, 0, 0.5, 0.5, 0, 0, 0))
    data = np.full(BD_SIDE * BD_SIDE, BD_BG, np.float32)
    law = birth_death_law(BD_MEANN, BD_AREA, 0.5, 1.0)  # PV_Trials[3], PV_Trials[4]
    return p, data, sig, off, back, law


def check_birth_death(n_trace, law):
    """Mean and total variation of the post-burn-in N against the law; the
    tolerances are about three standard errors of these chain lengths."""
    emp = np.bincount(n_trace, minlength=129)[:129] / n_trace.size
    mean_law = float((np.arange(129) * law).sum())
    assert abs(n_trace.mean() - mean_law) < 0.1, (n_trace.mean(), mean_law)
    assert 0.5 * np.abs(emp - law).sum() < 0.03
    return mean_law


def test_birth_death_stationary_law(oracle):
    p, data, sig, off, back, law = birth_death_setup()
    op = oracle.McmcParams.from_buffer_copy(p)
    trace = []
    for seed in (1, 2):
        r = oracle.rjmcmc(data, BD_SIDE, BD_SIDE, op, off, sig, back, bg0=BD_BG, seed=seed, record=True)
        assert r["status"] == 0
        st = r["steps"][p.n_burnin:]
        assert np.all(st["llr"][(st["flags"] & 2) != 0] == 1.0)  # the likelihood is exactly flat
        trace.append(st["n"])
        # acceptance per N against min(1, PR) of the two ratios
        kind, acc = st["kind"], (st["flags"] & 1) != 0
        nprev = np.concatenate([[st["n"][0]], st["n"][:-1]])
        for n in range(2, 8):
            for k, pr in ((3, 0.5 / BD_AREA * BD_MEANN / (n + 1)), (4, 2.0 * n / BD_MEANN)):
                sel = (kind == k) & (nprev == n)
                want = min(1.0, pr)
                assert abs(acc[sel].mean() - want) < 4 * np.sqrt(want * (1 - want) / sel.sum()) + 1e-3
    mean_law = check_birth_death(np.concatenate(trace), law)
    # the check separates the law from plausible slips in the ratios
    for wrong in (birth_death_law(BD_MEANN, BD_AREA, 0.5, 0.5),            # raw P instead of PV_Trials
                  birth_death_law(BD_MEANN, BD_AREA, 0.5, 1.0, n_plus=2.0),  # MeanN / (N + 2)
                  birth_death_law(BD_MEANN, 1.0, 0.5, 1.0),                  # area term dropped
                  birth_death_law(BD_MEANN * 1.1, BD_AREA, 0.5, 1.0)):       # MeanN off by 10 %
        assert abs(float((np.arange(129) * wrong).sum()) - mean_law) > 0.3
