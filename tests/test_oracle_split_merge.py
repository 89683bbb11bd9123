"""Pins the oracle's split, merge, gSplit, gMerge and backFore
(oracle/rjmcmc.c; LibRJMCMC.cpp:972-1255, 1389-1627) by checks that do not
restate their ratio formulas:

  - split's proposal map has the Jacobian determinant I / (1 - u1)^2 that both
    ratios use, measured by central differences of the oracle's own map;
    merge's map inverts it and returns the split's u1, u2, u3;
  - reversibility: one split driven by a tape of chosen draws, then the merge
    of the two emitters it made, from the split's proposed state: the two
    log ratios cancel to float rounding (LibRJMCMC.cpp:1030-1070 against
    :1201-1217), in the burn-in phase (raw P) and in the trials phase
    (cumulative PV_Trials), with P_split != P_merge so that a ratio inverted
    on one side shows;
  - decomposition: split's and merge's log ratios equal an assembly from
    independent pieces: the Poisson log-likelihood ratio in float64, the
    intensity prior from the Stat table, the Poisson prior on N (scipy), the
    numerical Jacobian, and the normal density of u2, u3 (scipy), plus the
    reference's own constants, each cited (the position prior 1/Area with
    Area = (Xsize + Bnd)(Ysize + Bnd), :1058-1059, and the move-probability
    factor P[1]/P[2], :1066);
  - group pair: gSplit then gMerge of the same group gives a ratio product of
    (MeanN/(N+1))^2, not 1: gMerge divides by PrM = NEm/MeanN (:1562), the
    inverse of merge's MeanN/NEm (:1211), so its prior on N enters inverted.
    The oracle keeps that quirk, since the chain must follow the reference;
  - backFore: with only backFore steps, each emitter's Signal label is a
    two-state Metropolis chain whose stationary law is
    P(Signal = 1) = S / (S + B), S and B the signal and background pdfs at its
    intensity (:1596-1621);
  - planted mistakes in the oracle (a Jacobian power, an inverted P ratio, N
    in place of N + 1, a dropped factor, swapped pdfs) each fail one of these
    checks.
"""
import math

import numpy as np
import pytest
from scipy import stats

from bamf_amd import synth
from rj_common import OFF_PDF, SIG_PDF, params

SIDE = 16
RHO = 3.0 / 256  # MeanN = 3 on 16 x 16
BURN_P = (0.3, 0.25, 0.15, 0.15, 0.15, 0.0, 0.0, 0.0)   # cumulative [.3, .55, .7, .85, 1]
TRIAL_P = (0.5, 0.2, 0.1, 0.1, 0.1, 0.0, 0.0, 0.0)      # cumulative [.5, .7, .8, .9, 1]
GROUP_P = (0.3, 0.1, 0.1, 0.1, 0.1, 0.2, 0.1, 0.0)      # cumulative [.3, .4, .5, .6, .7, .9, 1, 1]
BACK_ZERO = np.zeros(1001, np.float32)


def raw(u):
    """The raw 31-bit draw whose uniform (float)r / 2^31 is nearest u."""
    return np.uint32(min(max(int(round(u * 2 ** 31)), 0), 2 ** 31 - 1))


def urand(r):
    return np.float32(np.float32(r) / np.float32(2 ** 31))


def randn_raws(t):
    """Twelve equal draws whose Irwin-Hall sum minus 6 is close to t."""
    return [raw((t + 6.0) / 12.0)] * 12


def randn_value(raws):
    tmp = np.float32(0)
    for r in raws:
        tmp = np.float32(tmp + urand(r))
    return np.float32(tmp - np.float32(6))


def dispatch_u(kind, phase):
    return {("burn", "split"): 0.4, ("burn", "merge"): 0.6, ("trial", "split"): 0.6, ("trial", "merge"): 0.75,
            ("burn", "gsplit"): 0.8, ("burn", "gmerge"): 0.95}[(phase, kind)]


def one_jump(oracle, data, state, tape, phase, p_burn=BURN_P, p_trial=TRIAL_P, is_backg=0, back=None):
    """Runs the oracle chain for exactly one step on a tape; returns the step
    record and the proposed state (I, X, Y) of its chain record."""
    I, X, Y = (np.asarray(v, np.float32) for v in state)
    nb, nt = (1, 0) if phase == "burn" else (0, 1)
    p = params(nb=nb, nt=nt, p_burnin=p_burn, p_trials=p_trial, rho=RHO, is_backg=is_backg)
    r = oracle.rjmcmc(data, SIDE, SIDE, oracle.McmcParams.from_buffer_copy(p), OFF_PDF, SIG_PDF,
                      BACK_ZERO if back is None else back, bg0=10.0, src_kind=oracle.SRC_TAPE,
                      tape=np.array(tape, np.uint32), n0=len(I), i0=I, x0=X, y0=Y, chain=True)
    assert r["status"] == 0
    st = r["steps"][0]
    prop = None
    if st["flags"] & 2:
        em = r["chain"]["prop_em"]
        prop = (em[:, 2].copy(), em[:, 0].copy(), em[:, 1].copy())
    return st, prop


def make_case(rng, n):
    """A 16 x 16 ROI drawn from the forward model, and a chain state near its truth."""
    I = rng.uniform(700, 1800, n).astype(np.float32)
    X = rng.uniform(4, 12, n).astype(np.float32)
    Y = rng.uniform(4, 12, n).astype(np.float32)
    mu = synth.roi_expected(I, X, Y, 10.0, SIDE, SIDE, 1.3)
    data = rng.poisson(mu).astype(np.float32)
    jit = lambda v, s: (v + rng.normal(0, s, n)).astype(np.float32)  # noqa: E731
    return data, (jit(I, 50.0), jit(X, 0.2), jit(Y, 0.2))


def split_tape(phase, n, jj, u1, t2, t3, accept_u=0.999):
    return ([raw(dispatch_u("split", phase)), raw((jj + 0.5) / n), raw(u1)] + randn_raws(t2) + randn_raws(t3)
            + [raw(accept_u)])


def neighbours(X, Y, jj, dmin):
    """Indices within dmin of jj, in index order, with the oracle's distance (double of float differences)."""
    out = []
    for nn in range(len(X)):
        if nn == jj:
            continue
        dx, dy = float(np.float32(X[nn] - X[jj])), float(np.float32(Y[nn] - Y[jj]))
        if math.sqrt(dx * dx + dy * dy) < dmin:
            out.append(nn)
    return out


def merge_tape(phase, state, jj, ii, accept_u=0.999):
    """Merge of emitter jj with ii (ii must be among jj's neighbours within 3 sigma)."""
    n = len(state[0])
    nb = neighbours(state[1], state[2], jj, 3 * 1.3)
    assert ii in nb
    mm = nb.index(ii)
    return [raw(dispatch_u("merge", phase)), raw((jj + 0.5) / n), raw((mm + 0.5) / len(nb)), raw(accept_u)]


def split_then_merge(oracle, rng, n, phase):
    data, E = make_case(rng, n)
    jj = int(rng.integers(0, n))
    u1 = float(rng.uniform(0.25, 0.75))
    t2, t3 = rng.uniform(-0.8, 0.8, 2)
    st_s, T = one_jump(oracle, data, E, split_tape(phase, n, jj, u1, t2, t3), phase)
    assert st_s["jump_type"] == 4 and not (st_s["flags"] & 8) and st_s["n_prop"] == n + 1
    st_m, E2 = one_jump(oracle, data, T, merge_tape(phase, T, jj, n), phase)
    assert st_m["jump_type"] == 5 and not (st_m["flags"] & 8) and st_m["n_prop"] == n
    return data, E, T, E2, jj, st_s, st_m


# ---------------------------------------------------------------- maps
@pytest.mark.parametrize("u1", [0.2, 0.5, 0.8])
def test_split_map_jacobian(oracle, u1):
    v0 = np.array([1100.0, 7.3, 8.1, u1, 0.31, -0.22])

    def f(v):
        return oracle.split_map(*v).astype(np.float64)

    h = np.array([4.0, 0.05, 0.05, 1e-3, 0.02, 0.02])
    J = np.zeros((6, 6))
    for j in range(6):
        e = np.zeros(6)
        e[j] = h[j]
        J[:, j] = (f(v0 + e) - f(v0 - e)) / (2 * h[j])
    got = abs(np.linalg.det(J))
    want = v0[0] / (1 - u1) ** 2
    assert abs(got / want - 1) < 2e-3, (got, want)
    # the power is what matters: (1 - u1)^1 or ^3 is off by 20 % or more
    for wrong in (v0[0] / (1 - u1), v0[0] / (1 - u1) ** 3):
        assert abs(got / wrong - 1) > 0.19


def test_merge_inverts_split(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        I, mx, my = rng.uniform(200, 3000), rng.uniform(0, 16), rng.uniform(0, 16)
        u1, u2, u3 = rng.uniform(0.05, 0.95), rng.normal(0, 0.5), rng.normal(0, 0.5)
        s = oracle.split_map(I, mx, my, u1, u2, u3)
        m = oracle.merge_map(*s)
        np.testing.assert_allclose(m[:3], [I, mx, my], rtol=2e-6, atol=2e-5)
        np.testing.assert_allclose(m[3:], [u1, u2, u3], rtol=1e-4, atol=5e-5)


# ---------------------------------------------------------------- reversibility
@pytest.mark.parametrize("phase", ["burn", "trial"])
def test_split_merge_ratios_cancel(oracle, phase):
    rng = np.random.default_rng(11 if phase == "burn" else 12)
    for trial in range(24):
        n = 1 + trial % 4
        data, E, T, E2, jj, st_s, st_m = split_then_merge(oracle, rng, n, phase)
        # the merge restores the state
        for a, b in zip(E, E2):
            np.testing.assert_allclose(b, a, rtol=1e-5, atol=1e-4)
        total = st_s["log_pr"] + st_m["log_pr"]
        assert abs(total) < 2e-3, (trial, n, st_s["log_pr"], st_m["log_pr"])


def _cancel_residual(oracle, phase, seed, trials=8):
    rng = np.random.default_rng(seed)
    worst = 0.0
    for trial in range(trials):
        n = 1 + trial % 4
        _, _, _, _, _, st_s, st_m = split_then_merge(oracle, rng, n, phase)
        worst = max(worst, abs(st_s["log_pr"] + st_m["log_pr"]))
    return worst


@pytest.mark.parametrize("slip", ["split_jacob_pow1", "split_p_inverted", "split_prm_n", "split_no_area",
                                  "merge_jacob_pow3", "merge_prm_n1"])
def test_planted_split_merge_slips_break_cancellation(oracle, slip):
    assert _cancel_residual(oracle, "trial", 21) < 2e-3
    with oracle.planted_slip(slip):
        assert _cancel_residual(oracle, "trial", 21) > 0.05


# ---------------------------------------------------------------- decomposition
def _model(oracle, state, bg=10.0):
    I, X, Y = (np.asarray(v, np.float32) for v in state)
    return oracle.model(I, X, Y, bg, 0.0, 0.0, SIDE, SIDE, 1.3).astype(np.float64)


def _loglik_ratio(oracle, data, a, b):
    """log L(b) - log L(a), Poisson, float64 from the float32 models (the log D! terms cancel)."""
    ma, mb = _model(oracle, a), _model(oracle, b)
    d = data.astype(np.float64)
    return float(np.sum(d * np.log(mb / ma) - (mb - ma)))


def _log_prior_i(oracle, v):
    return math.log(oracle.lib().orc_stat_findpdf(SIG_PDF.ctypes.data, len(SIG_PDF), 10.0, float(v)))


def _log_jacobian(oracle, I, mux, muy, u1, u2, u3):
    v0 = np.array([I, mux, muy, u1, u2, u3], np.float64)
    h = np.array([4.0, 0.05, 0.05, 1e-3, 0.02, 0.02])
    J = np.zeros((6, 6))
    for j in range(6):
        e = np.zeros(6)
        e[j] = h[j]
        J[:, j] = (oracle.split_map(*(v0 + e)).astype(np.float64) - oracle.split_map(*(v0 - e))) / (2 * h[j])
    return math.log(abs(np.linalg.det(J)))


def _split_pieces(oracle, I, mux, muy, u1, u2, u3, n, p):
    """log of the proposal-independent part of a split of one emitter into two (state N -> N + 1)."""
    lam = RHO * SIDE * SIDE
    s = oracle.split_map(I, mux, muy, u1, u2, u3)
    prior_i = _log_prior_i(oracle, s[0]) + _log_prior_i(oracle, s[3]) - _log_prior_i(oracle, I)
    prior_n = stats.poisson.logpmf(n + 1, lam) - stats.poisson.logpmf(n, lam)
    dens_u = stats.norm.logpdf(u2, 0, 0.5) + stats.norm.logpdf(u3, 0, 0.5)  # u1 ~ U(0, 1): density 1
    log_j = _log_jacobian(oracle, I, mux, muy, u1, u2, u3)
    area_ref = -math.log((SIDE + 2.0) * (SIDE + 2.0))  # LibRJMCMC.cpp:1058-1059, Bnd_out = 2
    p_ref = math.log(p[1] / p[2])                      # LibRJMCMC.cpp:1066
    return prior_i + prior_n - dens_u + log_j + area_ref + p_ref


@pytest.mark.parametrize("phase", ["burn", "trial"])
def test_split_and_merge_ratios_decompose(oracle, phase):
    rng = np.random.default_rng(31 if phase == "burn" else 32)
    P = np.array(BURN_P, np.float32) if phase == "burn" else np.cumsum(np.array(TRIAL_P, np.float32),
                                                                         dtype=np.float32)
    for trial in range(12):
        n = 1 + trial % 4
        data, E, T, E2, jj, st_s, st_m = split_then_merge(oracle, rng, n, phase)
        I, mux, muy = float(E[0][jj]), float(E[1][jj]), float(E[2][jj])
        u1 = float(np.float32(T[0][jj] / np.float32(I)))
        mm = oracle.merge_map(T[0][jj], T[1][jj], T[2][jj], T[0][n], T[1][n], T[2][n])
        # split: the u's it drew, recovered by the (pinned) merge map
        want_s = _loglik_ratio(oracle, data, E, T) + _split_pieces(oracle, I, mux, muy, float(mm[3]), float(mm[4]),
                                                                    float(mm[5]), n, P)
        assert abs(st_s["log_pr"] - want_s) < 2e-3, (trial, st_s["log_pr"], want_s)
        # merge from T: the reverse move, the same pieces with the opposite sign
        want_m = _loglik_ratio(oracle, data, T, E2) - _split_pieces(oracle, float(mm[0]), float(mm[1]), float(mm[2]),
                                                                     float(mm[3]), float(mm[4]), float(mm[5]), n, P)
        assert abs(st_m["log_pr"] - want_m) < 2e-3, (trial, st_m["log_pr"], want_m)
        assert abs(u1 - float(mm[3])) < 1e-4


# ---------------------------------------------------------------- group pair
def group_case(rng, k):
    """One tight group of k emitters (all within 2 sigma of the last) and two far ones."""
    c = rng.uniform(6.5, 9.5, 2)
    ang = rng.uniform(0, 2 * np.pi, k)
    rad = rng.uniform(0.5, 1.1, k)
    X = np.concatenate([c[0] + rad * np.cos(ang), [c[0] - 5.5, c[0] + 5.0]]).astype(np.float32)
    Y = np.concatenate([c[1] + rad * np.sin(ang), [c[1] + 5.5, c[1] - 5.0]]).astype(np.float32)
    X[k - 1], Y[k - 1] = np.float32(c[0]), np.float32(c[1])
    I = rng.uniform(600, 1500, k + 2).astype(np.float32)
    mu = synth.roi_expected(I, X, Y, 10.0, SIDE, SIDE, 1.3)
    return rng.poisson(mu).astype(np.float32), (I, X, Y)


def gsplit_tape(n, jj, num, t1, t2, u3, accept_u=0.999):
    np_ = num + 1
    return ([raw(dispatch_u("gsplit", "burn")), raw((jj + 0.5) / n), raw(0.01 / num)] + [raw(0.0)] * num
            + randn_raws(t1) + randn_raws(t2) + [raw(u3 * np_)] + [raw(accept_u)])


def gmerge_tape(n, jj, num, accept_u=0.999):
    return [raw(dispatch_u("gmerge", "burn")), raw((jj + 0.5) / n), raw(0.01 / num)] + [raw(0.0)] * num + \
        [raw(accept_u)]


def _group_residual(oracle, seed, trials=10):
    """max |log PR(gSplit) + log PR(gMerge) - 2 log(MeanN/(N+1))| over group pairs."""
    rng = np.random.default_rng(seed)
    lam = RHO * SIDE * SIDE
    worst = 0.0
    for trial in range(trials):
        k = 2 + trial % 3
        data, E = group_case(rng, k)
        n = len(E[0])
        jj = k - 1
        nb = neighbours(E[1], E[2], jj, 2 * 1.3)
        assert nb == list(range(k - 1)), nb
        st_s, T = one_jump(oracle, data, E, gsplit_tape(n, jj, len(nb), 0.2 * (trial % 3 - 1), 0.15, 0.4 / k),
                           "burn", p_burn=GROUP_P)
        assert st_s["jump_type"] == 8 and st_s["n_prop"] == n + 1
        nb_m = neighbours(T[1], T[2], n, 3 * 1.3)
        assert nb_m == list(range(k)), nb_m  # the new emitter sees exactly the group
        st_m, E2 = one_jump(oracle, data, T, gmerge_tape(n + 1, n, len(nb_m)), "burn", p_burn=GROUP_P)
        assert st_m["jump_type"] == 9 and st_m["n_prop"] == n
        for a, b in zip(E, E2):
            np.testing.assert_allclose(b, a, rtol=1e-5, atol=1e-4)
        quirk = 2 * math.log(lam / (n + 1))  # gMerge's PrM = NEm/MeanN, LibRJMCMC.cpp:1562
        worst = max(worst, abs(st_s["log_pr"] + st_m["log_pr"] - quirk))
    return worst


def test_group_pair_ratio_product(oracle):
    assert _group_residual(oracle, 41) < 2e-3


@pytest.mark.parametrize("slip", ["gsplit_jacob_np", "gsplit_us_no_np", "gmerge_p_inverted"])
def test_planted_group_slips_break_the_product(oracle, slip):
    with oracle.planted_slip(slip):
        assert _group_residual(oracle, 41, trials=6) > 0.05


# ---------------------------------------------------------------- backFore
BACK_PDF_EXP = synth.gamma_pdf_table(1, 300, 10.0, 1001)  # exponential background pdf, mean 300


def _backfore_label_freq(oracle, I, chains=3000, steps=60):
    p = params(nb=steps, nt=0, p_burnin=(0, 0, 0, 0, 0, 0, 0, 1), rho=RHO, is_backg=1)
    op = oracle.McmcParams.from_buffer_copy(p)
    data = np.full(SIDE * SIDE, 10.0, np.float32)
    X = np.full(len(I), 8.0, np.float32)
    ones = np.zeros(len(I))
    for r in range(chains):
        res = oracle.rjmcmc(data, SIDE, SIDE, op, OFF_PDF, SIG_PDF, BACK_PDF_EXP, bg0=10.0, seed=77, roi=r,
                            n0=len(I), i0=np.asarray(I, np.float32), x0=X, y0=X)
        assert res["status"] == 0 and res["n"] == len(I)
        np.testing.assert_array_equal(res["I"], np.asarray(I, np.float32))
        ones += res["sig"]
    return ones / chains


def _label_law(oracle, I):
    S = np.array([oracle.lib().orc_stat_findpdf(SIG_PDF.ctypes.data, len(SIG_PDF), 10.0, float(v)) for v in I])
    B = np.array([oracle.lib().orc_stat_findpdf(BACK_PDF_EXP.ctypes.data, len(BACK_PDF_EXP), 10.0, float(v))
                  for v in I])
    return S / (S + B)


def test_backfore_label_law(oracle):
    I = np.array([250.0, 500.0, 1400.0], np.float32)
    want = _label_law(oracle, I)
    assert want.min() < 0.3 and want.max() > 0.8  # labels of both kinds are likely
    got = _backfore_label_freq(oracle, I)
    se = np.sqrt(want * (1 - want) / 3000)
    assert np.all(np.abs(got - want) < 4 * se), (got, want)
    with oracle.planted_slip("backfore_swapped"):
        bad = _backfore_label_freq(oracle, I, chains=1000)
    assert np.any(np.abs(bad - want) > 8 * se)
