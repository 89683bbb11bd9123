"""Pins the oracle's gSplit / gMerge (oracle/rjmcmc.c, LibRJMCMC.cpp:1389-1584).

  - gSplit's proposal map has the Jacobian determinant the reference's ratio
    uses, I0 / (1 - u3)^(Np + 1), measured by central differences of the
    oracle's own map (a wrong centring term, a dropped (1 - u3) or an
    exponent off by one changes it by tens of percent);
  - gMerge's map inverts gSplit's: merging the new emitter back into the same
    members restores their intensities and positions and returns the split's
    u1, u2, u3 and I0;
  - in a chain: the dispatch (gSplit/gMerge reached through P_Burnin[4..6]
    also in the trials phase), the draw counts, the accept rule, N and the
    conserved total intensity of accepted group jumps, and the counters that
    these jumps leave alone.
The assembly of the posterior ratios from these pieces is pinned by the
gSplit-then-gMerge ratio product in tests/test_oracle_split_merge.py.
"""
import numpy as np
import pytest

from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params, run_oracle_chains


def _group(rng, n_all, np_):
    I = rng.uniform(300, 2000, n_all).astype(np.float32)
    X = rng.uniform(3, 13, n_all).astype(np.float32)
    Y = rng.uniform(3, 13, n_all).astype(np.float32)
    idx = rng.choice(n_all, np_, replace=False).astype(np.int32)
    return I, X, Y, idx


@pytest.mark.parametrize("np_", [2, 3, 5])
def test_gsplit_map_jacobian(oracle, np_):
    rng = np.random.default_rng(30 + np_)
    I, X, Y, idx = _group(rng, 8, np_)
    u = np.array([0.21, -0.17, 0.6 / np_])

    def f(v):
        Iv, Xv, Yv = I.copy(), X.copy(), Y.copy()
        Iv[idx], Xv[idx], Yv[idx] = v[0:np_], v[np_:2 * np_], v[2 * np_:3 * np_]
        I1, X1, Y1, new, _ = oracle.gsplit_map(Iv, Xv, Yv, idx, v[-3], v[-2], v[-1])
        return np.concatenate([I1, X1, Y1, new]).astype(np.float64)

    v0 = np.concatenate([I[idx], X[idx], Y[idx], u]).astype(np.float64)
    h = np.concatenate([np.full(np_, 4.0), np.full(2 * np_, 0.05), [0.01, 0.01, 1e-3]])
    J = np.zeros((len(v0), len(v0)))
    for j in range(len(v0)):
        e = np.zeros(len(v0))
        e[j] = h[j]
        J[:, j] = (f(v0 + e) - f(v0 - e)) / (2 * h[j])
    I0 = float(I[idx].astype(np.float64).sum())
    want = I0 / (1 - u[2]) ** (np_ + 1)
    got = abs(np.linalg.det(J))
    assert abs(got / want - 1) < 5e-3, (got, want)
    # and the exponent is what matters: (1 - u3)^Np or ^(Np + 2) would be off by > 10 %
    assert abs(I0 / (1 - u[2]) ** np_ / want - 1) > 0.1


@pytest.mark.parametrize("np_", [2, 4])
def test_gmerge_inverts_gsplit(oracle, np_):
    rng = np.random.default_rng(50 + np_)
    I, X, Y, idx = _group(rng, 9, np_)
    u1, u2, u3 = 0.13, -0.22, 0.7 / np_
    I1, X1, Y1, (In, Xn, Yn), I0 = oracle.gsplit_map(I, X, Y, idx, u1, u2, u3)
    # the intensity is conserved exactly up to rounding
    assert abs((I1.astype(np.float64).sum() + In) / I0 - 1) < 1e-6
    Ip, Xp, Yp = np.append(I, In), np.append(X, Xn), np.append(Y, Yn)
    Ip[idx], Xp[idx], Yp[idx] = I1, X1, Y1
    I2, X2, Y2, (U1, U2, U3), I0m = oracle.gmerge_map(Ip, Xp, Yp, len(I), idx)
    np.testing.assert_allclose(I2, I[idx], rtol=1e-5)
    np.testing.assert_allclose(X2, X[idx], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(Y2, Y[idx], rtol=1e-5, atol=1e-5)
    assert abs(U1 - u1) < 1e-4 and abs(U2 - u2) < 1e-4 and abs(U3 - u3) < 1e-5
    assert abs(I0m / I0 - 1) < 1e-6


def clustered_rois(rng, n, side=16, k=6):
    """ROIs holding one tight cluster of k emitters, so group jumps find neighbours."""
    from bamf_amd import synth

    offs = np.arange(n + 1, dtype=np.int64) * side * side
    data = np.zeros(offs[-1], np.float32)
    inits = []
    for r in range(n):
        c = rng.uniform(5, side - 5, 2)
        X = (c[0] + rng.normal(0, 0.9, k)).astype(np.float32)
        Y = (c[1] + rng.normal(0, 0.9, k)).astype(np.float32)
        I = rng.gamma(5, 200, k).astype(np.float32)
        mu = synth.roi_expected(I, X, Y, 10.0, side, side, 1.3)
        data[offs[r]:offs[r + 1]] = rng.poisson(mu)
        inits.append((I, X, Y))
    return {"data": data, "offset": offs, "xsize": np.full(n, side), "ysize": np.full(n, side)}, inits


GROUP_P = dict(p_burnin=(0.3, 0.1, 0.1, 0.1, 0.1, 0.15, 0.15, 0), p_trials=(0.5, 0.1, 0.1, 0.1, 0.1, 0, 0, 0))


def test_group_jumps_in_chain(oracle):
    rng = np.random.default_rng(8)
    rois, inits = clustered_rois(rng, 10)
    p = params(nb=1500, nt=1500, **GROUP_P)
    res, tapes = run_oracle_chains(oracle, rois, p, seed=17, n0=inits)
    n_gs = n_gm = n_gm_trials = acc_g = 0
    for ri, (r, t) in enumerate(zip(res, tapes)):
        assert r["status"] == 0
        st = r["steps"]
        k, jt, d, fl = st["kind"], st["jump_type"], st["draws"], st["flags"]
        assert st["draws"].sum() == len(t)
        burn = np.arange(len(st)) < p.n_burnin
        # trials: rtmp >= PV_Trials[4] = 0.9 falls to P_Burnin's cumulative
        # [0.7, 0.85) -> gSplit (empty here) and [0.85, 1.0) -> gMerge
        assert not np.any(k[~burn] == 5)
        n_gm_trials += int(np.sum(k[~burn] == 6))
        rec = (fl & 2) != 0
        nprev = np.concatenate([[len(inits[ri][0])], st["n"][:-1]])
        gs = rec & (k == 5)
        gm = rec & (k == 6)
        n_gs += int(gs.sum())
        n_gm += int(gm.sum())
        assert np.all(jt[gs] == 8) and np.all(jt[gm] == 9)
        assert np.all(st["n_prop"][gs] == nprev[gs] + 1)
        assert np.all(st["n_prop"][gm] == nprev[gm] - 1)
        # draws: rtmp, Ua, closestMulti (1 + Np - 1), 2 randn, u3, accept -> 28 + Np, Np >= 2
        assert np.all(d[gs] >= 30)
        # gMerge: rtmp, Ua, closestMulti (1 + Np), accept -> 4 + Np, Np >= 1
        assert np.all(d[gm] >= 5)
        # no record: too few emitters or no neighbours; then no accept draw either
        nr = ~rec & ((k == 5) | (k == 6))
        assert np.all(st["n"][nr] == nprev[nr])
        # accept rule U < PR on the last draw
        ends = np.cumsum(d)
        for m in (gs, gm):
            u = (t[ends[m] - 1].astype(np.float64) / 2147483648.0).astype(np.float32)
            assert np.array_equal((fl[m] & 1) != 0, u < st["pr"][m])
        # accepted group jumps conserve the total intensity and change N by one
        ph_prev = np.concatenate([[np.float32(inits[ri][0].sum())], st["photons"][:-1]])
        a = (gs | gm) & ((fl & 1) != 0)
        acc_g += int(a.sum())
        assert np.all(np.abs(st["photons"][a] / ph_prev[a] - 1) < 1e-4)
        assert np.all(st["n"][gs & ((fl & 1) != 0)] == nprev[gs & ((fl & 1) != 0)] + 1)
        # counters count only the five reference jumps
        c = r["counters"]
        for kind in range(5):
            assert c[kind] == 1 + np.sum(k == kind)
    assert n_gs > 100 and n_gm > 100 and n_gm_trials > 100 and acc_g > 5, (n_gs, n_gm, n_gm_trials, acc_g)
