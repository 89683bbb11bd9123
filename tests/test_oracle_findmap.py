"""Pins the oracle's findMap / mcmc.m summary restatement (oracle/findmap.py)
on hand-built Chain records: the second-half range round(NLoop/2)..NLoop
with its empty last entry, the count histogram and its first-maximum rule,
the N = 0 placeholder states, the background statistics, k-means on
well-separated clusters (whose means it must return exactly, by brute force),
and the per-emitter summary in closed form."""
import numpy as np

from bamf_amd.rjmcmc_abi import CHAIN_REC_DTYPE
from oracle import findmap as fm


def build(states):
    """states: list of (bg, abg, bbg, [(x, y, I, signal), ...]); [] is an N = 0 state."""
    rec = np.zeros(len(states), CHAIN_REC_DTYPE)
    em = []
    for k, (bg, abg, bbg, ems) in enumerate(states):
        rec[k]["n"] = len(ems)
        rec[k]["em_index"] = len(em)
        rec[k]["bg"], rec[k]["abg"], rec[k]["bbg"] = bg, abg, bbg
        em.extend(ems if ems else [(-100.0, -100.0, 0.0, -1.0)])
    return rec, np.asarray(em, np.float32)


def test_most_frequent_range_and_empty_entry():
    # nloop = 7: Chain(4..7) are records 3..5 and the empty entry
    s1 = (10.0, 0.0, 0.0, [(1.0, 1.0, 100.0, 1.0)])
    s2 = (11.0, 0.0, 0.0, [(1.0, 1.0, 100.0, 1.0), (5.0, 5.0, 200.0, 1.0)])
    s0 = (12.0, 0.0, 0.0, [])
    rec, em = build([s2, s2, s2, s1, s0, s2])
    # range counts: record 3 -> 1, record 4 -> 0 (placeholder), record 5 -> 2, empty -> 0
    mf, bg, cl = fm.find_map(rec, em, 7, fb=1)
    assert mf == 0  # 0 occurs twice (placeholder and the empty entry)
    assert bg[0] == 12.0 and bg[1] == 0.0  # only the placeholder state carries a BG
    assert cl.shape == (0, 6)
    # nloop = 6: Chain(3..6) -> records 2, 3, 4 and the empty entry: counts 2, 1, 0, 0
    mf, _, _ = fm.find_map(rec, em, 6, fb=1)
    assert mf == 0
    rec2, em2 = build([s1, s1, s2, s2, s2, s1])
    # nloop = 6: records 2..4 (count 2) and the empty entry (0) -> 2
    mf, bg, cl = fm.find_map(rec2, em2, 6, fb=1)
    assert mf == 2 and bg[0] == 11.0
    assert np.allclose(cl[:, :3], [[1.0, 1.0, 100.0], [5.0, 5.0, 200.0]])
    assert np.all(cl[:, 3:] == 0.0)
    # background emitters (Signal 0) are not counted for FB = 1, and are for FB = 0
    s_mix = (9.0, 0.1, 0.2, [(1.0, 1.0, 100.0, 0.0), (5.0, 5.0, 200.0, 1.0)])
    rec3, em3 = build([s_mix] * 5)
    assert fm.find_map(rec3, em3, 5, fb=1)[0] == 1
    assert fm.find_map(rec3, em3, 5, fb=0)[0] == 1
    assert fm.find_map(rec3, em3, 5, fb=1)[2][0, 0] == 5.0
    bgs = fm.find_map(rec3, em3, 5, fb=1)[1]
    assert np.allclose(bgs[[0, 2, 3]], [9.0, 0.1, 0.2])


def test_kmeans_recovers_separated_clusters():
    rng = np.random.default_rng(3)
    centres = np.array([[3.0, 4.0], [10.0, 12.0], [14.0, 2.0]])
    states = []
    for k in range(400):
        order = rng.permutation(3)
        ems = [(centres[c, 0] + rng.normal(0, 0.1), centres[c, 1] + rng.normal(0, 0.1), 1000.0 + c, 1.0)
               for c in order]
        states.append((10.0 + rng.normal(0, 0.5), 0.0, 0.0, ems))
    rec, em = build(states)
    mf, bg, cl = fm.find_map(rec, em, len(states) + 1, fb=1)
    assert mf == 3
    # brute force: group the points by their true centre
    pts = em[rec[(len(states) + 2) // 2 - 1]["em_index"]:]
    got = sorted(map(tuple, cl[:, :3]))
    want = []
    for c in range(3):
        sel = pts[:, 2] == 1000.0 + c
        want.append((pts[sel, 0].astype(np.float64).mean(), pts[sel, 1].astype(np.float64).mean(), 1000.0 + c))
    assert np.allclose(got, sorted(want), rtol=1e-12, atol=1e-12)
    sds = sorted(cl[:, 3])
    assert all(0.07 < s < 0.13 for s in sds)


def test_chain_summary_closed_form():
    states = [(10.0, 0, 0, [(float(k), 2.0 * k, 100.0 + k, 1.0), (1.0, 1.0, 50.0, 1.0)]) for k in range(10)]
    rec, em = build(states)
    n, out = fm.chain_summary(rec, em, 11)  # range: Chain(6..11) -> records 5..9 and the empty entry
    ks = np.arange(5, 10, dtype=np.float64)
    assert n == 2
    assert np.allclose(out[0, :3], [ks.mean(), 2 * ks.mean(), 100 + ks.mean()])
    assert np.allclose(out[0, 3:], [ks.std(ddof=1), 2 * ks.std(ddof=1), ks.std(ddof=1)])
    assert np.allclose(out[1], [1.0, 1.0, 50.0, 0.0, 0.0, 0.0])
    states[7] = (10.0, 0, 0, [(1.0, 1.0, 50.0, 1.0)])
    rec, em = build(states)
    assert fm.chain_summary(rec, em, 11) == (None, None)
