"""Training targets from chain final states (proposal.chain_targets) against a
pixel-by-pixel restatement of the rule: the emitter at (X, Y) adds its
intensity to pixel (floor(X + 1/2), floor(Y + 1/2)) clipped into the ROI; a
pixel with a positive sum is occupied and carries the log of that sum."""
import math

import numpy as np

from bamf_amd.proposal import chain_targets


def brute(emitters, side):
    occ = np.zeros((side, side), np.float32)
    logi = np.zeros((side, side), np.float32)
    for px in range(side):
        for py in range(side):
            tot = 0.0
            for I, X, Y in emitters:
                ex = min(max(math.floor(X + 0.5), 0), side - 1)
                ey = min(max(math.floor(Y + 0.5), 0), side - 1)
                if (ex, ey) == (px, py) and I > 0:
                    tot += I
            if tot > 0:
                occ[px, py] = 1
                logi[px, py] = math.log(tot)
    return occ, logi


def final_rows(sets, cap):
    f = np.zeros((len(sets), 4 + 3 * cap), np.float32)
    for r, em in enumerate(sets):
        f[r, 0] = len(em)
        for k, (I, X, Y) in enumerate(em):
            f[r, 4 + k], f[r, 4 + cap + k], f[r, 4 + 2 * cap + k] = I, X, Y
    return f


def test_chain_targets_rule():
    side, cap = 8, 6
    sets = [
        [],                                                   # empty state
        [(500.0, 3.2, 4.7)],                                  # one emitter: pixel (3, 5)
        [(100.0, 2.49, 2.51), (300.0, 2.2, 2.6)],             # two in one pixel: the sum
        [(200.0, -0.4, 7.8), (50.0, 7.6, -0.45)],             # out past the borders: clipped
        [(0.0, 1.0, 1.0), (-5.0, 4.0, 4.0), (80.0, 6.5, 0.5)],  # zero/negative intensity: not occupied
    ]
    occ, logi = chain_targets(final_rows(sets, cap), side, cap)
    for r, em in enumerate(sets):
        bo, bl = brute(em, side)
        assert np.array_equal(occ[r], bo), r
        np.testing.assert_allclose(logi[r], bl, rtol=1e-6, atol=0)
    assert occ[2].sum() == 1 and abs(logi[2][2, 3] - math.log(400.0)) < 1e-5
    # stale entries past N are ignored
    f = final_rows([[(500.0, 3.2, 4.7), (900.0, 1.0, 1.0)]], cap)
    f[0, 0] = 1
    occ, _ = chain_targets(f, side, cap)
    assert occ[0].sum() == 1 and occ[0][3, 5] == 1


def test_chain_targets_random_states():
    """Random states of up to 11 emitters (shared pixels among them)."""
    rng = np.random.default_rng(3)
    side, cap, n = 16, 40, 30
    sets = []
    for _ in range(n):
        k = rng.integers(0, 12)
        sets.append([(float(rng.gamma(5, 200)), float(rng.uniform(0, side)), float(rng.uniform(0, side)))
                     for _ in range(k)])
    occ, logi = chain_targets(final_rows(sets, cap), side, cap)
    for r in range(n):
        bo, bl = brute(sets[r], side)
        assert np.array_equal(occ[r], bo)
        np.testing.assert_allclose(logi[r], bl, rtol=1e-6)
