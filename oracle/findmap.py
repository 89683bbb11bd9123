"""CPU restatement of findMap (@RJ/findMap.m:1-112) and of the refinement
summary of mcmc.m:144-166, over Chain records.  TEST INFRASTRUCTURE ONLY:
only tests/ use it.

Chain records are the device's layout (bamf_chain_rec array `rec` with
fields n, em_index, bg, abg, bbg; emitter entries `em` of (X, Y, Photons,
Signal) rows).  Chain(nn), nn = 1..nloop, is record nn - 1 for nn < nloop;
Chain(nloop) is the empty entry cRJMCMC leaves (cRJMCMC.cpp:159-174).

k-means: MATLAB's kmeans (findMap.m:79) seeds at random; this restatement,
like the device, seeds with the Signal == FB positions of the last state
(in the range) with the most frequent count, then runs Lloyd iterations (at
most 100, kmeans' MaxIter) until no centre moves, nearest centre by squared
distance in single precision (first of equal distances), centres as the
mean of their points formed from sums of the positions rounded to 2^-24,
an empty cluster keeping its centre.  Cluster assignment is therefore
parity unpinned against the reference; the rest is the reference's.
"""
from __future__ import annotations

import math

import numpy as np

FIX = 2.0 ** 24


def mround(x: float) -> int:
    """MATLAB round for x >= 0 (half away from zero)."""
    return int(math.floor(x + 0.5))


def _entries(rec, em, k):
    r = rec[k]
    e = em[int(r["em_index"]):int(r["em_index"]) + max(int(r["n"]), 1)]
    return r, e


def find_map(rec, em, nloop: int, fb: int, cap: int = 128, max_iter: int = 100):
    """Returns (most_frequent, bg_stats[4], clusters (k, 6): X, Y, I means then SDs)."""
    it = int(nloop)
    lo = mround(it / 2)
    # findMap.m:3-11: NN over nn = round(Iter/2):Iter; the empty entry and
    # N = 0 states (Signal -1) leave 0
    nn_counts = []
    for nn in range(lo, it + 1):
        if nn == it:
            nn_counts.append(0)
            continue
        _, e = _entries(rec, em, nn - 1)
        sig = e[:, 3]
        nn_counts.append(int(np.sum(sig == fb)) if np.all(sig != -1) else 0)
    freq = np.bincount(np.asarray(nn_counts), minlength=max(nn_counts) + 1)  # hist(NN, 0:MaxN)
    mf = int(np.argmax(freq))  # findMap.m:13-19: the first of equal maxima
    # findMap.m:33-56: the states with that count
    X, Y, I, bg, abg, bbg = [], [], [], [], [], []
    last = None
    for nn in range(lo, it):
        r, e = _entries(rec, em, nn - 1)
        m = e[:, 3] == fb
        if int(m.sum()) != mf:
            continue
        X.extend(e[m, 0])
        Y.extend(e[m, 1])
        I.extend(e[m, 2])
        bg.append(float(r["bg"]))
        abg.append(float(r["abg"]))
        bbg.append(float(r["bbg"]))
        last = nn
    bgs = np.array([np.mean(bg), np.std(bg, ddof=1) if len(bg) > 1 else 0.0, np.mean(abg), np.mean(bbg)]
                   if bg else [np.nan] * 4)
    k = min(mf, cap)
    if k < 1 or last is None:
        return mf, bgs, np.zeros((0, 6))
    pts = np.stack([np.asarray(X, np.float32), np.asarray(Y, np.float32)], axis=1)
    _, e = _entries(rec, em, last - 1)
    seed = e[e[:, 3] == fb][:k, :2].astype(np.float32)
    cx, cy = seed[:, 0].copy(), seed[:, 1].copy()
    fx = np.rint(pts[:, 0].astype(np.float64) * FIX).astype(np.int64)
    fy = np.rint(pts[:, 1].astype(np.float64) * FIX).astype(np.int64)

    def assign(cx, cy):
        dx = pts[:, 0:1] - cx[None, :]
        dy = pts[:, 1:2] - cy[None, :]
        d = (dx * dx).astype(np.float32) + (dy * dy).astype(np.float32)
        return np.argmin(d, axis=1)  # the first of equal distances

    for _ in range(max_iter):
        a = assign(cx, cy)
        nx, ny = cx.copy(), cy.copy()
        for c in range(k):
            sel = a == c
            n = int(sel.sum())
            if n:
                nx[c] = np.float32(float(fx[sel].sum()) / FIX / n)
                ny[c] = np.float32(float(fy[sel].sum()) / FIX / n)
        moved = np.any(nx != cx) or np.any(ny != cy)
        cx, cy = nx, ny
        if not moved:
            break
    a = assign(cx, cy)
    out = np.full((k, 6), np.nan)
    vals = np.stack([np.asarray(X, np.float64), np.asarray(Y, np.float64), np.asarray(I, np.float64)], axis=1)
    for c in range(k):
        v = vals[a == c]
        if len(v):
            out[c, :3] = v.mean(axis=0)
            out[c, 3:] = v.std(axis=0, ddof=1) if len(v) > 1 else 0.0  # findMap.m:86-95
    return mf, bgs, out


def chain_summary(rec, em, nloop: int):
    """mcmc.m:144-166: per emitter index, mean and SD of X, Y, Photons over
    nn = round(NLoop/2)..NLoop.  Returns (n, (n, 6)) or (None, None) when N varies."""
    it = int(nloop)
    lo = mround(it / 2)
    rows = []
    for nn in range(lo, it):
        _, e = _entries(rec, em, nn - 1)
        rows.append(e[:, :3].astype(np.float64))
    if not rows or any(len(r) != len(rows[0]) for r in rows):
        return None, None
    a = np.stack(rows)  # (records, n, 3)
    out = np.concatenate([a.mean(axis=0), a.std(axis=0, ddof=1) if len(rows) > 1 else np.zeros_like(a[0])], axis=1)
    return len(rows[0]), out
