"""Posterior summaries of batched chains, and their agreement with a
reference posterior (the long oracle chains of
tests/golden/posterior_batched_uneven.npz, written by
tests/golden/make_posterior_fixture.py).

A posterior is given as independent samples of the chain state (the final
states of independent chains): N and, per sample, N emitters (x, y).

  - localization error of a posterior: every sample's emitters are matched
    one-to-one to the ground-truth emitters at minimum total distance (pairs
    further than `cutoff` pixels apart count as unmatched); per ROI the mean
    over samples of the summed squared distances and of the matched count;
    the RMSE pools the ROIs with equal weight.  This is the posterior
    expectation of the squared localization error, a smooth functional whose
    Monte Carlo error shrinks as 1/sqrt(samples);
  - MAP-N and MAP-N localizations (findMap.m:29-112 in spirit: the most
    frequent N, then the positions of the samples with that N clustered into
    MAP-N groups by k-means from a deterministic start), reported beside it.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import linear_sum_assignment


def map_n(ns: np.ndarray) -> int:
    vals, counts = np.unique(np.asarray(ns), return_counts=True)
    return int(vals[np.argmax(counts)])  # ties: the smallest N


def map_localizations(ns: np.ndarray, pos: list[np.ndarray], iters: int = 60) -> np.ndarray:
    """ns[k] emitters per sample, pos[k] their (ns[k], 2) positions; returns the
    (MAP-N, 2) cluster means of the samples at MAP-N."""
    m = map_n(ns)
    if m == 0:
        return np.zeros((0, 2))
    sel = [p for n, p in zip(ns, pos) if n == m]
    pts = np.concatenate(sel).astype(np.float64)
    centres = sel[0].astype(np.float64).copy()
    for _ in range(iters):
        d = ((pts[:, None, :] - centres[None, :, :]) ** 2).sum(-1)
        lab = np.argmin(d, axis=1)
        new = centres.copy()
        for j in range(m):
            if np.any(lab == j):
                new[j] = pts[lab == j].mean(0)
        if np.allclose(new, centres, atol=1e-9):
            break
        centres = new
    return centres


def matched_errors(est: np.ndarray, truth: np.ndarray, cutoff: float = 2.0) -> np.ndarray:
    """Distances (pixels) of the matched estimate/truth pairs closer than cutoff."""
    if len(est) == 0 or len(truth) == 0:
        return np.zeros(0)
    d = np.sqrt(((est[:, None, :] - truth[None, :, :]) ** 2).sum(-1))
    r, c = linear_sum_assignment(np.minimum(d, cutoff * 10))
    e = d[r, c]
    return e[e < cutoff]


def sample_errors(ns, pos, truth: np.ndarray, cutoff: float = 2.0) -> tuple[float, float]:
    """Mean over samples of (summed squared matched distance, matched count), pixels."""
    sq = cnt = 0.0
    for n, p in zip(ns, pos):
        e = matched_errors(np.asarray(p, np.float64)[:n], truth, cutoff)
        sq += float(np.sum(e ** 2))
        cnt += e.size
    k = max(len(ns), 1)
    return sq / k, cnt / k


def posterior_rmse_nm(per_roi: list[tuple[float, float]], pixel_nm: float = 100.0) -> float:
    """RMSE (nm) from per-ROI (mean summed squared error, mean matched count)."""
    sq = sum(a for a, _ in per_roi)
    cnt = sum(b for _, b in per_roi)
    return float(np.sqrt(sq / cnt) * pixel_nm) if cnt > 0 else float("nan")


def localization_rmse_nm(ests: list[np.ndarray], truths: list[np.ndarray], pixel_nm: float = 100.0,
                         cutoff: float = 2.0) -> tuple[float, int]:
    """RMSE (nm) over all ROIs' matched pairs, and the number of pairs."""
    errs = np.concatenate([matched_errors(e, t, cutoff) for e, t in zip(ests, truths)])
    return float(np.sqrt(np.mean(errs ** 2)) * pixel_nm), int(errs.size)


def load_reference(path) -> dict:
    """The reference posterior fixture: ROIs, truth, parameters, the oracle
    samples' N per ROI and their localization errors."""
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}

    def split(flat, lens):
        out, at = [], 0
        for n in lens:
            out.append(flat[at:at + n])
            at += n
        return out

    d["oracle_n_roi"] = split(d["oracle_n"].astype(np.int64), d["oracle_n_len"])
    d["oracle_loc_roi"] = [a.astype(np.float64) for a in split(d["oracle_loc"], d["oracle_loc_len"])]
    d["oracle_err_roi"] = [tuple(e) for e in d["oracle_err"]]
    d["truth_roi"] = [a.astype(np.float64) for a in split(d["truth_xy"], d["truth_len"])]
    return d


def replicated_batch(ref: dict, k: int, key_base: int = 100000):
    """A device batch holding every reference ROI k times (ROI r's copies at
    r*k .. r*k + k - 1) with distinct stream keys, largest ROIs first in the
    work queue."""
    from .rjmcmc import RoiSet

    side = ref["side"].astype(np.int32)
    o = ref["offset"]
    data = np.concatenate([np.tile(ref["data"][o[r]:o[r + 1]], k) for r in range(len(side))])
    xs = np.repeat(side, k)
    offs = np.zeros(len(xs) + 1, np.int64)
    offs[1:] = np.cumsum(xs.astype(np.int64) ** 2)
    keys = np.arange(len(xs), dtype=np.uint32) + np.uint32(key_base)
    bg0 = np.full(len(xs), float(ref["bg0"]), np.float32)
    order = np.argsort(-(xs.astype(np.int64) ** 2), kind="stable").astype(np.int32)
    return RoiSet(data, offs, xs, xs, bg0=bg0, keys=keys, order=order)


def reference_plan(ref: dict):
    """A chain plan with the reference posterior's parameters and priors."""
    from .rjmcmc import Plan, make_params

    p = make_params(n_burnin=int(ref["n_burnin"]), n_trials=int(ref["n_trials"]), rho=float(ref["rho"]),
                    p_burnin=tuple(ref["p_burnin"]), p_trials=tuple(ref["p_trials"]))
    return Plan(p, ref["off_pdf"], ref["sig_pdf"], np.zeros_like(ref["sig_pdf"]))


def final_samples(final: np.ndarray, n_roi: int, k: int, cap: int):
    """Per reference ROI: the N and (N, 2) positions of its k chains' final states
    (rows of a run's final-state output, bamf_rjmcmc_out.d_final layout)."""
    ns, pos_all = [], []
    for r in range(n_roi):
        rows = final[r * k:(r + 1) * k]
        n = rows[:, 0].astype(np.int64)
        pos = [np.stack([row[4 + cap:4 + cap + m], row[4 + 2 * cap:4 + 2 * cap + m]], 1) for row, m in zip(rows, n)]
        ns.append(n)
        pos_all.append(pos)
    return ns, pos_all


def agreement(ref: dict, ns, pos) -> dict:
    """KS p-values of the samples' N against the reference's (per ROI and pooled),
    the posterior localization RMSE of both, and MAP-N."""
    from scipy import stats

    n_roi = len(ns)
    ps = [float(stats.ks_2samp(ns[r], ref["oracle_n_roi"][r]).pvalue) for r in range(n_roi)]
    pooled = float(stats.ks_2samp(np.concatenate(ns), np.concatenate(ref["oracle_n_roi"])).pvalue)
    errs = [sample_errors(ns[r], pos[r], ref["truth_roi"][r]) for r in range(n_roi)]
    rmse_g, rmse_o = posterior_rmse_nm(errs), posterior_rmse_nm(ref["oracle_err_roi"])
    locs = [map_localizations(ns[r], pos[r]) for r in range(n_roi)]
    return {"ks_p": ps, "ks_p_pooled": pooled, "rmse_nm": rmse_g, "oracle_rmse_nm": rmse_o,
            "rmse_ratio": rmse_g / rmse_o, "map_n": [map_n(v) for v in ns],
            "oracle_map_n": [map_n(v) for v in ref["oracle_n_roi"]],
            "map_rmse_nm": localization_rmse_nm(locs, ref["truth_roi"])[0],
            "oracle_map_rmse_nm": localization_rmse_nm(ref["oracle_loc_roi"], ref["truth_roi"])[0],
            "mean_n": [float(v.mean()) for v in ns], "oracle_mean_n": [float(v.mean()) for v in ref["oracle_n_roi"]]}


def split_snapshots(snap_n: np.ndarray, snap_em: np.ndarray) -> list[np.ndarray]:
    """Per-sample (N, 2) position arrays from the oracle's flat snapshot records."""
    out, at = [], 0
    for n in snap_n:
        out.append(snap_em[at:at + n, :2])
        at += n
    return out
