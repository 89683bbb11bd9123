"""The refinement of @RJ/mcmc.m and the ClustInfo assembly of
@RJ/analyzeROIs.m:112-199, over the device chain, findMap and summary.

Per tile (a sub-region of makeSubRegions.m, bamf_amd.tiles.Tiling):
  1. the RJMCMC chain with rjmcmc.m's parameters;
  2. findMap for signal (FB = 1) and background (FB = 0) emitters
     (bamf_find_map);
  3. when the signal MostFrequent is not 0: a fixed-dimension chain of moves
     only (P = [1 0 0 0 0 0 0 0], mcmc.m:103) started from the signal SMD,
     summarised over its second half (bamf_chain_summary, mcmc.m:144-166);
  4. on the host: emitters in the tile's overlap bands removed, positions
     shifted to frame coordinates (mcmc.m:42-84,167-231), and the tile's
     entries appended to ClustInfo (analyzeROIs.m:156-199).
analyze_tiles runs steps 1-3 for every tile of a Tiling as three batched
launches; mcmc() is the single-ROI mirror of mcmc.m over the cRJMCMC host
function.
"""
from __future__ import annotations

import math

import numpy as np

from . import _hip
from ._lib import BamfError
from .rjmcmc import (STAT_NAMES, Plan, RoiSet, chain_summary, crjmcmc, fetch, find_map, make_params,
                     run_batch)

SMD_FIELDS = ("X", "Y", "I", "X_SE", "Y_SE", "I_SE")
CLUST_FIELDS = ["X", "Y", "I", "X_SE", "Y_SE", "I_SE", "X_kmeans", "Y_kmeans", "I_kmeans", "X_SE_kmeans",
                "Y_SE_kmeans", "I_SE_kmeans", "Xbg", "Ybg", "Ibg", "Xbg_SE", "Ybg_SE", "Ibg_SE", "FrameNum",
                "FrameNumBg", "FrameNum_kmeans", "ROInum", "Bg", "ABg", "BBg", "JumpXRJ", "JumpIRJ", "JumpBgRJ",
                "JumpXFg", "JumpIFg", "JumpXBg", "JumpIBg", "SplitRJ", "MergeRJ", "BirthRJ", "DeathRJ", "JumpMC",
                "ROIcornerX", "ROIcornerY"]


def mround(x: float) -> int:
    """MATLAB round (half away from zero)."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


# ---------------------------------------------------------------- parameters
def rj_params(r: dict, drift_x: float = 0.0, drift_y: float = 0.0):
    """The MCMC struct of rjmcmc.m:49-78."""
    return make_params(psf_sigma=r["PSF_Sigma"], n_burnin=r["N_Burnin"], n_trials=r["N_Trials"],
                       p_burnin=r["P_Burnin"], p_trials=r["P_Trials"], rho=r["Rho"], x_std_fg=r["XstdFg"],
                       i_std_fg=r["IstdFg"], bg_std=r["BGstd"], dx=r["P_dP"], grid_zoom=r["SRZoom"],
                       icutoff=r.get("MinPhoton", 0.0), is_backg=r.get("IsBackg", 0), drift_x=drift_x,
                       drift_y=drift_y)


def refine_params(r: dict, drift_x: float = 0.0, drift_y: float = 0.0):
    """The MCMC struct of mcmc.m:88-118: N_TrialsMCMC / N_BurninMCMC, moves
    only, I_stdBg = IstdFg / 1.5, X_stdBg = XstdFg / 1.5, IsBackg 0."""
    p = make_params(psf_sigma=r["PSF_Sigma"], n_burnin=r["N_BurninMCMC"], n_trials=r["N_TrialsMCMC"],
                    p_burnin=(1, 0, 0, 0, 0, 0, 0, 0), p_trials=(1, 0, 0, 0, 0, 0, 0, 0), rho=r["Rho"],
                    x_std_fg=r["XstdFg"], i_std_fg=r["IstdFg"], bg_std=r["BGstd"], dx=r["P_dP"],
                    grid_zoom=r["SRZoom"], icutoff=r.get("MinPhoton", 0.0), is_backg=0, drift_x=drift_x,
                    drift_y=drift_y)
    p.i_std_bg = float(r["IstdFg"]) / 1.5  # a double in MATLAB, read as single by the mex
    p.x_std_bg = float(r["XstdFg"]) / 1.5
    return p


def offset_bg0(r: dict) -> np.float32:
    """E_active.BG of rjmcmc.m:46: P_dP times the 0-based index of P_Offset's first maximum."""
    p = np.asarray(r["P_Offset"], np.float64).ravel()
    return np.float32(r["P_dP"] * int(np.flatnonzero(p == p.max())[0]))


def drift_at(r: dict, key: str, frame: int) -> float:
    """RJstruct.XDrift(frame) / YDrift(frame) (1-based), 0 when the field is empty."""
    d = r.get(key)
    if d is None or np.size(d) == 0:
        return 0.0
    return float(np.asarray(d, np.float64).ravel()[frame - 1])


def stat_from_counters(c, n_trials: int) -> dict:
    """Stat of cRJMCMC.cpp:283-305 from the 30 jump counters (proposed and
    accepted, for all, signal and background emitters) of a chain of n_trials
    jumps (burn-in included)."""
    c = np.asarray(c, np.float64)
    out = {"NTrials": float(n_trials)}
    with np.errstate(invalid="ignore", divide="ignore"):
        for i, nm in enumerate(("Jump", "Split", "Merge", "Birth", "Death")):
            out[f"Accept_{nm}"] = float(c[5 + i] / c[i])
            out[f"Accept_{nm}Fg"] = float(c[15 + i] / c[10 + i])
            out[f"Accept_{nm}Bg"] = float(c[25 + i] / c[20 + i])
    return {k: out[k] for k in STAT_NAMES}


# ---------------------------------------------------------------- host glue (mcmc.m)
def tile_window(roi_struct: dict, roi_id: int) -> dict:
    """mcmc.m:43-69 for ROIid (1-based): the tile's block indices and the
    window inside it whose emitters it keeps (the overlap bands toward
    neighbouring tiles excluded)."""
    ind = int(np.asarray(roi_struct["ROI_Ind"]).ravel()[roi_id - 1])
    xsub, ysub, ov = int(roi_struct["Xsub"]), int(roi_struct["Ysub"]), int(roi_struct["Overlap"])
    ix = (ind - 1) // xsub + 1
    iy = ind - (ix - 1) * xsub  # the reference indexes both axes with Xsub
    bx, by = 0, 0
    ex = int(np.asarray(roi_struct["Xsize"]).ravel()[roi_id - 1])
    ey = int(np.asarray(roi_struct["Ysize"]).ravel()[roi_id - 1])
    ox = oy = 0
    if ix != xsub:
        ex -= ov
    if iy != ysub:
        ey -= ov
    if ix != 1:
        bx += ov
        ox = ov
    if iy != 1:
        by += ov
        oy = ov
    return {"ix": ix, "iy": iy, "bx": bx, "ex": ex, "by": by, "ey": ey, "ox": ox, "oy": oy}


def _outside(X, Y, w):
    X, Y = np.asarray(X, np.float64), np.asarray(Y, np.float64)
    return (X <= w["bx"]) | (X > w["ex"]) | (Y <= w["by"]) | (Y > w["ey"])


def _shift(w, roi_struct):
    step = int(np.asarray(roi_struct["Xsize"]).ravel()[0]) - int(roi_struct["Overlap"])  # Xsize(1) on both axes
    return (w["ix"] - 1) * step - w["ox"], (w["iy"] - 1) * step - w["oy"]


def found_map_from_chain(chain: list) -> dict:
    """mcmc.m:133-166 on cRJMCMC's Chain entries (the AcceptChain of a
    fixed-dimension run): X, Y, Photons of entries round(L/2)..L stacked as
    rows in single precision; FoundMAP.X = mean(X) (per column; a single row
    is averaged along itself, as MATLAB's mean does), X_SE(nn) = std(X(:, nn))."""
    L = len(chain)
    rows = [(np.asarray(e["X"], np.float32).ravel(), np.asarray(e["Y"], np.float32).ravel(),
             np.asarray(e["Photons"], np.float32).ravel())
            for e in chain[max(mround(L / 2), 1) - 1:L] if e.get("X") is not None]
    if not rows:
        return {k: np.zeros(0, np.float32) for k in SMD_FIELDS}
    if any(len(r[0]) != len(rows[0][0]) for r in rows):
        raise BamfError("the chain's states differ in N (cat(1, X, x) fails, mcmc.m:137)")
    out = {}
    for k, name in enumerate(("X", "Y", "I")):
        M = np.stack([r[k] for r in rows]).astype(np.float32)
        if M.shape[0] == 1:
            out[name] = np.array([M.mean(dtype=np.float64)], np.float32)
        else:
            out[name] = M.mean(axis=0, dtype=np.float64).astype(np.float32)
        sd = M.std(axis=0, ddof=1, dtype=np.float64) if M.shape[0] > 1 else np.zeros(M.shape[1])
        out[name + "_SE"] = sd.astype(np.float32)
    return out


def trim(found: dict, smd: dict, smdbg: dict, w: dict, roi_struct: dict, xdrift: float = 0.0,
         ydrift: float = 0.0, xstd_cutoff=None, istd_cutoff=None):
    """mcmc.m:71-84 and 167-231: FoundMAP, SMDkmeans and SMDbg in frame
    coordinates with the emitters outside the tile's window removed
    (FoundMAP by its own positions, the SMDs by theirs less the frame's drift),
    then, with cutoffs given, FoundMAP emitters whose X_SE, Y_SE or I_SE
    exceed them.  Returns (FoundMAP, SMDkmeans, SMDbg) as dicts of float32
    arrays; FoundMAP also carries ROIcorner."""
    sx, sy = _shift(w, roi_struct)
    size = roi_struct["SizeData"]
    corner = [(w["ix"] - 1) * size[1] / roi_struct["Xsub"], (w["iy"] - 1) * size[0] / roi_struct["Ysub"]]

    def place(d, drop):
        out = {}
        for k in SMD_FIELDS:
            v = np.asarray(d.get(k, []), np.float64).ravel()
            if k == "X":
                v = v + sx
            elif k == "Y":
                v = v + sy
            out[k] = v[~drop].astype(np.float32) if v.size else v.astype(np.float32)
        return out

    def kmeans_drop(d):
        X = np.asarray(d.get("X", []), np.float64).ravel()
        Y = np.asarray(d.get("Y", []), np.float64).ravel()
        return _outside(X - xdrift, Y - ydrift, w) if X.size else np.zeros(0, bool)

    smdk = place(smd, kmeans_drop(smd))
    smdb = place(smdbg, kmeans_drop(smdbg))
    X = np.asarray(found["X"], np.float64).ravel()
    fm = place(found, _outside(X, np.asarray(found["Y"], np.float64).ravel(), w))
    if xstd_cutoff is not None:
        for key, cut in (("X_SE", xstd_cutoff), ("Y_SE", xstd_cutoff), ("I_SE", istd_cutoff)):
            keep = ~(fm[key] > cut)
            fm = {k: v[keep] for k, v in fm.items()}
    fm["ROIcorner"] = corner
    return fm, smdk, smdb


def smd_of(fm_out: dict, r: int) -> tuple[dict, int]:
    """SMD (findMap.m:52-112) of ROI r from bamf_find_map's outputs, and MostFrequent."""
    mf = int(fm_out["n"][r])
    k = min(mf, fm_out["em"].shape[1])
    em = fm_out["em"][r, :k].astype(np.float64)
    smd = {f: em[:, j].copy() for j, f in enumerate(SMD_FIELDS)}
    bg = fm_out["bg"][r]
    smd.update({"Bg": float(bg[0]), "Bg_SE": float(bg[1]), "ABg": float(bg[2]), "BBg": float(bg[3])})
    return smd, mf


class ClustInfo:
    """ClustInfo of analyzeROIs.m:72-110,156-199: single-precision row
    vectors grown tile by tile."""

    def __init__(self):
        self.parts = {k: [] for k in CLUST_FIELDS}

    def add_background(self, smd: dict) -> None:
        for k in ("Bg", "ABg", "BBg"):
            self.parts[k].append(np.atleast_1d(np.float32(smd[k])))

    def add(self, found: dict, smdk: dict, smdbg: dict, stat_rj: dict, stat_mc: dict, frame: int,
            loop_num: int = 1) -> None:
        P = self.parts
        for k in SMD_FIELDS:
            P[k].append(found[k].astype(np.float32))
            P[k + "_kmeans" if k in ("X", "Y", "I") else k.replace("_SE", "_SE_kmeans")].append(
                smdk[k].astype(np.float32))
        for src, dst in (("X", "Xbg"), ("Y", "Ybg"), ("I", "Ibg"), ("X_SE", "Xbg_SE"), ("Y_SE", "Ybg_SE"),
                         ("I_SE", "Ibg_SE")):
            P[dst].append(smdbg[src].astype(np.float32))
        nf, nk, nb = found["X"].size, smdk["X"].size, smdbg["X"].size
        P["FrameNum_kmeans"].append(np.full(nk, frame, np.float32))
        for dst, src in (("JumpXRJ", "Accept_Jump"), ("JumpIRJ", "Accept_Jump"), ("JumpBgRJ", "Accept_JumpBg"),
                         ("JumpXFg", "Accept_JumpFg"), ("JumpIFg", "Accept_JumpFg"), ("JumpXBg", "Accept_JumpBg"),
                         ("JumpIBg", "Accept_JumpBg"), ("SplitRJ", "Accept_SplitFg"), ("MergeRJ", "Accept_MergeFg"),
                         ("BirthRJ", "Accept_BirthFg"), ("DeathRJ", "Accept_DeathFg")):
            P[dst].append(np.atleast_1d(np.float32(stat_rj[src])))
        P["JumpMC"].append(np.atleast_1d(np.float32(stat_mc["Accept_JumpFg"])))
        P["FrameNum"].append(np.full(nf, frame, np.float32))
        P["FrameNumBg"].append(np.full(nb, frame, np.float32))
        P["ROInum"].append(np.full(nf, loop_num, np.float32))
        P["ROIcornerX"].append(np.full(nf, found["ROIcorner"][0], np.float32))
        P["ROIcornerY"].append(np.full(nf, found["ROIcorner"][1], np.float32))

    def arrays(self) -> dict:
        return {k: (np.concatenate(v).astype(np.float32) if v else np.zeros(0, np.float32))
                for k, v in self.parts.items()}


# ---------------------------------------------------------------- single ROI
def mcmc(data, rjstruct: dict, xsize: int, ysize: int, drift_y: float, drift_x: float, scmos_var, sampled_psf_in,
         opt_model_in, roi_struct: dict, smd: dict, smdbg: dict, roi_id: int, frame_id: int, xstd_cutoff=None,
         istd_cutoff=None, seed: int = 0):
    """@RJ/mcmc.m for one ROI: the moves-only chain from SMD (E_active.X =
    SMD.X - DriftX, ..., BG = SMD.Bg), run through cRJMCMC, summarised and
    trimmed.  Returns (FoundMAP, SMDkmeans, SMDbg, AcceptChain, Stat)."""
    d0 = np.asarray(data, np.float32)
    if d0.ndim > 2 and d0.shape[2] > 1:
        raise BamfError("The input data to rjmcmc() must be a single ROI.")
    d = np.atleast_2d(d0.reshape(d0.shape[0], -1) if d0.ndim > 2 else d0)[:ysize, :xsize]
    p = refine_params(rjstruct, drift_x, drift_y)
    sx = np.asarray(smd["X"], np.float64).ravel()
    e_active = {"N": sx.size, "I": np.asarray(smd["I"], np.float32), "X": (sx - drift_x).astype(np.float32),
                "Y": (np.asarray(smd["Y"], np.float64).ravel() - drift_y).astype(np.float32),
                "BG": np.float32(smd["Bg"])}
    sc = None
    if scmos_var is not None and np.ndim(scmos_var) == 2:
        sc = np.asarray(scmos_var, np.float32)[:ysize, :xsize]
    psf, opt = None, 0
    if sampled_psf_in is not None and np.ndim(sampled_psf_in) > 0 and np.size(sampled_psf_in) > 0 and \
            opt_model_in is not None and np.size(opt_model_in) > 0:
        opt = int(np.asarray(opt_model_in).ravel()[0])
        psf = sampled_psf_in if opt == 1 else None
    _, _, chain, _, stat = crjmcmc(p, d, xsize, ysize, e_active, rjstruct["P_Offset"], rjstruct["P_EmPhotons"],
                                   rjstruct["P_BgEmPhotons"], sc, psf, opt, seed)
    found = found_map_from_chain(chain)
    w = tile_window(roi_struct, roi_id)
    fm, smdk, smdb = trim(found, smd, smdbg, w, roi_struct, drift_at(rjstruct, "XDrift", frame_id),
                          drift_at(rjstruct, "YDrift", frame_id), xstd_cutoff, istd_cutoff)
    fm["Stat"] = stat
    return fm, smdk, smdb, chain, stat


# ---------------------------------------------------------------- batched tiles
def analyze_tiles(tiling, rjstruct: dict, seed: int = 0, loop_num: int = 1, temp_frame: int | None = None,
                  rec_cap: int | None = None, em_per_record: int = 16, cap: int = 128, debug: bool = False):
    """Steps 1-4 for every tile of a Tiling whose run() has been called:
    the RJMCMC chains, both findMaps and the refinement chains as batched
    device launches (one set per distinct frame drift), then the host glue
    in analyzeROIs.m's order (frame pp outer, tile mm inner).  Returns
    ClustInfo as a dict of float32 arrays (with debug, also the device
    outputs it was assembled from)."""
    _hip.synchronize()
    st = tiling.struct()
    n = tiling.n
    temp_frame = tiling.nframes if temp_frame is None else temp_frame
    frames = np.asarray(st["FrameNum"]).astype(np.int64)  # 1-based frame of each tile
    gframe = frames + (loop_num - 1) * temp_frame
    dx = np.array([drift_at(rjstruct, "XDrift", int(f)) for f in gframe])
    dy = np.array([drift_at(rjstruct, "YDrift", int(f)) for f in gframe])
    bg0 = offset_bg0(rjstruct)
    pr = rj_params(rjstruct)
    total_rj = pr.n_burnin + pr.n_trials
    caps_rj = (rec_cap or total_rj + 1, (rec_cap or total_rj + 1) * em_per_record + 1024)
    pm = refine_params(rjstruct)
    total_mc = pm.n_burnin + pm.n_trials
    caps_mc = (total_mc + 1, (total_mc + 1) * em_per_record + 1024)
    base = tiling.roi_set()
    offs_all = base.offsets
    res = {"mf": np.zeros(n, np.int64), "smd": [None] * n, "smdbg": [None] * n, "stat_rj": [None] * n,
           "stat_mc": [None] * n, "found": [None] * n}
    for (gx, gy) in sorted(set(zip(dx.tolist(), dy.tolist()))):
        idx = np.flatnonzero((dx == gx) & (dy == gy))
        sub = lambda sel, **kw: RoiSet(  # noqa: E731
            None, offs_all[sel], base.xsize[sel], base.ysize[sel],
            device=(tiling.d_pixels, _hip.DeviceBuffer.from_array(offs_all[sel]),
                    _hip.DeviceBuffer.from_array(base.xsize[sel]), _hip.DeviceBuffer.from_array(base.ysize[sel])),
            keys=sel.astype(np.uint32), **kw)
        # 1-2: RJMCMC and findMap
        plan = Plan(rj_params(rjstruct, gx, gy), rjstruct["P_Offset"], rjstruct["P_EmPhotons"],
                    rjstruct["P_BgEmPhotons"])
        rs = sub(idx, bg0=np.full(idx.size, bg0, np.float32))
        outs = run_batch(plan, rs, seed=seed, chain_caps=caps_rj, sync=False, final_cap=cap)
        _hip.synchronize()
        got = fetch(outs, idx.size, total_rj, 64, cap, 0)
        if np.any(got.status != 0):
            raise BamfError(f"RJMCMC status {np.unique(got.status)}")
        fm1, fm0 = find_map(outs, idx.size, fb=1, cap=cap), find_map(outs, idx.size, fb=0, cap=cap)
        del outs
        refine = []
        for j, t in enumerate(idx):
            smd, mf = smd_of(fm1, j)
            smdbg, _ = smd_of(fm0, j)
            res["mf"][t], res["smd"][t], res["smdbg"][t] = mf, smd, smdbg
            res["stat_rj"][t] = stat_from_counters(got.counters[j], total_rj)
            if mf != 0:
                refine.append(t)
        if not refine:
            continue
        # 3: moves-only chains from the signal SMD
        sel = np.array(refine, np.int64)
        n0 = np.array([res["smd"][t]["X"].size for t in sel], np.int32)
        eoff = np.zeros(sel.size, np.int64)
        eoff[1:] = np.cumsum(n0)[:-1]
        cat = lambda f, d: np.concatenate([np.asarray(res["smd"][t][f], np.float64) - d for t in sel])  # noqa: E731
        em0 = (cat("I", 0.0).astype(np.float32), cat("X", gx).astype(np.float32), cat("Y", gy).astype(np.float32),
               eoff)
        rs2 = sub(sel, bg0=np.array([res["smd"][t]["Bg"] for t in sel], np.float32), n0=n0, em0=em0)
        plan2 = Plan(refine_params(rjstruct, gx, gy), rjstruct["P_Offset"], rjstruct["P_EmPhotons"],
                     rjstruct["P_BgEmPhotons"])
        outs2 = run_batch(plan2, rs2, seed=seed + 1, chain_caps=caps_mc, sync=False, final_cap=cap)
        _hip.synchronize()
        got2 = fetch(outs2, sel.size, total_mc, 64, cap, 0)
        cs = chain_summary(outs2, sel.size, cap=cap)
        del outs2
        for j, t in enumerate(sel):
            if cs["status"][j] != 0:
                raise BamfError(f"chain summary status {cs['status'][j]} on tile {t}")
            k = int(cs["n"][j])
            em = cs["em"][j, :k]
            res["found"][t] = {f: em[:, c].astype(np.float32) for c, f in enumerate(SMD_FIELDS)}
            res["stat_mc"][t] = stat_from_counters(got2.counters[j], total_mc)
    out = assemble_clust_info(res, st, rjstruct, loop_num, temp_frame)
    if debug:
        out["_device"] = res
    return out


def assemble_clust_info(res: dict, roi_struct: dict, rjstruct: dict, loop_num: int = 1,
                        temp_frame: int | None = None) -> dict:
    """Step 4 in analyzeROIs.m's order (frame pp outer, tile mm inner) from
    the per-tile results: res["mf"] (signal MostFrequent), res["smd"] and
    res["smdbg"] (findMap SMDs), res["found"] (the refinement summaries),
    res["stat_rj"] and res["stat_mc"] (the two chains' Stat)."""
    frames = np.asarray(roi_struct["FrameNum"]).astype(np.int64).ravel()
    n = frames.size
    nframes = int(roi_struct["SizeData"][2]) if len(roi_struct["SizeData"]) > 2 else 1
    per_frame = n // nframes
    temp_frame = nframes if temp_frame is None else temp_frame
    ci = ClustInfo()
    for t in range(n):
        pp = int(frames[t])
        mm = t - (pp - 1) * per_frame + 1
        ci.add_background(res["smd"][t])
        if res["mf"][t] == 0:
            continue
        w = tile_window(roi_struct, mm)
        fm, smdk, smdb = trim(res["found"][t], res["smd"][t], res["smdbg"][t], w, roi_struct,
                              drift_at(rjstruct, "XDrift", pp), drift_at(rjstruct, "YDrift", pp))
        ci.add(fm, smdk, smdb, res["stat_rj"][t], res["stat_mc"][t], pp + (loop_num - 1) * temp_frame, loop_num)
    return ci.arrays()
