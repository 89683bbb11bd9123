"""The host side of the mcmc.m refinement and the ClustInfo assembly
(bamf_amd.refine) against a statement-by-statement restatement of
@RJ/mcmc.m:42-84,144-231 (index lists, unique, deletions field by field)
and hand-worked windows, and the chain summary against the oracle's."""
import math

import numpy as np

from bamf_amd import refine
from oracle import findmap as ofm


def roi_struct(ypix=50, xpix=50, nframes=2, box=16, overlap=3):
    """makeSubRegions.m:42-101's ROI_Struct, restated with its loops."""
    xsub, ysub = math.ceil(xpix / box), math.ceil(ypix / box)
    n = xsub * ysub * nframes
    st = {"Xsize": np.zeros(n, int), "Ysize": np.zeros(n, int), "FrameNum": np.zeros(n, int),
          "ROI_Ind": np.zeros(n, int), "SizeData": [ypix, xpix, nframes], "Xsub": xsub, "Ysub": ysub,
          "Overlap": overlap}
    for frame in range(1, nframes + 1):
        for ix in range(1, xsub + 1):
            for iy in range(1, ysub + 1):
                ex_, ey_ = box + 2 * overlap, box + 2 * overlap
                if iy == 1:
                    ey_ = box + overlap
                if ix == 1:
                    ex_ = box + overlap
                if iy == ysub:
                    ey_ = ypix - box * (ysub - 1) + overlap
                if ix == xsub:
                    ex_ = xpix - box * (xsub - 1) + overlap
                if xsub == 1:
                    ex_ = xpix
                if ysub == 1:
                    ey_ = ypix
                ind = xsub * ysub * (frame - 1) + ysub * (ix - 1) + iy
                st["Xsize"][ind - 1], st["Ysize"][ind - 1] = ex_, ey_
                st["FrameNum"][ind - 1], st["ROI_Ind"][ind - 1] = frame, ysub * (ix - 1) + iy
    return st


def matlab_mcmc_tail(found, smd, smdbg, rs, roi_id, frame_id, xdrift, ydrift, cut=None):
    """mcmc.m:43-84 and 167-231 as written: index vectors, unique, deletions."""
    Ind = rs["ROI_Ind"][roi_id - 1]
    Xsub, Ysub, Overlap = rs["Xsub"], rs["Ysub"], rs["Overlap"]
    ix = math.floor((Ind - 1) / Xsub) + 1
    iy = Ind - (ix - 1) * Xsub
    BeginROIx = BeginROIy = 0
    EndROIx, EndROIy = rs["Xsize"][roi_id - 1], rs["Ysize"][roi_id - 1]
    BeginOverlapX = BeginOverlapY = 0
    if ix != Xsub:
        EndROIx -= Overlap
    if iy != Ysub:
        EndROIy -= Overlap
    if ix != 1:
        BeginROIx += Overlap
        BeginOverlapX = Overlap
    if iy != 1:
        BeginROIy += Overlap
        BeginOverlapY = Overlap

    def find(v):
        return [k for k, x in enumerate(v) if x]

    def index(X, Y, dx, dy):
        a = find([x - dx <= BeginROIx for x in X]) + find([x - dx > EndROIx for x in X])
        b = find([y - dy <= BeginROIy for y in Y]) + find([y - dy > EndROIy for y in Y])
        return sorted(set(a + b))

    XD, YD = xdrift[frame_id - 1], ydrift[frame_id - 1]
    Index_k = index(smd["X"], smd["Y"], XD, YD)
    Index_b = index(smdbg["X"], smdbg["Y"], XD, YD)
    F = {k: list(map(float, found[k])) for k in refine.SMD_FIELDS}
    Index = index(F["X"], F["Y"], 0.0, 0.0)
    step = rs["Xsize"][0] - Overlap
    F["X"] = [x + (ix - 1) * step - BeginOverlapX for x in F["X"]]
    F["Y"] = [y + (iy - 1) * step - BeginOverlapY for y in F["Y"]]

    def moved(d, drop):
        out = {}
        for k in refine.SMD_FIELDS:
            v = list(map(float, d[k]))
            if k == "X":
                v = [x + (ix - 1) * step - BeginOverlapX for x in v]
            if k == "Y":
                v = [y + (iy - 1) * step - BeginOverlapY for y in v]
            out[k] = [x for j, x in enumerate(v) if j not in drop]
        return out

    K, B = moved(smd, Index_k), moved(smdbg, Index_b)
    for k in refine.SMD_FIELDS:
        F[k] = [x for j, x in enumerate(F[k]) if j not in Index]
    if cut is not None:
        xc, ic = cut
        for key, c in (("X_SE", xc), ("Y_SE", xc), ("I_SE", ic)):
            # X(FoundMAP.X_SE > c) = []; ...; the tested field itself last
            for k in [f for f in refine.SMD_FIELDS if f != key] + [key]:
                mask = [v > c for v in F[key]]
                F[k] = [x for x, m in zip(F[k], mask) if not m]
    corner = [(ix - 1) * rs["SizeData"][1] / Xsub, (iy - 1) * rs["SizeData"][0] / Ysub]
    return F, K, B, corner


def rand_set(rng, k, lo, hi):
    return {"X": rng.uniform(lo, hi, k), "Y": rng.uniform(lo, hi, k), "I": rng.gamma(5, 200, k),
            "X_SE": rng.uniform(0, 0.2, k), "Y_SE": rng.uniform(0, 0.2, k), "I_SE": rng.uniform(0, 80, k)}


def test_tile_windows_by_hand():
    rs = roi_struct()  # 50 x 50, box 16, overlap 3: 4 x 4 tiles of sides 19 / 22 / 22 / 5
    assert list(rs["Xsize"][:4]) == [19, 19, 19, 19] and rs["Xsize"][4] == 22 and rs["Xsize"][12] == 5
    w = refine.tile_window(rs, 1)  # corner tile: keeps x in (0, 16], y in (0, 16]
    assert (w["ix"], w["iy"], w["bx"], w["ex"], w["by"], w["ey"]) == (1, 1, 0, 16, 0, 16)
    w = refine.tile_window(rs, 6)  # interior (ix 2, iy 2): (3, 19] in both
    assert (w["ix"], w["iy"], w["bx"], w["ex"], w["by"], w["ey"], w["ox"]) == (2, 2, 3, 19, 3, 19, 3)
    w = refine.tile_window(rs, 16)  # last tile: side 5, no band past the frame edge
    assert (w["ix"], w["iy"], w["bx"], w["ex"], w["ox"], w["oy"]) == (4, 4, 3, 5, 3, 3)
    # an interior emitter at tile (2, 2) position 5.5 lands at frame x 16 + 5.5 - 3
    fm, _, _ = refine.trim({k: np.array([v]) for k, v in zip(refine.SMD_FIELDS, (5.5, 7.25, 900, .1, .1, 20))},
                           {k: np.zeros(0) for k in refine.SMD_FIELDS},
                           {k: np.zeros(0) for k in refine.SMD_FIELDS}, refine.tile_window(rs, 6), rs)
    assert fm["X"][0] == np.float32(18.5) and fm["Y"][0] == np.float32(20.25) and fm["ROIcorner"] == [12.5, 12.5]


def test_trim_matches_mcmc_m_statements():
    rng = np.random.default_rng(4)
    rs = roi_struct(ypix=60, xpix=60, nframes=3)
    ntile = len(rs["Xsize"]) // 3
    xdrift, ydrift = rng.normal(0, 0.4, 3), rng.normal(0, 0.4, 3)
    for trial in range(200):
        roi_id = int(rng.integers(1, ntile + 1))
        frame = int(rng.integers(1, 4))
        side = rs["Xsize"][roi_id - 1]
        found = rand_set(rng, int(rng.integers(0, 9)), -1.0, side + 1.0)
        smd = rand_set(rng, int(rng.integers(0, 9)), -1.0, side + 1.0)
        smdbg = rand_set(rng, int(rng.integers(0, 5)), -1.0, side + 1.0)
        # emitters exactly on the window edges
        if found["X"].size:
            found["X"][0] = 3.0
        if smd["X"].size:
            smd["X"][0] = 3.0 + xdrift[frame - 1]
        cut = (0.1, 50.0) if trial % 2 else None
        w = refine.tile_window(rs, roi_id)
        fm, k, b = refine.trim(found, smd, smdbg, w, rs, xdrift[frame - 1], ydrift[frame - 1],
                               *(cut if cut else (None, None)))
        F, K, B, corner = matlab_mcmc_tail(found, smd, smdbg, rs, roi_id, frame, xdrift, ydrift, cut)
        for f in refine.SMD_FIELDS:
            np.testing.assert_array_equal(fm[f], np.float32(F[f]), err_msg=f"found {f} trial {trial}")
            np.testing.assert_array_equal(k[f], np.float32(K[f]), err_msg=f"kmeans {f}")
            np.testing.assert_array_equal(b[f], np.float32(B[f]), err_msg=f"bg {f}")
        assert fm["ROIcorner"] == corner


def chain_from(states):
    """cRJMCMC Chain entries of fixed-dimension states, plus the records the oracle reads."""
    chain, rec, em = [], [], []
    for X, Y, I in states:
        chain.append({"X": np.asarray(X, np.float64), "Y": np.asarray(Y, np.float64),
                      "Photons": np.asarray(I, np.float64), "N": float(len(X))})
        rec.append((len(X), 0, 1, len(em), 0.0, 0.0, 10.0, 0.0, 0.0))
        em.extend([(x, y, i, 1.0) for x, y, i in zip(X, Y, I)])
    chain.append({"X": None})
    from bamf_amd.rjmcmc_abi import CHAIN_REC_DTYPE
    return chain, np.array(rec, CHAIN_REC_DTYPE), np.array(em, np.float32)


def test_found_map_matches_oracle_summary():
    rng = np.random.default_rng(9)
    for L, k in ((41, 3), (40, 1), (7, 5)):
        base = rng.uniform(2, 14, (k, 3)) * [1, 1, 100]
        states = [(base[:, 0] + rng.normal(0, .05, k), base[:, 1] + rng.normal(0, .05, k),
                   base[:, 2] + rng.normal(0, 20, k)) for _ in range(L - 1)]
        states = [tuple(np.float32(a) for a in s) for s in states]
        chain, rec, em = chain_from(states)
        got = refine.found_map_from_chain(chain)
        n, want = ofm.chain_summary(rec, em, len(chain))
        assert n == k
        for c, f in enumerate(refine.SMD_FIELDS):
            np.testing.assert_allclose(got[f], want[:, c], rtol=2e-6, atol=1e-6, err_msg=f)
    # states of different N: the reference's cat fails
    chain, _, _ = chain_from([([1.0], [1.0], [5.0]), ([1.0, 2.0], [1.0, 2.0], [5.0, 6.0])] * 2)
    try:
        refine.found_map_from_chain(chain)
        raise AssertionError("expected an error")
    except Exception as e:
        assert "differ in N" in str(e)


def test_clust_info_assembly_order_and_fields():
    rs = roi_struct(ypix=40, xpix=40, nframes=2)  # 3 x 3 tiles per frame
    n = len(rs["Xsize"])
    rng = np.random.default_rng(2)
    res = {"mf": np.zeros(n, int), "smd": [], "smdbg": [], "found": [None] * n, "stat_rj": [None] * n,
           "stat_mc": [None] * n}
    stat = {k: float(rng.uniform()) for k in refine.STAT_NAMES}
    for t in range(n):
        side = rs["Xsize"][t]
        s = rand_set(rng, 2, 4.0, min(side, 16) - 1.0)
        s.update({"Bg": 10.0 + t, "Bg_SE": 0.1, "ABg": 0.01 * t, "BBg": -0.01 * t})
        res["smd"].append(s)
        res["smdbg"].append(rand_set(rng, 1, 4.0, min(side, 16) - 1.0))
        if t % 3 != 1:
            res["mf"][t] = 2
            res["found"][t] = {k: np.asarray(v, np.float32) for k, v in rand_set(rng, 2, 4.0, 12.0).items()}
            res["stat_rj"][t] = stat
            res["stat_mc"][t] = dict(stat, Accept_JumpFg=0.5 + t)
    ci = refine.assemble_clust_info(res, rs, {}, loop_num=2, temp_frame=2)
    assert set(ci) == set(refine.CLUST_FIELDS)
    assert np.array_equal(ci["Bg"], np.float32([10.0 + t for t in range(n)]))  # every tile
    kept = [t for t in range(n) if t % 3 != 1]
    assert np.array_equal(ci["JumpMC"], np.float32([0.5 + t for t in kept]))
    # frame numbers pp + (LoopNum - 1) * TempFrame, ROInum = LoopNum
    nf = [refine.trim(res["found"][t], res["smd"][t], res["smdbg"][t],
                      refine.tile_window(rs, t - (rs["FrameNum"][t] - 1) * 9 + 1), rs)[0]["X"].size for t in kept]
    want_frames = np.concatenate([np.full(c, rs["FrameNum"][t] + 2, np.float32) for c, t in zip(nf, kept)])
    assert np.array_equal(ci["FrameNum"], want_frames) and np.all(ci["ROInum"] == 2)
    assert ci["X"].size == sum(nf) == ci["ROIcornerX"].size
    assert ci["JumpXRJ"].size == len(kept) and np.all(ci["JumpXRJ"] == np.float32(stat["Accept_Jump"]))
