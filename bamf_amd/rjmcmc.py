"""Host mirror of the chain interfaces: @RJ/rjmcmc.m (one ROI, the cRJMCMC
mex argument layout) and a batched runner over many ROIs.

All chains run on the device through bamf_rjmcmc_run; there is no host
fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _hip
from ._lib import BamfError, check, lib
from .rjmcmc_abi import (CHAIN_REC_DTYPE, LEARNED, PLAIN, REPLAY, STEP_DTYPE, McmcParams, Replay, RjOut, RoiBatch,
                         Tuning)

COUNTER_NAMES = [f"{k}_{j}" for k in ("N", "A", "NFg", "AFg", "NBg", "ABg")
                 for j in ("jump", "split", "merge", "birth", "death")]


def make_params(psf_sigma=1.3, n_burnin=3000, n_trials=3000, p_burnin=(0.4, 0.15, 0.15, 0.15, 0.15, 0, 0, 0),
                p_trials=(0.6, 0.1, 0.1, 0.1, 0.1, 0, 0, 0), rho=3 / 256, x_std_fg=0.1, i_std_fg=10.0, bg_std=1.0,
                split_std=0.5, bnd_out=2.0, abbg_std=0.02, dx=10.0, grid_zoom=20, icutoff=0.0, is_backg=0,
                drift_x=0.0, drift_y=0.0) -> McmcParams:
    """The MCMC struct as @RJ/rjmcmc.m:49-78 fills it (I_stdBg = IstdFg/2, X_stdBg = XstdFg/2,
    Split_std 0.5, Bnd_out 2, ABBG_std 0.02, DX = P_dP)."""
    p = McmcParams()
    p.psf_sigma = psf_sigma
    p.n_burnin = n_burnin
    p.n_trials = n_trials
    p.i_std_fg = i_std_fg
    p.i_std_bg = np.float32(i_std_fg) / np.float32(2)
    p.x_std_fg = x_std_fg
    p.x_std_bg = np.float32(x_std_fg) / np.float32(2)
    for i in range(8):
        p.p_burnin[i] = p_burnin[i]
        p.p_trials[i] = p_trials[i]
    p.rho = rho
    p.bnd_out = bnd_out
    p.grid_zoom = grid_zoom
    p.split_std = split_std
    p.icutoff = icutoff
    p.bg_std = bg_std
    p.abbg_std = abbg_std
    p.dx = dx
    p.drift_x = drift_x
    p.drift_y = drift_y
    p.is_backg = is_backg
    return p


class Plan:
    """bamf_rjmcmc_plan: parameters plus the three prior tables on the device."""

    def __init__(self, params: McmcParams, off_pdf, sig_pdf, back_pdf, sampled_psf=None):
        """sampled_psf: a 4-D SampledPSF array (y, x, phase_y, phase_x), the chains then
        using NumModel (OptModel = 1, LibRJMCMC.cpp:62-170) for the model image."""
        self.params = params
        self._tables = [np.ascontiguousarray(t, np.float32) for t in (off_pdf, sig_pdf, back_pdf)]
        h = ctypes.c_void_p()
        o, s, b = self._tables
        check(lib().bamf_rjmcmc_plan_create(ctypes.byref(params), o.ctypes.data, len(o), s.ctypes.data, len(s),
                                            b.ctypes.data, len(b), ctypes.byref(h)), "bamf_rjmcmc_plan_create")
        self.handle = h
        self._proposal = None
        if sampled_psf is not None:
            self.set_sampled_psf(sampled_psf)

    def set_sampled_psf(self, psf) -> None:
        """OptModel = 1 with SampledPSF = psf (4-D, MATLAB index order); None: the erf model."""
        if psf is None:
            check(lib().bamf_rjmcmc_set_sampled_psf(self.handle, None, None), "bamf_rjmcmc_set_sampled_psf")
            return
        a = np.asarray(psf, np.float32)
        if a.ndim != 4:
            raise BamfError("the sampled PSF is a 4-D array (y, x, phase_y, phase_x)")
        flat = np.ascontiguousarray(a.ravel(order="F"))
        dims = np.array(a.shape, np.int32)
        check(lib().bamf_rjmcmc_set_sampled_psf(self.handle, flat.ctypes.data, dims.ctypes.data),
              "bamf_rjmcmc_set_sampled_psf")

    def set_proposal(self, d_q, d_mu, alpha: float = 0.2, halfwidth: float = 0.35, beta: float = 0.5) -> None:
        """Learned proposal maps (device buffers in the batch's pixel layout)."""
        check(lib().bamf_rjmcmc_set_proposal(self.handle, d_q.vp if d_q is not None else None,
                                             d_mu.vp if d_mu is not None else None, ctypes.c_float(alpha),
                                             ctypes.c_float(halfwidth), ctypes.c_float(beta)),
              "bamf_rjmcmc_set_proposal")
        self._proposal = (d_q, d_mu)  # keep the buffers alive while the plan uses them

    def autotune(self, rois: "RoiSet", sample_jumps: int = 400, stream=None) -> dict:
        """Times the launch settings on a short run of `rois`; the fastest stays in the plan."""
        t = Tuning()
        check(lib().bamf_rjmcmc_autotune(self.handle, ctypes.byref(rois.batch), int(sample_jumps), ctypes.byref(t),
                                         stream.vp if stream else None), "bamf_rjmcmc_autotune")
        return {"chains_per_simd": list(t.chains_per_simd), "concurrent": int(t.concurrent), "ms": float(t.ms),
                "small_max_slots": int(t.small_max_slots)}

    def set_tuning(self, chains_per_simd=(4, 4), concurrent: int = 2, small_max_slots: int = 4) -> None:
        """concurrent: 2 both size classes in one launch (4 chains per SIMD), 1 a launch per
        class side by side, 0 one after the other; chains_per_simd per class (3, 4 or 5);
        small_max_slots: ROIs of at most this many pixel slots per lane run in the 4-slot
        class (1..4), the rest in the 16-slot class."""
        t = Tuning()
        t.chains_per_simd[0], t.chains_per_simd[1] = int(chains_per_simd[0]), int(chains_per_simd[1])
        t.concurrent = int(concurrent)
        t.small_max_slots = int(small_max_slots)
        check(lib().bamf_rjmcmc_set_tuning(self.handle, ctypes.byref(t)), "bamf_rjmcmc_set_tuning")

    def tuning(self) -> dict:
        t = Tuning()
        check(lib().bamf_rjmcmc_get_tuning(self.handle, ctypes.byref(t)), "bamf_rjmcmc_get_tuning")
        return {"chains_per_simd": list(t.chains_per_simd), "concurrent": int(t.concurrent),
                "small_max_slots": int(t.small_max_slots)}

    def set_intensity_proposal(self, kind: str = "window", inflation: float = 1.2, single_share: float = 0.25) -> None:
        """Learned mode's independence intensity proposal: "window" (the network's log-uniform
        window) or "fit" (Gaussian fit of the emitter's conditional likelihood); single_share is
        the share of moves that are single-emitter moves (0.25 as the reference, up to 0.5)."""
        k = {"window": 0, "fit": 1}[kind]
        check(lib().bamf_rjmcmc_set_intensity_proposal(self.handle, k, ctypes.c_float(inflation),
                                                       ctypes.c_float(single_share)),
              "bamf_rjmcmc_set_intensity_proposal")

    def set_split_proposal(self, share: float = 0.5, radius: float = 2.6) -> None:
        """Learned mode's split: with probability share the first child is placed in a pixel
        drawn from the occupancy map within +-radius pixels of the parent (both split and
        merge ratios score the mixture density); share 0 is the reference's split."""
        check(lib().bamf_rjmcmc_set_split_proposal(self.handle, ctypes.c_float(share), ctypes.c_float(radius)),
              "bamf_rjmcmc_set_split_proposal")

    def set_segments(self, n_segments: int, graph: bool = False) -> None:
        """Run every chain as n_segments consecutive launches (optionally replayed from a HIP graph)."""
        check(lib().bamf_rjmcmc_set_segments(self.handle, int(n_segments), 1 if graph else 0),
              "bamf_rjmcmc_set_segments")

    def __del__(self):
        try:
            if self.handle:
                lib().bamf_rjmcmc_plan_destroy(self.handle)
        except Exception:
            pass


@dataclass
class BatchResult:
    status: np.ndarray
    counters: np.ndarray
    n_final: np.ndarray
    final: np.ndarray
    hist: np.ndarray
    logpost: np.ndarray | None
    trace_n: np.ndarray | None
    trace_photons: np.ndarray | None
    steps: np.ndarray | None
    pgrid: np.ndarray | None = None
    pback: np.ndarray | None = None
    alg_flops: np.ndarray | None = None


class RoiSet:
    """A device-resident batch of ROIs (bamf_roi_batch)."""

    def __init__(self, data: np.ndarray | None, offsets: np.ndarray, xsize, ysize, rho=None, bg0=None, keys=None,
                 order=None, scmos=None, n0=None, em0=None, device=None):
        """device = (d_data, d_offsets, d_xsize, d_ysize) DeviceBuffers already
        holding the batch (bamf_make_subregions output); data is then None and
        offsets/xsize/ysize are host copies of what they hold."""
        self.n_roi = len(xsize)
        self.xsize = np.ascontiguousarray(xsize, np.int32)
        self.ysize = np.ascontiguousarray(ysize, np.int32)
        self.offsets = np.ascontiguousarray(offsets[: self.n_roi], np.int64)
        self.total_pixels = int(np.max(self.offsets + self.xsize.astype(np.int64) * self.ysize)) if self.n_roi else 0
        if np.any(self.xsize < 1) or np.any(self.xsize > 32) or np.any(self.ysize < 1) or np.any(self.ysize > 32):
            raise BamfError("ROI sides must be in [1, 32]")
        up = _hip.DeviceBuffer.from_array
        if device is not None:
            self.d_data, self.d_off, self.d_xs, self.d_ys = device
        else:
            self.d_data = up(np.ascontiguousarray(data, np.float32))
            self.d_off = up(self.offsets)
            self.d_xs = up(self.xsize)
            self.d_ys = up(self.ysize)
        self.bufs = {}
        b = RoiBatch()
        b.n_roi = self.n_roi
        b.d_pix_offset, b.d_xsize, b.d_ysize, b.d_data = self.d_off.ptr, self.d_xs.ptr, self.d_ys.ptr, self.d_data.ptr
        for name, arr, dt in (("d_rho", rho, np.float32), ("d_bg0", bg0, np.float32), ("d_key", keys, np.uint32),
                              ("d_order", order, np.int32), ("d_scmos", scmos, np.float32)):
            if arr is not None:
                self.bufs[name] = up(np.ascontiguousarray(arr, dt))
                setattr(b, name, self.bufs[name].ptr)
        if n0 is not None:
            i0, x0, y0, eoff = em0
            for name, arr, dt in (("d_n0", n0, np.int32), ("d_em0_offset", eoff, np.int64), ("d_i0", i0, np.float32),
                                  ("d_x0", x0, np.float32), ("d_y0", y0, np.float32)):
                self.bufs[name] = up(np.ascontiguousarray(arr, dt))
                setattr(b, name, self.bufs[name].ptr)
        self.batch = b

    @classmethod
    def from_synth(cls, rois: dict, **kw):
        return cls(rois["data"], rois["offset"], rois["side"], rois["side"], **kw)


def lpt_order(xsize: np.ndarray, ysize: np.ndarray, expected_n: np.ndarray | None = None) -> np.ndarray:
    """Longest-processing-time-first queue order: ROIs sorted by estimated cost
    (pixels, plus the emitter count's share of stencil work), largest first."""
    cost = (xsize.astype(np.float64) * ysize)
    if expected_n is not None:
        cost = cost * (1.0 + 0.05 * np.asarray(expected_n, np.float64))
    return np.argsort(-cost, kind="stable").astype(np.int32)


def run_batch(plan: Plan, rois: RoiSet, seed: int = 0, mode: int = PLAIN, replay=None, *, hist_len: int = 64,
              final_cap: int = 128, trace_len: int = 0, trace_thin: int = 1, record_steps: bool = False,
              grids: bool | dict = False, chain_caps: tuple | None = None, stream=None, sync: bool = True,
              outs=None) -> BatchResult | dict:
    """Runs every chain of the batch; returns host copies of the outputs.
    chain_caps = (records, emitter entries) per ROI turns the Chain /
    ProposedChain records on.  grids: True for one zoomed grid per ROI, or a
    placement (bamf_amd.tiles.Tiling.placement) for one stitched image."""
    n = rois.n_roi
    p = plan.params
    total = p.n_burnin + p.n_trials
    if outs is None:
        outs = alloc_outputs(n, total, hist_len, final_cap, trace_len, record_steps, grids, rois, p.grid_zoom,
                             chain_caps)
    o = outs["struct"]
    o.trace_thin = trace_thin
    rp = None
    if mode == REPLAY:
        rp = replay["struct"]
    check(lib().bamf_rjmcmc_run(plan.handle, ctypes.byref(rois.batch), int(seed), int(mode),
                                ctypes.byref(rp) if rp is not None else None, ctypes.byref(o),
                                stream.vp if stream else None), "bamf_rjmcmc_run")
    if not sync:
        return outs
    _hip.synchronize()
    return fetch(outs, n, total, hist_len, final_cap, trace_len)


def alloc_outputs(n, total, hist_len, final_cap, trace_len, record_steps, grids, rois, zoom, chain_caps=None,
                  logpost=True):
    """logpost=False leaves out the accepted-jump log-PR total (a diagnostic with no counterpart
    in the reference), and with it the double-precision log of each accepted jump."""
    B = _hip.DeviceBuffer
    d = {"status": B(4 * n), "counters": B(4 * 30 * n), "final": B(4 * n * (4 + 4 * final_cap)),
         "hist": B(4 * n * hist_len), "flops": B(8 * n)}
    o = RjOut()
    o.d_status, o.d_counters, o.d_final, o.final_cap = d["status"].ptr, d["counters"].ptr, d["final"].ptr, final_cap
    o.d_n_hist, o.hist_len = d["hist"].ptr, hist_len
    if logpost:
        d["logpost"] = B(8 * n)
        o.d_logpost = d["logpost"].ptr
    o.d_alg_flops = d["flops"].ptr
    if trace_len > 0:
        d["trace_n"] = B(2 * n * trace_len)
        d["trace_ph"] = B(4 * n * trace_len)
        o.d_trace_n, o.d_trace_photons, o.trace_len = d["trace_n"].ptr, d["trace_ph"].ptr, trace_len
    if record_steps:
        d["steps"] = B(STEP_DTYPE.itemsize * n * total)
        o.d_steps = d["steps"].ptr
    if isinstance(grids, dict):
        # the tiles' grids placed in one stitched image (bamf_amd.tiles.Tiling.placement)
        npx = int(grids["pixels"])
        d["pgrid"] = B.zeros(4 * npx)
        d["pback"] = B.zeros(4 * npx)
        o.d_pgrid, o.d_pback = d["pgrid"].ptr, d["pback"].ptr
        o.d_grid_offset, o.d_grid_place = grids["offset"].ptr, grids["place"].ptr
        d["grid_placement"] = grids
    elif grids:
        sizes = rois.xsize.astype(np.int64) * rois.ysize * zoom * zoom
        goff = np.zeros(n, np.int64)
        goff[1:] = np.cumsum(sizes)[:-1]
        d["pgrid"] = B.zeros(4 * int(sizes.sum()))
        d["pback"] = B.zeros(4 * int(sizes.sum()))
        d["goff"] = B.from_array(goff)
        o.d_pgrid, o.d_pback, o.d_grid_offset = d["pgrid"].ptr, d["pback"].ptr, d["goff"].ptr
        d["grid_sizes"] = sizes
        d["grid_off_host"] = goff
    if chain_caps is not None:
        rc, ec = (int(v) for v in chain_caps)
        for k in ("acc_rec", "prop_rec"):
            d[k] = B(CHAIN_REC_DTYPE.itemsize * n * max(rc, 1))
        for k in ("acc_em", "prop_em"):
            d[k] = B(16 * n * max(ec, 1))
        d["rec_count"] = B(8 * 3 * n)
        o.d_acc_rec, o.d_prop_rec = d["acc_rec"].ptr, d["prop_rec"].ptr
        o.d_acc_em, o.d_prop_em, o.d_rec_count = d["acc_em"].ptr, d["prop_em"].ptr, d["rec_count"].ptr
        o.rec_cap, o.em_cap = rc, ec
        d["chain_caps"] = (rc, ec)
    d["struct"] = o
    return d


def fetch(d, n, total, hist_len, final_cap, trace_len) -> BatchResult:
    fin = d["final"].download(np.float32).reshape(n, 4 + 4 * final_cap)
    res = BatchResult(
        status=d["status"].download(np.int32),
        counters=d["counters"].download(np.int32).reshape(n, 30),
        n_final=fin[:, 0].astype(np.int32), final=fin,
        hist=d["hist"].download(np.int32).reshape(n, hist_len),
        logpost=d["logpost"].download(np.float64) if "logpost" in d else None,
        trace_n=d["trace_n"].download(np.int16).reshape(n, trace_len) if "trace_n" in d else None,
        trace_photons=d["trace_ph"].download(np.float32).reshape(n, trace_len) if "trace_ph" in d else None,
        steps=d["steps"].download(STEP_DTYPE).reshape(n, total) if "steps" in d else None)
    res.alg_flops = d["flops"].download(np.uint64)
    if "pgrid" in d:
        res.pgrid = d["pgrid"].download(np.float32)
        res.pback = d["pback"].download(np.float32)
    if "chain_caps" in d:
        rc, ec = d["chain_caps"]
        res.chain = {"count": d["rec_count"].download(np.int64).reshape(n, 3),
                     "acc_rec": d["acc_rec"].download(CHAIN_REC_DTYPE)[:n * rc].reshape(n, rc),
                     "prop_rec": d["prop_rec"].download(CHAIN_REC_DTYPE)[:n * rc].reshape(n, rc),
                     "acc_em": d["acc_em"].download(np.float32)[:n * ec * 4].reshape(n, ec, 4),
                     "prop_em": d["prop_em"].download(np.float32)[:n * ec * 4].reshape(n, ec, 4)}
    return res


def find_map(outs: dict, n_roi: int, fb: int = 1, cap: int = 128, stream=None) -> dict:
    """findMap (@RJ/findMap.m) for every ROI of a run whose outputs (from
    alloc_outputs with chain_caps) are still on the device.  Returns
    MostFrequent, the background terms and per-cluster X, Y, Photons means
    and standard deviations (SMD.X .. SMD.I_SE)."""
    if "chain_caps" not in outs:
        raise BamfError("find_map needs the run's chain records (chain_caps)")
    rc, ec = outs["chain_caps"]
    B = _hip.DeviceBuffer
    st, nm, bg, em = B(4 * n_roi), B(4 * n_roi), B(16 * n_roi), B(24 * n_roi * cap)
    check(lib().bamf_find_map(outs["acc_rec"].vp, outs["acc_em"].vp, outs["rec_count"].vp, rc, ec, int(n_roi),
                              int(fb), int(cap), st.vp, nm.vp, bg.vp, em.vp, stream.vp if stream else None),
          "bamf_find_map")
    _hip.synchronize()
    return {"status": st.download(np.int32), "n": nm.download(np.int32),
            "bg": bg.download(np.float32).reshape(n_roi, 4), "em": em.download(np.float32).reshape(n_roi, cap, 6)}


def chain_summary(outs: dict, n_roi: int, cap: int = 128, stream=None) -> dict:
    """The refinement summary of mcmc.m:144-166 (fixed-dimension chains): per
    emitter index, means and standard deviations of X, Y and Photons."""
    if "chain_caps" not in outs:
        raise BamfError("chain_summary needs the run's chain records (chain_caps)")
    rc, ec = outs["chain_caps"]
    B = _hip.DeviceBuffer
    st, nn, em = B(4 * n_roi), B(4 * n_roi), B(24 * n_roi * cap)
    check(lib().bamf_chain_summary(outs["acc_rec"].vp, outs["acc_em"].vp, outs["rec_count"].vp, rc, ec, int(n_roi),
                                   int(cap), st.vp, nn.vp, em.vp, stream.vp if stream else None), "bamf_chain_summary")
    _hip.synchronize()
    return {"status": st.download(np.int32), "n": nn.download(np.int32),
            "em": em.download(np.float32).reshape(n_roi, cap, 6)}


def make_replay(tapes: list[np.ndarray], ref_steps: np.ndarray) -> dict:
    """Device replay input from per-ROI raw draw tapes and the recorded step records."""
    lens = np.array([len(t) for t in tapes], np.int64)
    offs = np.zeros(len(tapes), np.int64)
    offs[1:] = np.cumsum(lens)[:-1]
    flat = np.concatenate(tapes).astype(np.uint32) if len(tapes) else np.zeros(1, np.uint32)
    ref = np.ascontiguousarray(ref_steps)
    if ref.dtype != STEP_DTYPE:
        ref = ref.astype(STEP_DTYPE)
    B = _hip.DeviceBuffer.from_array
    d = {"tape": B(flat), "off": B(offs), "len": B(lens), "ref": B(ref)}
    r = Replay()
    r.d_tape, r.d_tape_offset, r.d_tape_len, r.d_ref_steps = d["tape"].ptr, d["off"].ptr, d["len"].ptr, d["ref"].ptr
    d["struct"] = r
    return d


STAT_NAMES = ["NTrials", "Accept_Jump", "Accept_Split", "Accept_Merge", "Accept_Birth", "Accept_Death",
              "Accept_JumpFg", "Accept_SplitFg", "Accept_MergeFg", "Accept_BirthFg", "Accept_DeathFg",
              "Accept_JumpBg", "Accept_SplitBg", "Accept_MergeBg", "Accept_BirthBg", "Accept_DeathBg"]
CHAIN_FIELDS = ["N", "X", "Y", "Photons", "Signal", "LLR", "PR", "JumpType", "Accepted", "BG", "ABG", "BBG"]


def _chain_entries(rec, meta, em, nrec):
    """The Chain struct array of cRJMCMC.cpp:159-229: NLoop entries, the last
    left empty; N is 0 for the empty state (one placeholder emitter of zero
    photons); LLR, PR, JumpType and Accepted are the accepted chain's for both
    outputs (cRJMCMC.cpp:265-272)."""
    out = []
    for k in range(nrec - 1):
        h, m = rec[k], meta[k]
        nn = max(int(h["n"]), 1)
        e = em[int(h["em_index"]):int(h["em_index"]) + nn].astype(np.float64)
        out.append({"N": 0.0 if e[0, 2] == 0 else float(h["n"]), "X": e[:, 0], "Y": e[:, 1], "Photons": e[:, 2],
                    "Signal": e[:, 3], "LLR": float(m["llr"]), "PR": float(m["pr"]),
                    "JumpType": float(m["jump_type"]), "Accepted": float(m["accepted"]), "BG": float(h["bg"]),
                    "ABG": float(h["abg"]), "BBG": float(h["bbg"])})
    if nrec > 0:
        out.append({f: None for f in CHAIN_FIELDS})
    return out


def crjmcmc(params: McmcParams, data: np.ndarray, xsize: int, ysize: int, e_active: dict, off_pdf, sig_pdf,
            back_pdf, scmos=None, sampled_psf=None, opt_model: int = 0, seed: int = 0):
    """cRJMCMC (cRJMCMC.cpp:52-305) for one ROI through the host gateway
    bamf_crjmcmc.  data: Ysize x Xsize (MATLAB orientation; any numpy order is
    read column-major); e_active: N, I, X, Y, BG.  Returns (PGrid, PBack, Chain,
    ProposedChain, Stat): PGrid and PBack (Ysize*Zoom, Xsize*Zoom) float32,
    the chains as lists of NLoop dicts with cRJMCMC's fields, Stat a dict."""
    d = np.asarray(data, np.float32)
    if d.size != xsize * ysize:
        raise BamfError("Data must hold Xsize * Ysize pixels")
    flat = np.asarray(d, order="F").ravel(order="F")
    n0 = int(e_active.get("N", 0))
    f32 = lambda v: np.ascontiguousarray(np.asarray(v, np.float32).ravel())  # noqa: E731
    i0, x0, y0 = (f32(e_active.get(k, [])) for k in ("I", "X", "Y"))
    if min(i0.size, x0.size, y0.size) < n0:
        raise BamfError("E_active.I, X and Y must hold N entries")
    off, sig, back = f32(off_pdf), f32(sig_pdf), f32(back_pdf)
    sc = None if scmos is None or np.size(scmos) == 0 else np.asarray(scmos, np.float32).ravel(order="F")
    if sc is not None and sc.size != flat.size:
        raise BamfError("sCMOScorr must be empty or hold Xsize * Ysize pixels")
    psf_flat = psf_size = None
    if sampled_psf is not None:
        sp = np.asarray(sampled_psf, np.float32)
        psf_size = np.array(list(sp.shape) + [1] * (4 - sp.ndim), np.int32)
        psf_flat = np.ascontiguousarray(sp.ravel(order="F"))
    z = int(params.grid_zoom)
    pg = np.zeros(xsize * ysize * z * z, np.float32)
    pb = np.zeros_like(pg)
    total = int(params.n_burnin) + int(params.n_trials)
    rec_cap, em_cap = total + 1, total * 16 + 1024
    acc, prop = np.zeros(rec_cap, CHAIN_REC_DTYPE), np.zeros(rec_cap, CHAIN_REC_DTYPE)
    ae, pe = np.zeros((em_cap, 4), np.float32), np.zeros((em_cap, 4), np.float32)
    counts = np.zeros(3, np.int64)
    stat = np.zeros(16, np.float64)
    ptr = lambda a: None if a is None or a.size == 0 else a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    check(lib().bamf_crjmcmc(ctypes.byref(params), ptr(flat), int(xsize), int(ysize), n0, ptr(i0), ptr(x0), ptr(y0),
                             ctypes.c_float(float(e_active.get("BG", 0.0))), ptr(off), off.size, ptr(sig), sig.size,
                             ptr(back), back.size, ptr(sc), ptr(psf_flat), ptr(psf_size), int(opt_model), int(seed),
                             ptr(pg), ptr(pb), ptr(acc), ptr(prop), ptr(ae), ptr(pe), rec_cap, em_cap, ptr(counts),
                             ptr(stat)), "bamf_crjmcmc")
    if counts[1] > em_cap or counts[2] > em_cap:
        raise BamfError("chain record capacity exceeded")
    nrec = int(counts[0])
    shape = (ysize * z, xsize * z)
    return (pg.reshape(shape, order="F"), pb.reshape(shape, order="F"), _chain_entries(acc, acc, ae, nrec),
            _chain_entries(prop, acc, pe, nrec), dict(zip(STAT_NAMES, stat.tolist())))


def rjmcmc(data: np.ndarray, rjstruct: dict, ysize: int, xsize: int, drift_y: float = 0.0, drift_x: float = 0.0,
           scmos_var=None, sampled_psf_in=None, opt_model_in=None, seed: int = 0):
    """@RJ/rjmcmc.m for one ROI: the MCMC struct, initial state, sCMOS and
    SampledPSF defaults of rjmcmc.m:40-102, then cRJMCMC (crjmcmc).  Returns
    (Chain, Stat, PImage, PBackg, ProposedChain); PImage and PBackg are
    SRZoom times the size of Data with the ROI's block filled
    (rjmcmc.m:95-96,104-105)."""
    d0 = np.asarray(data, np.float32)
    if d0.ndim > 2 and d0.shape[2] > 1:
        raise BamfError("The input data to rjmcmc() must be a single ROI.")
    d0 = d0.reshape(d0.shape[0], -1) if d0.ndim > 2 else np.atleast_2d(d0)
    r = rjstruct
    p_off = np.asarray(r["P_Offset"], np.float64).ravel()
    bg0 = np.float32(r["P_dP"] * int(np.flatnonzero(p_off == p_off.max())[0]))  # rjmcmc.m:46
    params = make_params(psf_sigma=r["PSF_Sigma"], n_burnin=r["N_Burnin"], n_trials=r["N_Trials"],
                         p_burnin=r["P_Burnin"], p_trials=r["P_Trials"], rho=r["Rho"], x_std_fg=r["XstdFg"],
                         i_std_fg=r["IstdFg"], bg_std=r["BGstd"], dx=r["P_dP"], grid_zoom=r["SRZoom"],
                         icutoff=r.get("MinPhoton", 0.0), is_backg=r.get("IsBackg", 0), drift_x=drift_x,
                         drift_y=drift_y)
    d = d0[:ysize, :xsize]
    sc = None  # rjmcmc.m:79-81: a scalar or absent sCMOSVar is zeros (adds nothing to the model)
    if scmos_var is not None and np.ndim(scmos_var) == 2:
        sc = np.asarray(scmos_var, np.float32)[:ysize, :xsize]
    psf, opt = None, 0  # rjmcmc.m:82-85,98-101: no SampledPSF means OptModel 0
    if sampled_psf_in is not None and np.ndim(sampled_psf_in) > 0 and np.size(sampled_psf_in) > 0 and \
            opt_model_in is not None and np.size(opt_model_in) > 0:
        opt = int(np.asarray(opt_model_in).ravel()[0])
        psf = sampled_psf_in if opt == 1 else None
    e_active = {"N": 0, "I": [], "X": [], "Y": [], "BG": bg0}
    pimg, pbg, chain, proposed, stat = crjmcmc(params, d, xsize, ysize, e_active, r["P_Offset"], r["P_EmPhotons"],
                                               r["P_BgEmPhotons"], sc, psf, opt, seed)
    z = int(params.grid_zoom)
    PImage = np.zeros((z * d0.shape[0], z * d0.shape[1]), np.float32)
    PBackg = np.zeros_like(PImage)
    PImage[:z * ysize, :z * xsize] = pimg
    PBackg[:z * ysize, :z * xsize] = pbg
    return chain, stat, PImage, PBackg, proposed
