"""Benchmark: ROI posteriors/sec of batched RJMCMC on one MI355X.

Workload (default, the batched uneven plain configuration): 20,000 synthetic
ROIs with side drawn uniformly from {8, 12, 16, 24, 32} px, emitter density
uniform in 0.5-5 /um^2 at 100 nm pixels, intensity Gamma(5, 200), background
uniform 5-30, sigma 1.3 px; 10,000 jumps per chain (5,000 burn-in + 5,000
post-burn-in) with P_Burnin = [.4 .15 .15 .15 .15 0 0 0] and
P_Trials = [.6 .1 .1 .1 .1 0 0 0].

A step is one batched run of every chain of the workload (the persistent
launch over both ROI size classes, or the launch setting the autotuner picks).  An ROI posterior is an ROI whose
post-burn-in chain reaches ESS >= 200, the ESS being the smaller of the
emitter-count and total-photon traces' ESS (Geyer's initial monotone
sequence estimator).  value = such ROIs per second: the number of ROIs whose
chain reaches ESS 200 within the configured length, over the median step
time (HIP events around each step on the launch stream, inputs resident in
HBM).  Reported beside it: the fixed-ESS rate extrapolated to every ROI
(each ROI's share of the step rescaled to the length its chain would need,
bamf_amd.ess.fixed_ess_rate), the summed-ESS rate, the learned-proposal
chain on the same ROIs, and the posterior agreement of both chains with the
long oracle reference chains of tests/golden/posterior_batched_uneven.npz
(emitter-count KS p and localization RMSE against the oracle's).  The CPU
baseline is the oracle chain (a C restatement of the
reference's RJMCMC class) on a sample of the same ROIs over all host cores,
timed before the GPU is touched.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
SEED = 20261203
# Algorithmic FLOP of one XYNB fit (7 x 7, 20 iterations, separable form):
# init 2.0 k, 20 x 1.95 k per iteration, CRLB + LL 2.55 k, 4 x 4 inverse 0.12 k.
MLE_FLOP_PER_FIT = 43_800
DX = 10.0


def workload(name: str):
    from bamf_amd import synth

    if name == "batched_uneven":
        rois = synth.make_rois(20000, seed=SEED + 2, sides=(8, 12, 16, 24, 32), density=(0.5, 5.0),
                               bg=(5.0, 30.0))
        nb, nt = 5000, 5000
        rho = 2.75 * 0.01  # mean density 2.75 / um^2 at 0.01 um^2 per pixel
    elif name == "plain_16":
        rois = synth.make_rois(256, seed=SEED + 1, sides=(16,), mean_n=3.0, n_cap=10, bg=(10.0, 10.0))
        nb, nt = 3000, 3000
        rho = 3.0 / 256
    else:
        raise SystemExit(f"unknown workload {name}")
    sig = synth.gamma_pdf_table(5, 200, DX, 1001)
    off = synth.gamma_pdf_table(4, 5, DX, 1001)  # offset prior: mean 20, sd 10 photons/px
    back = np.zeros(1001, np.float32)
    bg0 = float(DX * int(np.argmax(off)))  # @RJ/rjmcmc.m:47
    return rois, nb, nt, rho, sig, off, back, bg0


def params(nb, nt, rho):
    from bamf_amd.rjmcmc import make_params

    return make_params(n_burnin=nb, n_trials=nt, rho=rho, p_burnin=(0.4, 0.15, 0.15, 0.15, 0.15, 0, 0, 0),
                       p_trials=(0.6, 0.1, 0.1, 0.1, 0.1, 0, 0, 0))


# ---------------------------------------------------------------- CPU baseline
_CPU = {}


def _cpu_worker(idx):
    from oracle import pyoracle as po

    rois, p, off, sig, back, bg0, psf = _CPU["args"]
    op = po.McmcParams.from_buffer_copy(p)
    out = []
    for r in idx:
        o = rois["offset"]
        s = int(rois["side"][r])
        t0 = time.perf_counter()
        res = po.rjmcmc(rois["data"][o[r]:o[r + 1]], s, s, op, off, sig, back, bg0=bg0, seed=SEED, roi=int(r),
                        record=False, psf=psf)
        dt = time.perf_counter() - t0
        st = res["steps"][p.n_burnin:]
        out.append((int(r), st["n"].copy(), st["photons"].copy(), dt))
    return out


def cpu_baseline(rois, p, off, sig, back, bg0, sample, cores, thin, psf=None):
    import multiprocessing as mp

    from bamf_amd.ess import fixed_ess_rate, roi_posteriors
    from oracle import pyoracle as po

    po.build()
    rng = np.random.default_rng(SEED)
    idx = np.sort(rng.choice(len(rois["side"]), size=sample, replace=False))
    _CPU["args"] = (rois, p, off, sig, back, bg0, psf)
    chunks = [idx[i::cores * 4] for i in range(cores * 4)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        parts = pool.map(_cpu_worker, chunks)
    dt = time.perf_counter() - t0
    res = sorted([x for part in parts for x in part])
    tn = np.stack([r[1][::thin] for r in res])
    tp = np.stack([r[2][::thin] for r in res])
    ess, ok = roi_posteriors(tn, tp)
    jumps = sample * (p.n_burnin + p.n_trials)
    # each ROI's share of the wall time: its own chain time over the worker count
    share = np.array([r[3] for r in res]) * (dt / sum(r[3] for r in res))
    burn = p.n_burnin / (p.n_burnin + p.n_trials)
    return {"value": float(ok.sum() / dt), "unit": "ROI posteriors/s", "cores": cores, "kind": "port",
            "extrapolated_fixed_ess_rate": fixed_ess_rate(ess, share, burn),
            "sample": f"{sample} of the workload's ROIs (seeded random subset), full {p.n_burnin}+{p.n_trials}-jump "
                      f"chains of the oracle C chain, {cores} forked worker processes, {dt:.1f} s wall",
            "chains_per_s": float(sample / dt), "jumps_per_s": float(jumps / dt),
            "strict_posteriors_per_s": float(ok.sum() / dt), "median_ess": float(np.median(ess)),
            "ess_sum_rate": float(ess.sum() / 200.0 / dt)}


# ---------------------------------------------------------------- accuracy
POSTERIOR_REF = ROOT / "tests" / "golden" / "posterior_batched_uneven.npz"


def posterior_agreement(mode, prm=None, k=1024, cap=96, split_share=0.0):
    """The chain's posterior against the long oracle reference chains: k
    independent chains per reference ROI of the reference's length, final
    states as samples; emitter-count KS p-values and the localization RMSE
    against the oracle's (bamf_amd.posterior)."""
    from bamf_amd import _hip, posterior, proposal
    from bamf_amd.rjmcmc import LEARNED, run_batch

    if not POSTERIOR_REF.exists():
        return None
    ref = posterior.load_reference(POSTERIOR_REF)
    plan = posterior.reference_plan(ref)
    rois = posterior.replicated_batch(ref, k)
    keep = None
    if mode == LEARNED:
        mm = proposal.MapMaker(prm)
        dq, dmu = _hip.DeviceBuffer(rois.total_pixels * 4), _hip.DeviceBuffer(rois.total_pixels * 4)
        mm.run(rois, dq, dmu, 0.05)
        _hip.synchronize()
        mm.close()
        plan.set_proposal(dq, dmu, alpha=0.2, halfwidth=0.35, beta=1.0)
        plan.set_intensity_proposal("fit", 1.3, 0.5)
        plan.set_split_proposal(split_share, 2.6)
        keep = (dq, dmu)
    res = run_batch(plan, rois, seed=SEED + 77, mode=mode, final_cap=cap, hist_len=128)
    assert np.all(res.status == 0)
    ns, pos = posterior.final_samples(res.final, len(ref["side"]), k, cap)
    a = posterior.agreement(ref, ns, pos)
    del keep
    return {"reference": f"tests/golden/posterior_batched_uneven.npz (10 ROIs of this workload, "
                         f"{int(ref['chains'])} oracle chains of {int(ref['n_burnin']) + int(ref['n_trials'])} jumps "
                         f"per ROI)", "chains_per_roi": k, "ks_p_min": min(a["ks_p"]),
            "ks_p_pooled": a["ks_p_pooled"], "rmse_nm": a["rmse_nm"], "oracle_rmse_nm": a["oracle_rmse_nm"],
            "rmse_vs_oracle": a["rmse_ratio"] - 1.0}


def localization_rmse_nm(final, truth, cap, pixel_nm=100.0):
    """RMSE of final-state emitters matched one-to-one to the truth (greedy
    nearest pairs within 1 px), in nm."""
    d2 = []
    for r, (I, X, Y) in enumerate(truth):
        n = int(final[r, 0])
        if n == 0 or len(X) == 0:
            continue
        fx = final[r, 4 + cap:4 + cap + n]
        fy = final[r, 4 + 2 * cap:4 + 2 * cap + n]
        dd = (fx[:, None] - X[None, :]) ** 2 + (fy[:, None] - Y[None, :]) ** 2
        used_f, used_t = set(), set()
        for k in np.argsort(dd, axis=None):
            i, j = divmod(int(k), dd.shape[1])
            if dd[i, j] > 1.0:
                break
            if i in used_f or j in used_t:
                continue
            used_f.add(i)
            used_t.add(j)
            d2.append(dd[i, j])
    return float(math.sqrt(np.mean(d2)) * pixel_nm) if d2 else float("nan")


def read_traffic(profile: Path):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    if not profile.exists():
        return None
    vals = {}
    for line in profile.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration(mib: int) -> int:
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE of its dispatches
    give the counters' bytes-per-KiB scale for the traffic figure."""
    from bamf_amd import _hip

    n = mib << 20
    src, dst = _hip.DeviceBuffer.zeros(n), _hip.DeviceBuffer(n)
    s = _hip.Stream()
    for _ in range(3):
        _hip.copy_d2d(dst, src, n, s)
    s.synchronize()
    print(json.dumps({"copy_bytes": n, "dispatches": 3}), flush=True)
    return 0


# ---------------------------------------------------------------- front end
def _time_gpu(fn, stream, warmup, steps):
    from bamf_amd import _hip

    for _ in range(warmup):
        fn()
    stream.synchronize()
    ev0, ev1 = _hip.Event(), _hip.Event()
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    ev1.synchronize()
    return ev0.elapsed_ms(ev1) / steps


def _cpu_pool(fn, chunks, cores):
    import multiprocessing as mp

    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        pool.map(fn, chunks)
    return time.perf_counter() - t0


_FE = {}


def _cpu_findroi(fr):
    from oracle import pyoracle as po

    st, s1, s2, box, mv = _FE["args"]
    for f in fr:
        po.find_roi(st[f], st.shape[2], st.shape[1], 1, s1, s2, box, mv)


def _cpu_mle(rng_):
    from oracle import pyoracle as po

    boxes, sigma, ft, sc = _FE["args"]
    b = boxes[rng_[0]:rng_[1]]
    if sc is not None:  # sCMOS: box corners and the variance / gain^2 map
        po.mle_scmos(("Basic", "Sigma", "SigmaYX").index(ft), b, sc[0][rng_[0]:rng_[1]], sc[1], sigma, 20)
    elif ft == "Basic":
        po.mle_xynb(b, sigma, 20)
    elif ft == "Sigma":
        po.mle_xynbs(b, sigma, 20)
    elif ft == "SigmaYX":
        po.mle_xynbsxsy(b, sigma, 20)
    else:
        z = _ZCAL
        po.mle_xynbz(b, (sigma, z["Ax"], z["Ay"], z["Bx"], z["By"], z["Gamma"], z["D"], sigma), 20)


def bench_find_roi(args):
    """FindROI over 2000 frames of 256 x 256 (Poisson emitters at 0.3 /um^2,
    sigma 1.3, bg 10), sigma1 1.3, sigma2 2.6, Boxsize 7, default Minval."""
    from bamf_amd import _hip, synth
    from bamf_amd.find_roi import FindRoiDevice, default_minval

    nf, size = 2000, 256
    frames = synth.make_frames(nf, size, seed=SEED + 3)
    s1, s2, box, mv = 1.3, 2.6, 7, default_minval()
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po

        po.build()
        cores = args.cpu_cores or min(os.cpu_count() or 1, 16)
        sample = args.cpu_sample_frames
        _FE["args"] = (frames[:sample], s1, s2, box, mv)
        dt = _cpu_pool(_cpu_findroi, [list(range(sample))[i::cores] for i in range(cores)], cores)
        print(f"[bench] cpu baseline: {sample} frames in {dt:.1f} s", file=sys.stderr, flush=True)
        cpu = {"value": sample * size * size / dt, "unit": "pixels/s", "cores": cores, "kind": "port",
               "sample": f"first {sample} frames, oracle C FindROI (restated findROI.m + cuda_FindROI.cu), "
                         f"{cores} forked workers, {dt:.1f} s wall"}
    d = _hip.DeviceBuffer.from_array(frames)
    dev = FindRoiDevice(size, size, nf, capacity=4_000_000)
    s = _hip.Stream()
    ms = _time_gpu(lambda: dev.run(d, s1, s2, box, mv, stream=s), s, args.warmup, args.steps)
    n, *_ = dev.results()
    px = nf * size * size
    alg_bytes = 4.125 * px  # stack read once + 1 bit per pixel of the Res map
    return {"metric": "FindROI pixels/s (2000 x 256 x 256 stack, DoG 1.3/2.6, Boxsize 7)", "value": px / (ms / 1e3),
            "unit": "pixels/s", "ms_per_step": ms, "dtype": "f32", "higher_is_better": True,
            "config": {"workload": "find_roi", "frames": nf, "size": size, "detections": int(n)},
            "roofline": {"bound": "hbm", "achieved": None, "peak": PEAK_HBM_GBS,
                         "unit": "GB/s", "frac": None, "kernel": "whole pipeline (one bamf_find_roi call)",
                         "traffic": read_traffic(ROOT / "profiles" / "findroi_pipeline_find_roi_2000x256x256_pmc.txt"),
                         "alg_bytes_per_launch": alg_bytes},
            "cpu_baseline": cpu}


_ZCAL = {"Ax": 0.05, "Ay": -0.05, "Bx": 0.02, "By": 0.03, "Gamma": 0.35, "D": 0.6}  # astigmatism for 'Z'


def front_end_crops(n: int, sz: int = 7):
    """The first n FindROI boxes of the front-end stack: 2000 frames of 256 x 256
    from the device forward model (emitters at 0.3 /um^2 with 100 nm pixels,
    Gamma(5, 200) photons, background 10, sigma 1.3), FindROI with sigma 1.3 / 2.6,
    Boxsize sz and the default minimum, cropped on the device in find() order.
    Returns the device buffer of the boxes and a host copy."""
    from bamf_amd import _hip
    from bamf_amd._lib import check, lib
    from bamf_amd.find_roi import FindRoiDevice, default_minval

    nf, side = 2000, 256
    ncell = lib().bamf_synth_cells(side, side)
    d_frames = _hip.DeviceBuffer(nf * side * side * 4)
    d_truth = _hip.DeviceBuffer(nf * ncell * 128 * 12)
    d_cnt = _hip.DeviceBuffer(nf * ncell * 4)
    check(lib().bamf_synth_frames(nf, side, side, 0.003, 5.0, 200.0, 10.0, 1.3, SEED + 4, 128, 1, d_frames.vp,
                                  d_truth.vp, d_cnt.vp, None), "bamf_synth_frames")
    dev = FindRoiDevice(side, side, nf, capacity=nf * side * side // 64)
    dev.run(d_frames, 1.3, 2.6, sz, default_minval())
    _hip.synchronize()
    found = dev.results()[0]
    if found < n:
        raise SystemExit(f"the front-end stack yields {found} FindROI boxes, fewer than {n}")
    d_boxes = _hip.DeviceBuffer(n * sz * sz * 4)
    check(lib().bamf_find_roi_crop(d_frames.vp, side, side, dev.row.vp, dev.col.vp, dev.frame.vp, n, sz,
                                   d_boxes.vp, None), "bamf_find_roi_crop")
    _hip.synchronize()
    return d_boxes, d_boxes.download(np.float32).reshape(n, sz, sz), int(found)


def bench_gauss_mle(args):
    """gaussMLE: 100,000 7 x 7 boxes, sigma 1.3, 20 iterations, fit type
    args.fit_type ('Basic' is the hot path's XYNB kernel).  The boxes are the
    first 100,000 FindROI crops of the front-end stack (--mle-boxes crops), or
    single-emitter boxes drawn on the host (--mle-boxes synth)."""
    from bamf_amd import _hip, synth
    from bamf_amd.gauss_mle import gauss_mle_device

    n, sz, sigma = 100_000, 7, 1.3
    if args.mle_boxes == "crops":
        d, boxes, found = front_end_crops(n, sz)
    else:
        boxes, _ = synth.make_mle_boxes(n, sz, seed=SEED + 4, sigma=sigma)
        d, found = None, None
    ft = args.fit_type
    sc_host = None
    if args.camera == "SCMOS":
        if ft == "Z":
            raise SystemExit("--camera SCMOS times the 'Basic', 'Sigma' and 'SigmaYX' fits")
        mapsz = 256
        rng = np.random.default_rng(SEED + 9)
        sc_host = (rng.integers(0, mapsz - sz, (n, 2)).astype(np.float32),
                   rng.uniform(0.5, 3.0, (mapsz, mapsz)).astype(np.float32))
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po

        po.build()
        cores = args.cpu_cores or min(os.cpu_count() or 1, 16)
        sample = args.cpu_sample_fits
        _FE["args"] = (boxes[:sample], sigma, args.fit_type, sc_host)
        step = (sample + cores - 1) // cores
        dt = _cpu_pool(_cpu_mle, [(i, min(sample, i + step)) for i in range(0, sample, step)], cores)
        print(f"[bench] cpu baseline: {sample} fits in {dt:.1f} s", file=sys.stderr, flush=True)
        cpu = {"value": sample / dt, "unit": "fits/s", "cores": cores, "kind": "port",
               "sample": f"first {sample} boxes, oracle C {args.fit_type} fit restatement, {cores} forked workers, "
                         f"{dt:.1f} s wall"}
    if d is None:
        d = _hip.DeviceBuffer.from_array(boxes)
    dp, dc, dl = _hip.DeviceBuffer(24 * n), _hip.DeviceBuffer(24 * n), _hip.DeviceBuffer(4 * n)
    s = _hip.Stream()
    sg = (sigma, sigma) if ft == "Z" else sigma
    zs = _ZCAL if ft == "Z" else None
    scm = None
    if sc_host is not None:
        scm = (_hip.DeviceBuffer.from_array(sc_host[0]), _hip.DeviceBuffer.from_array(sc_host[1]), 256, None)
    ms = _time_gpu(lambda: gauss_mle_device(d, sg, sz, 20, n, dp, dc, dl, stream=s, fit_type=ft, zstruct=zs,
                                            scmos=scm), s, args.warmup, args.steps)
    basic = ft == "Basic" and scm is None
    flops = MLE_FLOP_PER_FIT * n if basic else None  # the algorithmic count is derived for the XYNB fit
    achieved = flops / (ms / 1e3) / 1e12 if basic else None
    return {"metric": f"gaussMLE {ft} {args.camera} fits/s (7 x 7 boxes, 20 iterations)", "value": n / (ms / 1e3),
            "unit": "fits/s", "ms_per_step": ms, "dtype": "f32", "higher_is_better": True,
            "config": {"workload": "gauss_mle", "fit_type": ft, "camera": args.camera, "fits": n, "box": sz,
                       "iterations": 20, "boxes": args.mle_boxes if found is None else
                       f"the first {n} of {found} FindROI crops of the 2000 x 256 x 256 front-end stack"},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / PEAK_FP32_TFLOPS if basic else None, "kernel": "k_mle_xynb",
                         "traffic": read_traffic(ROOT / "profiles" / "mle_xynb_gauss_mle_100k_7x7_pmc.txt")
                         if basic else None, "alg_flop_per_launch": flops},
            "cpu_baseline": cpu}


def bench_find_map(args):
    """findMap over the Chain records of the 256-ROI 16 x 16 workload (3000 + 3000 jumps)."""
    from bamf_amd import _hip
    from bamf_amd.rjmcmc import Plan, RoiSet, alloc_outputs, fetch, run_batch
    from bamf_amd.rjmcmc_abi import CHAIN_REC_DTYPE
    from bamf_amd._lib import check, lib

    rois, nb, nt, rho, sig, off, back, bg0 = workload("plain_16")
    p = params(nb, nt, rho)
    n = len(rois["side"])
    rs = RoiSet.from_synth(rois, bg0=np.full(n, bg0, np.float32))
    caps = (nb + nt + 1, 16 * (nb + nt + 1))
    outs = alloc_outputs(n, nb + nt, 64, 128, 0, False, False, rs, p.grid_zoom, caps)
    run_batch(Plan(p, off, sig, back), rs, seed=SEED, outs=outs, sync=False)
    res = fetch(outs, n, nb + nt, 64, 128, 0)
    assert np.all(res.status == 0)
    ch = res.chain
    # algorithmic bytes: the second half's records and emitter entries, read once
    nloop = ch["count"][:, 0]
    alg_bytes = 0
    for r in range(n):
        lo = (int(nloop[r]) + 1) // 2
        recs = ch["acc_rec"][r, lo - 1:int(nloop[r]) - 1]
        alg_bytes += len(recs) * CHAIN_REC_DTYPE.itemsize + int(np.maximum(recs["n"], 1).sum()) * 16
    cpu = None
    if not args.no_cpu:
        from oracle import findmap as fm

        sample = min(n, 64)
        t0 = time.perf_counter()
        for r in range(sample):
            fm.find_map(ch["acc_rec"][r], ch["acc_em"][r], int(nloop[r]), fb=1)
        dt = time.perf_counter() - t0
        cpu = {"value": sample / dt, "unit": "ROIs/s", "cores": 1, "kind": "port",
               "sample": f"first {sample} ROIs, oracle findMap restatement (numpy), 1 core, {dt:.1f} s"}
    B = _hip.DeviceBuffer
    st, nm, bg, em = B(4 * n), B(4 * n), B(16 * n), B(24 * n * 128)
    s = _hip.Stream()

    def call():
        check(lib().bamf_find_map(outs["acc_rec"].vp, outs["acc_em"].vp, outs["rec_count"].vp, caps[0], caps[1], n,
                                  1, 128, st.vp, nm.vp, bg.vp, em.vp, s.vp), "bamf_find_map")

    ms = _time_gpu(call, s, args.warmup, args.steps)
    return {"metric": "findMap ROIs/s (Chain records of 256 ROIs of 16 x 16, 3000 + 3000 jumps)", "value": n / (ms / 1e3),
            "unit": "ROIs/s", "ms_per_step": ms, "dtype": "f32 (k-means sums fixed point, statistics f64)",
            "higher_is_better": True,
            "config": {"workload": "find_map", "n_roi": n, "records": int(nloop.sum()),
                       "mean_map_n": float(nm.download(np.int32).mean())},
            "roofline": {"bound": "hbm", "achieved": None, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": None,
                         "kernel": "k_find_map", "traffic": None, "alg_bytes_per_launch": float(alg_bytes)},
            "cpu_baseline": cpu}


def _cpu_tiles(fr):
    from oracle import tiles as ot

    data, box, ov = _FE["args"]
    for f in fr:
        ot.make_subregions(data[:, :, f:f + 1], box, ov)
    return len(fr)


def bench_tiles(args):
    """Frame tiling (makeSubRegions.m, box 16, overlap 3) of 2000 frames of
    256 x 256 into the chain batch, with each tile's placement in the stitched
    posterior image at zoom 20."""
    from bamf_amd import _hip
    from bamf_amd.tiles import Tiling

    nf, size, box, ov, zoom = 2000, 256, 16, 3, 20
    rng = np.random.default_rng(SEED + 6)
    data = rng.poisson(10.0, size=(size, size, nf)).astype(np.float32)  # (y, x, f)
    cpu = None
    if not args.no_cpu:
        sample = nf  # the whole workload: about 3 s on one core
        _FE["args"] = (data[:, :, :sample], box, ov)
        dt = _cpu_pool(_cpu_tiles, [list(range(sample))], 1)
        print(f"[bench] cpu baseline: {sample} frames in {dt:.1f} s", file=sys.stderr, flush=True)
        cpu = {"value": sample * size * size / dt, "unit": "pixels/s", "cores": 1, "kind": "port",
               "sample": f"all {sample} frames, oracle numpy makeSubRegions restatement, 1 core, {dt:.1f} s wall"}
    d = _hip.DeviceBuffer.from_array(data.ravel(order="F"))
    t = Tiling(size, size, nf, box, ov)
    s = _hip.Stream()
    ms = _time_gpu(lambda: t.run(d, zoom=zoom, stream=s), s, args.warmup, args.steps)
    px = nf * size * size
    alg_bytes = 4.0 * px + 4.0 * t.npix + t.n * (8 + 4 * 4 + 8 + 20)  # frames in, tiles and metadata out
    return {"metric": "frame tiling pixels/s (2000 x 256 x 256, box 16, overlap 3)", "value": px / (ms / 1e3),
            "unit": "pixels/s", "ms_per_step": ms, "dtype": "f32", "higher_is_better": True,
            "config": {"workload": "tiles", "frames": nf, "size": size, "box": box, "overlap": ov, "tiles": t.n,
                       "tile_pixels": t.npix},
            "roofline": {"bound": "hbm", "achieved": None, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": None,
                         "kernel": "k_tiles", "traffic": None, "alg_bytes_per_launch": alg_bytes},
            "cpu_baseline": cpu}


# ---------------------------------------------------------------- dense low-SNR, end to end
def bench_dense_lowsnr(args):
    """Dense low-SNR stress, end to end: args.dense_frames frames of 512 x 512
    (default 5000) from the device forward model (8 emitters per um^2 at 100 nm
    pixels = 0.08 per pixel^2, Gamma(4, 125) photons (mean 500), background 50,
    sigma 1.3), FindROI with Boxsize 32 (Minval = mean photons / (4 pi sigma^2) / 4,
    the rule of findPriors.m), the first args.dense_rois detections cropped to
    32 x 32 ROIs, then plain and learned-proposal chains of 10,000 + 10,000
    jumps on them.  Generator rule for the per-ROI emitter count: the density is
    the configured one over the whole frame, and a ROI holds whatever emitters
    its 32 x 32 window covers (about 82 on average at this density, more than
    the 40 the configuration's text assumes; the measured counts are reported).
    A step: FindROI over every frame, the crops, and the chains."""
    from bamf_amd import _hip, proposal, synth
    from bamf_amd._lib import check, lib
    from bamf_amd.ess import fixed_ess_rate, roi_posteriors
    from bamf_amd.find_roi import FindRoiDevice
    from bamf_amd.rjmcmc import LEARNED, PLAIN, Plan, RoiSet, alloc_outputs, fetch, make_params, run_batch

    nf, side, box = args.dense_frames, 512, 32
    dens, gk, gt, bg, sigma = 0.08, 4.0, 125.0, 50.0, 1.3
    minval = (gk * gt) / (4 * math.pi * sigma ** 2) / 4
    ncell = int(lib().bamf_synth_cells(side, side))
    ccap = 64
    d_frames = _hip.DeviceBuffer(nf * side * side * 4)
    d_truth = _hip.DeviceBuffer(nf * ncell * ccap * 12)
    d_cnt = _hip.DeviceBuffer(nf * ncell * 4)
    stream = _hip.Stream()
    e0, e1 = _hip.Event(), _hip.Event()
    e0.record(stream)
    check(lib().bamf_synth_frames(nf, side, side, dens, gk, gt, bg, sigma, SEED + 5, ccap, 1, d_frames.vp,
                                  d_truth.vp, d_cnt.vp, stream.vp), "bamf_synth_frames")
    e1.record(stream)
    e1.synchronize()
    print(f"[bench] dense: {nf} frames of {side} x {side} generated in {e0.elapsed_ms(e1):.1f} ms", file=sys.stderr,
          flush=True)
    cap = nf * 1024
    fr = FindRoiDevice(side, side, nf, capacity=cap)
    n_chain = args.dense_rois
    d_rois = _hip.DeviceBuffer(n_chain * box * box * 4)
    offs = np.arange(n_chain + 1, dtype=np.int64) * box * box
    xs = np.full(n_chain, box, np.int32)
    rs = RoiSet(None, offs, xs, xs, bg0=np.full(n_chain, bg, np.float32),
                device=(d_rois, _hip.DeviceBuffer.from_array(offs), _hip.DeviceBuffer.from_array(xs),
                        _hip.DeviceBuffer.from_array(xs)))
    nb = nt = 10000
    p = make_params(n_burnin=nb, n_trials=nt, rho=dens, p_burnin=(0.4, 0.15, 0.15, 0.15, 0.15, 0, 0, 0),
                    p_trials=(0.6, 0.1, 0.1, 0.1, 0.1, 0, 0, 0))
    sig = synth.gamma_pdf_table(gk, gt, DX, 1001)
    off = synth.gamma_pdf_table(25, 2, DX, 1001)  # offset prior: mean 50, sd 10 photons/px
    back = np.zeros(1001, np.float32)
    plan = Plan(p, off, sig, back)
    trace_len, cap_em = nt // args.thin, 128
    outs = alloc_outputs(n_chain, nb + nt, 129, cap_em, trace_len, False, False, rs, p.grid_zoom, logpost=False)

    def front():
        fr.run(d_frames, sigma, 2 * sigma, box, minval, stream=stream)
        check(lib().bamf_find_roi_crop(d_frames.vp, side, side, fr.row.vp, fr.col.vp, fr.frame.vp, n_chain, box,
                                       d_rois.vp, stream.vp), "bamf_find_roi_crop")

    # detections and the truth of the chained ROIs (emitters inside each window)
    front()
    stream.synchronize()
    n_det, _, row, col, frm = fr.results()
    if n_det < n_chain:
        raise SystemExit(f"dense: only {n_det} detections for {n_chain} ROIs")
    row, col, frm = row[:n_chain] - 1, col[:n_chain] - 1, frm[:n_chain] - 1
    fmax = int(frm.max()) + 1
    tr = d_truth.download(np.float32, fmax * ncell * ccap * 3).reshape(fmax, ncell, ccap, 3)
    tc = d_cnt.download(np.int32, fmax * ncell).reshape(fmax, ncell)
    per_frame = [np.concatenate([tr[f, c, :tc[f, c]] for c in range(ncell)]) for f in range(fmax)]
    truth, n_true = [], np.zeros(n_chain, np.int64)
    for b in range(n_chain):
        em = per_frame[frm[b]]
        # a crop's row runs along the frame's contiguous axis (findROI.m passes Xsize = size(Image, 2)):
        # chain y = frame x - row corner, chain x = frame y - column corner
        m = (em[:, 0] >= row[b]) & (em[:, 0] < row[b] + box) & (em[:, 1] >= col[b]) & (em[:, 1] < col[b] + box)
        truth.append((em[m, 2], em[m, 1] - col[b], em[m, 0] - row[b]))
        n_true[b] = m.sum()
    print(f"[bench] dense: {n_det} detections ({n_det / nf:.1f} per frame); {n_chain} ROIs chained, truth "
          f"emitters per ROI mean {n_true.mean():.1f}, max {n_true.max()}", file=sys.stderr, flush=True)

    def timed(mode, pre=None):
        for w in range(args.warmup):
            front()
            if pre:
                pre()
            run_batch(plan, rs, seed=SEED + 2000 + w, mode=mode, trace_thin=args.thin, stream=stream, sync=False,
                      outs=outs, hist_len=129, final_cap=cap_em, trace_len=trace_len)
        stream.synchronize()
        steps, parts = [], []
        for k in range(args.steps):
            ev = [_hip.Event() for _ in range(4)]
            ev[0].record(stream)
            front()
            ev[1].record(stream)
            if pre:
                pre()
            ev[2].record(stream)
            run_batch(plan, rs, seed=SEED + 3000 + k, mode=mode, trace_thin=args.thin, stream=stream, sync=False,
                      outs=outs, hist_len=129, final_cap=cap_em, trace_len=trace_len)
            ev[3].record(stream)
            ev[3].synchronize()
            steps.append(ev[0].elapsed_ms(ev[3]))
            parts.append((ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2]), ev[2].elapsed_ms(ev[3])))
        ms = float(np.median(steps))
        part = np.median(np.array(parts), axis=0)
        res = fetch(outs, n_chain, nb + nt, 129, cap_em, trace_len)
        st = np.unique(res.status, return_counts=True)
        ess, ok = roi_posteriors(res.trace_n, res.trace_photons)
        rmse = localization_rmse_nm(res.final, truth, cap_em)
        cost = res.alg_flops.astype(np.float64)
        share = (part[2] / 1e3) * cost / cost.sum()
        extrap = fixed_ess_rate(ess, share, 0.5)
        print(f"[bench] dense {('plain', 'replay', 'learned')[mode]}: {ms:.1f} ms/step (FindROI+crop {part[0]:.1f}, "
              f"maps {part[1]:.1f}, chains {part[2]:.1f}), status {dict(zip(*[v.tolist() for v in st]))}, "
              f"mean ESS {ess.mean():.2f}, RMSE {rmse:.2f} nm", file=sys.stderr, flush=True)
        return {"value": float(ok.sum() / (ms / 1e3)), "ms_per_step": ms, "find_roi_crop_ms": float(part[0]),
                "maps_ms": float(part[1]), "chain_ms": float(part[2]), "roi_posteriors": int(ok.sum()),
                "extrapolated_fixed_ess_rate": extrap, "mean_ess": float(ess.mean()),
                "median_ess": float(np.median(ess)), "rmse_nm_final_state": rmse,
                "final_n_mean": float(res.n_final.mean()), "status_counts": {int(a): int(b) for a, b in zip(*st)},
                "alg_flop_per_step": float(cost.sum())}

    plain = timed(PLAIN)
    line = {"metric": "ROI posteriors/sec at fixed ESS & RMSE: plain vs learned-proposal RJMCMC",
            "value": plain["value"], "unit": "ROI posteriors/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": plain["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 (accept test f64)", "data": "synthetic (device forward-model frames)",
            "config": {"workload": "dense_lowsnr", "frames": nf, "frame_side": side, "density_per_um2": 8.0,
                       "photons": "Gamma(4, 125)", "bg": bg, "find_roi_box": box, "minval": minval,
                       "detections": int(n_det), "chained_rois": n_chain, "roi_side": box,
                       "truth_emitters_per_roi_mean": float(n_true.mean()), "truth_emitters_per_roi_max": int(n_true.max()),
                       "jumps_per_chain": nb + nt, "burnin": nb, "ess_target": 200},
            "roofline": {"bound": "valu", "achieved": plain["alg_flop_per_step"] / (plain["chain_ms"] / 1e3) / 1e12,
                         "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "kernel": "k_chain (32 x 32 class)",
                         "traffic": None},
            "cpu_baseline": None, "quality": plain}
    line["roofline"]["frac"] = line["roofline"]["achieved"] / PEAK_FP32_TFLOPS
    if args.proposal in ("learned", "both"):
        t0 = time.perf_counter()
        prm, _ = proposal.train(args.train_steps, batch=256, lr=3e-3, seed=SEED % 1000,
                                chain_labels=args.chain_labels)
        train_s = time.perf_counter() - t0
        mm = proposal.MapMaker(prm)
        dq, dmu = _hip.DeviceBuffer(rs.total_pixels * 4), _hip.DeviceBuffer(rs.total_pixels * 4)
        plan.set_proposal(dq, dmu, alpha=0.2, halfwidth=0.35, beta=1.0)
        plan.set_intensity_proposal("fit", 1.3, 0.5)
        plan.set_split_proposal(args.split_share, 2.6)
        learned = timed(LEARNED, pre=lambda: mm.run(rs, dq, dmu, stream=stream))
        learned["train_s"] = train_s
        learned["train_rois"] = args.train_steps * 256
        learned["train_chain_label_share"] = args.chain_labels
        line["learned"] = learned
        line["learned_over_plain"] = learned["value"] / plain["value"] if plain["value"] > 0 else None
        line["learned_over_plain_extrapolated"] = (learned["extrapolated_fixed_ess_rate"] /
                                                   plain["extrapolated_fixed_ess_rate"])
        mm.close()
    print(json.dumps(line), flush=True)
    return 0


def front_end_line(args):
    res = {"find_roi": bench_find_roi, "gauss_mle": bench_gauss_mle, "find_map": bench_find_map,
           "tiles": bench_tiles}[args.workload](args)
    print(f"[bench] {args.workload}: {res['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    if res["roofline"]["bound"] == "hbm":
        r = res["roofline"]
        r["achieved"] = r["alg_bytes_per_launch"] / (res["ms_per_step"] / 1e3) / 1e9
        r["frac"] = r["achieved"] / r["peak"]
    line = {"metric": res["metric"], "value": res["value"], "unit": res["unit"], "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": res["dtype"], "data": "synthetic (seeded forward-model frames/boxes)",
            "config": res["config"], "roofline": res["roofline"], "cpu_baseline": res["cpu_baseline"]}
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the measurement procedure of the workloads: 5 warm-up runs, then the median of 20 timed runs
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="batched_uneven",
                    choices=["batched_uneven", "plain_16", "find_roi", "gauss_mle", "find_map", "tiles",
                             "dense_lowsnr"])
    ap.add_argument("--dense-frames", type=int, default=5000, help="frames of --workload dense_lowsnr")
    ap.add_argument("--dense-rois", type=int, default=20000, help="ROIs chained by --workload dense_lowsnr")
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--cpu-sample-frames", type=int, default=2000)
    ap.add_argument("--cpu-sample-fits", type=int, default=100000)
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch configuration")
    ap.add_argument("--no-posterior", action="store_true",
                    help="skip the posterior comparison with the oracle reference chains")
    ap.add_argument("--thin", type=int, default=5)
    ap.add_argument("--proposal", default="both", choices=["plain", "learned", "both"],
                    help="chain proposal(s) timed: the value is the plain chain's, learned is reported beside it")
    ap.add_argument("--train-steps", type=int, default=4000,
                    help="Adam steps of 256 device-generated ROIs for the learned proposal (4000: 1,024,000 ROIs)")
    ap.add_argument("--segments", type=int, default=0,
                    help="run each chain as this many launches (state in HBM between them)")
    ap.add_argument("--graph", action="store_true", help="with --segments: replay the launches from a HIP graph")
    ap.add_argument("--no-graph-line", action="store_true",
                    help="skip timing the graph-replayed segmented run beside the one-launch run")
    ap.add_argument("--split-share", type=float, default=0.0,
                    help="learned mode: share of splits whose first child is drawn from the occupancy map")
    ap.add_argument("--chain-labels", type=float, default=0.125,
                    help="share of training steps labelled by plain-chain posterior samples instead of the truth")
    ap.add_argument("--fit-type", default="Basic", choices=["Basic", "Sigma", "SigmaYX", "Z"],
                    help="gaussMLE fit type of --workload gauss_mle")
    ap.add_argument("--mle-boxes", default="crops", choices=["crops", "synth"],
                    help="--workload gauss_mle: the first 100,000 FindROI crops of the front-end stack, or "
                         "single-emitter boxes drawn on the host")
    ap.add_argument("--camera", default="CCD", choices=["CCD", "SCMOS"],
                    help="gaussMLE camera of --workload gauss_mle (SCMOS: a 256 x 256 variance/gain^2 map)")
    ap.add_argument("--psf", default="erf", choices=["erf", "sampled"],
                    help="chain PSF model: the erf stencil (OptModel 0) or a sampled 72 x 72 x 4 x 4 table of the "
                         "same Gaussian (OptModel 1, NumModel); sampled runs time the plain chain only")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (the HBM counter calibration run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (ROIs shard with no exchange; "
              "a multi-GPU run would give each GPU its own share of the ROIs). Run with --gpus 1.", file=sys.stderr)
        return 2

    if args.copy_mib:
        return copy_calibration(args.copy_mib)
    if args.workload == "dense_lowsnr":
        return bench_dense_lowsnr(args)
    if args.workload in ("find_roi", "gauss_mle", "find_map", "tiles"):
        return front_end_line(args)
    rois, nb, nt, rho, sig, off, back, bg0 = workload(args.workload)
    p = params(nb, nt, rho)
    psf = None
    if args.psf == "sampled":
        from bamf_amd import synth

        psf = synth.sampled_gauss_psf(72, 4, p.psf_sigma)
        args.proposal = "plain"  # the learned proposals use the erf stencil
    cpu = None
    if not args.no_cpu:
        cores = args.cpu_cores or min(os.cpu_count() or 1, 16)
        print(f"[bench] cpu baseline: {args.cpu_sample} ROIs on {cores} cores", file=sys.stderr, flush=True)
        cpu = cpu_baseline(rois, p, off, sig, back, bg0, args.cpu_sample, cores, args.thin, psf)
        print(f"[bench] cpu baseline done: {cpu['chains_per_s']:.1f} chains/s", file=sys.stderr, flush=True)

    from bamf_amd import _hip, proposal
    from bamf_amd.ess import fixed_ess_rate, roi_posteriors
    from bamf_amd.rjmcmc import LEARNED, PLAIN, Plan, RoiSet, alloc_outputs, fetch, lpt_order, run_batch

    n = len(rois["side"])
    plan = Plan(p, off, sig, back, sampled_psf=psf)
    order = lpt_order(rois["side"], rois["side"], rois["n_true"])
    rs = RoiSet.from_synth(rois, bg0=np.full(n, bg0, np.float32), order=order)
    trace_len = nt // args.thin
    cap = 128
    outs = alloc_outputs(n, nb + nt, 129, cap, trace_len, False, False, rs, p.grid_zoom, logpost=False)
    stream = _hip.Stream()
    tuning = {"chains_per_simd": [4, 4], "concurrent": 2, "small_max_slots": 4, "autotuned": False}
    if not args.no_autotune:
        t0 = time.perf_counter()
        best = plan.autotune(rs, sample_jumps=400, stream=stream)
        tuning = {"chains_per_simd": best["chains_per_simd"], "concurrent": best["concurrent"],
                  "small_max_slots": best["small_max_slots"], "autotuned": True,
                  "sample_ms": best["ms"], "tune_s": time.perf_counter() - t0}
        print(f"[bench] autotune: {tuning}", file=sys.stderr, flush=True)
    tuning["launch"] = {2: "one launch over both size classes", 1: "a launch per size class, side by side",
                        0: "a launch per size class, one after the other"}[tuning["concurrent"]]
    if args.segments > 1:
        # each chain as consecutive launches, its state in HBM between them, optionally
        # the launch sequence captured once in a HIP graph and replayed with new seeds
        plan.set_segments(args.segments, graph=args.graph)
        tuning["segments"] = args.segments
        tuning["graph"] = bool(args.graph)

    def timed(mode, pre=None, pre_events=None):
        """Warm-up, then args.steps timed steps (HIP events on the launch stream)."""
        ev0, ev1 = _hip.Event(), _hip.Event()
        for w in range(args.warmup):
            if pre:
                pre()
            run_batch(plan, rs, seed=SEED + 1000 + w, mode=mode, trace_thin=args.thin, stream=stream, sync=False,
                      outs=outs, hist_len=129, final_cap=cap, trace_len=trace_len)
        stream.synchronize()
        _hip.synchronize()
        evs = [_hip.Event() for _ in range(args.steps + 1)]
        evs[0].record(stream)
        for k in range(args.steps):
            if pre:
                pre()
            run_batch(plan, rs, seed=SEED + k, mode=mode, trace_thin=args.thin, stream=stream, sync=False,
                      outs=outs, hist_len=129, final_cap=cap, trace_len=trace_len)
            evs[k + 1].record(stream)
        evs[-1].synchronize()
        step_ms = np.array([evs[k].elapsed_ms(evs[k + 1]) for k in range(args.steps)])
        ms = float(np.median(step_ms))
        res = fetch(outs, n, nb + nt, 129, cap, trace_len)
        assert np.all(res.status == 0), np.unique(res.status, return_counts=True)
        ess, ok = roi_posteriors(res.trace_n, res.trace_photons)
        rmse = localization_rmse_nm(res.final, rois["truth"], cap)
        print(f"[bench] {('plain', 'replay', 'learned')[mode]}: {ms:.1f} ms/step (median of {args.steps}; "
              f"mean {step_ms.mean():.1f}), mean ESS {ess.mean():.2f}, RMSE {rmse:.2f} nm", file=sys.stderr,
              flush=True)
        # each ROI's share of the launch: its algorithmic work's share
        cost = res.alg_flops.astype(np.float64)
        share = (ms / 1e3) * cost / cost.sum()
        return ms, res, ess, ok, rmse, fixed_ess_rate(ess, share, nb / (nb + nt)), float(step_ms.mean())

    ms, res, ess, ok, rmse, extrap, ms_mean = timed(PLAIN)
    graph_line = None
    if args.segments <= 1 and not args.no_graph_line and args.workload == "batched_uneven":
        # the same chains as 4 launches replayed from a HIP graph (state in HBM between
        # segments): equal results bit for bit, timed beside the one-launch run
        plan.set_segments(4, graph=True)
        gms, gres, *_ = timed(PLAIN)
        same = bool(np.array_equal(gres.n_final, res.n_final) and np.array_equal(gres.counters, res.counters))
        graph_line = {"segments": 4, "graph": True, "ms_per_step": gms, "same_chains": same}
        plan.set_segments(1)
    value = float(ok.sum() / (ms / 1e3))
    flops = float(res.alg_flops.astype(np.float64).sum())
    achieved = flops / (ms / 1e3) / 1e12
    traffic = read_traffic(ROOT / "profiles" / f"rjmcmc_chain_{args.workload}_pmc.txt") if psf is None else None
    plain = {"value": value, "ms_per_step": ms, "ms_per_step_mean": ms_mean, "strict_posteriors_per_s": value,
             "extrapolated_fixed_ess_rate": extrap, "roi_posteriors": int(ok.sum()),
             "ess_sum_rate": float(ess.sum() / 200.0 / (ms / 1e3)),
             "mean_ess": float(ess.mean()), "median_ess": float(np.median(ess)), "rmse_nm_final_state": rmse,
             "chains_per_s": n / (ms / 1e3), "jumps_per_s": n * (nb + nt) / (ms / 1e3)}
    line = {
        "metric": "ROI posteriors/sec at fixed ESS & RMSE: plain vs learned-proposal RJMCMC",
        "value": value, "unit": "ROI posteriors/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (accept test f64)", "data": "synthetic (seeded forward-model ROIs)",
        "config": {"workload": args.workload, "n_roi": n, "roi_sides": sorted(set(int(s) for s in rois["side"])),
                   "jumps_per_chain": nb + nt, "burnin": nb, "ess_target": 200, "proposal": "plain",
                   "parallelism": "1 GPU, one wavefront per chain, persistent launch draining one work queue (large ROIs first)",
                   "psf_model": "erf (OptModel 0)" if psf is None else
                   "sampled (OptModel 1), 72 x 72 x 4 x 4 table of the same Gaussian",
                   "launch_tuning": tuning, "graph_segments": graph_line},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": "k_chain (both size classes, one run)", "alg_flop_per_launch": flops},
        "cpu_baseline": cpu,
        "quality": plain,
    }
    if cpu is not None:
        line["gpu_over_cpu"] = value / cpu["value"] if cpu["value"] > 0 else None
    if args.workload == "batched_uneven" and psf is None and not args.no_posterior:
        line["quality"]["posterior_vs_oracle"] = posterior_agreement(PLAIN)
        print(f"[bench] plain posterior vs oracle: {line['quality']['posterior_vs_oracle']}", file=sys.stderr,
              flush=True)
    if args.proposal in ("learned", "both"):
        t0 = time.perf_counter()
        prm, hist = proposal.train(args.train_steps, batch=256, lr=3e-3, seed=SEED % 1000,
                                   chain_labels=args.chain_labels)
        train_s = time.perf_counter() - t0
        print(f"[bench] network trained: {args.train_steps} x 256 ROIs in {train_s:.1f} s", file=sys.stderr,
              flush=True)
        mm = proposal.MapMaker(prm)
        dq, dmu = _hip.DeviceBuffer(rs.total_pixels * 4), _hip.DeviceBuffer(rs.total_pixels * 4)
        plan.set_proposal(dq, dmu, alpha=0.2, halfwidth=0.35, beta=1.0)
        # intensities of births and single-emitter moves from the conditional-likelihood fit,
        # half of all moves single-emitter moves
        plan.set_intensity_proposal("fit", 1.3, 0.5)
        plan.set_split_proposal(args.split_share, 2.6)
        em0, em1 = _hip.Event(), _hip.Event()
        em0.record(stream)
        mm.run(rs, dq, dmu, stream=stream)
        em1.record(stream)
        em1.synchronize()
        map_ms = em0.elapsed_ms(em1)
        lms, lres, less, lok, lrmse, lextrap, lms_mean = timed(LEARNED,
                                                               pre=lambda: mm.run(rs, dq, dmu, stream=stream))
        lvalue = float(lok.sum() / (lms / 1e3))
        net_tflops = proposal.NET_FLOP_PER_CANVAS * n / (map_ms / 1e3) / 1e12
        line["learned"] = {
            "value": lvalue, "ms_per_step": lms, "ms_per_step_mean": lms_mean, "maps_ms": map_ms, "train_s": train_s,
            "train_rois": args.train_steps * 256, "train_chain_label_share": args.chain_labels,
            "strict_posteriors_per_s": lvalue,
            "extrapolated_fixed_ess_rate": lextrap, "roi_posteriors": int(lok.sum()),
            "ess_sum_rate": float(less.sum() / 200.0 / (lms / 1e3)),
            "mean_ess": float(less.mean()), "median_ess": float(np.median(less)), "rmse_nm_final_state": lrmse,
            "rmse_vs_plain": lrmse / rmse - 1.0 if rmse else None,
            "network_roofline": {"bound": "mfma", "achieved": net_tflops, "peak": 2500.0, "unit": "TFLOP/s",
                                 "frac": net_tflops / 2500.0, "kernel": "proposal maps (bf16 MFMA conv)"}}
        if args.workload == "batched_uneven" and not args.no_posterior:
            line["learned"]["posterior_vs_oracle"] = posterior_agreement(LEARNED, prm, split_share=args.split_share)
            print(f"[bench] learned posterior vs oracle: {line['learned']['posterior_vs_oracle']}", file=sys.stderr,
                  flush=True)
        line["learned_over_plain"] = lvalue / value if value > 0 else None
        line["learned_over_plain_extrapolated"] = lextrap / extrap if extrap > 0 else None
        mm.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
