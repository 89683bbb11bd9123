"""bamf_crjmcmc (the host-array gateway under the mex wrapper, with
cRJMCMC.cpp's arguments and outputs) against the batched entry point on the
same ROI and seed, and the standalone C++ bench harness."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from bamf_amd import synth
from bamf_amd.rjmcmc import PLAIN, Plan, RoiSet, run_batch
from bamf_amd._lib import lib
from bamf_amd.rjmcmc_abi import CHAIN_REC_DTYPE
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def crjmcmc(p, data, xs, ys, bg0, seed, n0=0, i0=None, x0=None, y0=None, rec_cap=0, em_cap=0, psf=None):
    z = p.grid_zoom
    pg = np.zeros(xs * ys * z * z, np.float32)
    pb = np.zeros_like(pg)
    acc = np.zeros(rec_cap, CHAIN_REC_DTYPE)
    prop = np.zeros(rec_cap, CHAIN_REC_DTYPE)
    ae = np.zeros((em_cap, 4), np.float32)
    pe = np.zeros((em_cap, 4), np.float32)
    counts = np.zeros(3, np.int64)
    stat = np.zeros(16, np.float64)
    rec = rec_cap > 0
    psf_flat = psf_size = None
    if psf is not None:
        psf_flat = np.ascontiguousarray(np.asarray(psf, np.float32).ravel(order="F"))
        psf_size = np.array(psf.shape, np.int32)
    rc = lib().bamf_crjmcmc(ctypes.byref(p), ptr(data), xs, ys, n0, ptr(i0), ptr(x0), ptr(y0), ctypes.c_float(bg0),
                            ptr(OFF_PDF), len(OFF_PDF), ptr(SIG_PDF), len(SIG_PDF), ptr(BACK_PDF), len(BACK_PDF), None,
                            ptr(psf_flat), ptr(psf_size), 0 if psf is None else 1, seed, ptr(pg), ptr(pb), ptr(acc) if rec else None, ptr(prop) if rec else None,
                            ptr(ae) if rec else None, ptr(pe) if rec else None, rec_cap, em_cap, ptr(counts),
                            ptr(stat))
    return rc, pg, pb, acc, prop, ae, pe, counts, stat


@pytest.mark.parametrize("side", [8, 17, 32])
def test_gateway_matches_batched_entry(side):
    rois = synth.make_rois(1, seed=90 + side, sides=(side,), mean_n=4.0)
    p = params(nb=700, nt=700)
    data = rois["data"].astype(np.float32)
    rc, pg, pb, acc, prop, ae, pe, counts, stat = crjmcmc(p, data, side, side, 10.0, 1234, rec_cap=1500,
                                                          em_cap=30000)
    assert rc == 0, lib().bamf_last_error()
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    rs = RoiSet.from_synth(rois, bg0=np.full(1, 10.0, np.float32))
    g = run_batch(plan, rs, seed=1234, mode=PLAIN, grids=True, chain_caps=(1500, 30000))
    assert g.status[0] == 0
    assert np.array_equal(pg.view(np.uint32), g.pgrid.view(np.uint32))
    assert np.array_equal(pb.view(np.uint32), g.pback.view(np.uint32))
    ch = g.chain
    assert np.array_equal(counts, ch["count"][0])
    n = int(counts[0])
    assert n > 0
    assert acc[:n].tobytes() == ch["acc_rec"][0, :n].tobytes()
    assert prop[:n].tobytes() == ch["prop_rec"][0, :n].tobytes()
    assert np.array_equal(ae[:counts[1]].view(np.uint32), ch["acc_em"][0, :counts[1]].view(np.uint32))
    assert np.array_equal(pe[:counts[2]].view(np.uint32), ch["prop_em"][0, :counts[2]].view(np.uint32))
    c = g.counters[0].astype(np.float64)
    assert stat[0] == p.n_trials + p.n_burnin
    with np.errstate(invalid="ignore", divide="ignore"):
        want = np.array([c[10 * gi + 5 + k] / c[10 * gi + k] for gi in range(3) for k in range(5)])
    assert np.array_equal(np.isnan(stat[1:]), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(stat[1:][ok], want[ok])


def test_gateway_initial_state_and_errors():
    rois = synth.make_rois(1, seed=5, sides=(12,), mean_n=2.0)
    p = params(nb=300, nt=300)
    data = rois["data"].astype(np.float32)
    i0 = np.array([800.0, 500.0], np.float32)
    x0 = np.array([4.2, 7.5], np.float32)
    y0 = np.array([3.1, 8.8], np.float32)
    rc, pg, *_ = crjmcmc(p, data, 12, 12, 12.0, 7, n0=2, i0=i0, x0=x0, y0=y0)
    assert rc == 0, lib().bamf_last_error()
    assert np.isfinite(pg).all() and pg.sum() > 0
    # same initial state through the batched entry point
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    rs = RoiSet(data, np.array([0, data.size]), [12], [12], bg0=[12.0], n0=[2], em0=(i0, x0, y0, [0]))
    g = run_batch(plan, rs, seed=7, grids=True)
    assert np.array_equal(pg.view(np.uint32), g.pgrid.view(np.uint32))
    # argument errors are reported, not run
    rc, *_ = crjmcmc(p, np.zeros(33 * 33, np.float32), 33, 33, 1.0, 1)
    assert rc != 0 and b"1..32" in lib().bamf_last_error()


def test_bench_harness_runs():
    exe = os.path.join(os.path.dirname(__file__), "..", "bamf_amd", "bamf_bench")
    assert os.path.exists(exe), "build the package first (bamf_amd.build)"
    r = subprocess.run([exe, "--rois", "300", "--jumps", "2000", "--steps", "1"], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    line = json.loads(r.stdout.decode().strip().splitlines()[-1])
    assert line["status_nonzero"] == 0 and line["n_roi"] == 300 and line["ms_per_step"] > 0


def test_gauss_mle_gateway_matches_device_entry():
    from bamf_amd.gauss_mle import fit_boxes

    boxes = np.ascontiguousarray(synth.make_mle_boxes(3000, 7, seed=17)[0], np.float32)
    n = boxes.shape[0]
    p, c, ll = fit_boxes(boxes, 1.3, 20)
    outs = [np.zeros(n, np.float32) for _ in range(9)]
    secs = np.zeros(1, np.float64)
    # (n, jj, ii) C order is the (sz, sz, N) column-major MATLAB stack
    rc = lib().bamf_gauss_mle_host(ptr(boxes), 7, n, ctypes.c_float(1.3), 20, *[ptr(o) for o in outs], ptr(secs))
    assert rc == 0, lib().bamf_last_error()
    for k in range(4):
        assert np.array_equal(outs[k].view(np.uint32), p[k].view(np.uint32))
        assert np.array_equal(outs[4 + k], np.sqrt(c[k]))
    assert np.array_equal(outs[8].view(np.uint32), ll.view(np.uint32))
    assert secs[0] > 0


def test_gauss_mle_fit_gateway_sigma_matches_device_entry():
    from bamf_amd.gauss_mle import fit_boxes

    boxes = np.ascontiguousarray(synth.make_mle_boxes(2000, 9, seed=18, sigma=1.5)[0], np.float32)
    n = boxes.shape[0]
    p, c, ll = fit_boxes(boxes, 1.3, 20, "Sigma")
    hp, hc, hl = np.zeros(5 * n, np.float32), np.zeros(5 * n, np.float32), np.zeros(n, np.float32)
    secs = np.zeros(1, np.float64)
    rc = lib().bamf_gauss_mle_fit_host(ptr(boxes), 9, n, ctypes.c_float(1.3), 20, 1, None, ptr(hp), ptr(hc),
                                       ptr(hl), ptr(secs))
    assert rc == 0, lib().bamf_last_error()
    assert np.array_equal(hp.reshape(5, n).view(np.uint32), p.view(np.uint32))
    assert np.array_equal(hc.reshape(5, n).view(np.uint32), c.view(np.uint32))
    assert np.array_equal(hl.view(np.uint32), ll.view(np.uint32))
    assert lib().bamf_gauss_mle_fit_host(ptr(boxes), 9, n, ctypes.c_float(1.3), 20, 7, None, ptr(hp), ptr(hc),
                                         ptr(hl), None) != 0
    assert lib().bamf_gauss_mle_fit_host(ptr(boxes), 9, n, ctypes.c_float(1.3), 20, 3, None, ptr(hp), ptr(hc),
                                         ptr(hl), None) != 0  # 'Z' without its calibration


def test_gauss_mle_scmos_gateway_matches_device_entry():
    from bamf_amd.gauss_mle import fit_boxes

    rng = np.random.default_rng(19)
    n, sz, mapsz = 1000, 7, 40
    boxes = np.ascontiguousarray(synth.make_mle_boxes(n, sz, seed=19)[0], np.float32)
    gain = rng.uniform(0.5, 30, (mapsz, mapsz)).astype(np.float32)
    coords = rng.integers(0, mapsz - sz, (n, 2)).astype(np.float32)
    p, c, ll = fit_boxes(boxes, 1.3, 20, "Sigma", None, coords, gain)
    hp, hc, hl = np.zeros(5 * n, np.float32), np.zeros(5 * n, np.float32), np.zeros(n, np.float32)
    rc = lib().bamf_gauss_mle_scmos_fit_host(ptr(boxes), sz, n, ptr(coords), ptr(gain), mapsz, ctypes.c_float(1.3),
                                             20, 1, None, None, ptr(hp), ptr(hc), ptr(hl), None)
    assert rc == 0, lib().bamf_last_error()
    assert np.array_equal(hp.reshape(5, n).view(np.uint32), p.view(np.uint32))
    assert np.array_equal(hc.reshape(5, n).view(np.uint32), c.view(np.uint32))
    assert np.array_equal(hl.view(np.uint32), ll.view(np.uint32))


def test_find_roi_gateway_matches_device_entry():
    from bamf_amd.find_roi import default_minval, find_roi

    frames = synth.make_frames(6, 64, seed=23, density_um2=1.0)  # (f, y, x), x contiguous
    img = np.ascontiguousarray(np.transpose(frames, (1, 2, 0)))  # any layout: read as MATLAB (Y, X, Z)
    smd, stack, res = find_roi(img)
    flat = np.asarray(img, order="F").ravel(order="F").astype(np.float32)
    ys, xs, nf = img.shape
    for cap in (4096, 3):
        xc, yc, fr = (np.zeros(cap, np.int32) for _ in range(3))
        st = np.zeros(cap * 49, np.float32)
        rs = np.zeros(flat.size, np.uint8)
        cnt = np.zeros(1, np.int64)
        rc = lib().bamf_find_roi_host(ptr(flat), ys, xs, nf, ctypes.c_float(1.3), ctypes.c_float(2.6), 7,
                                      ctypes.c_float(default_minval()), cap, ptr(xc), ptr(yc), ptr(fr), ptr(st),
                                      ptr(rs), ptr(cnt))
        assert rc == 0, lib().bamf_last_error()
        n = int(cnt[0])
        assert n == len(smd["XBoxCorner"]) and n > 3
        m = min(n, cap)
        assert np.array_equal(xc[:m], smd["XBoxCorner"][:m].astype(np.int32))
        assert np.array_equal(yc[:m], smd["YBoxCorner"][:m].astype(np.int32))
        assert np.array_equal(fr[:m], smd["FrameNum"][:m].astype(np.int32))
        want = np.asarray(stack, order="F").ravel(order="F")[:m * 49]
        assert np.array_equal(st[:m * 49], want)
        assert np.array_equal(rs, np.asarray(res, order="F").ravel(order="F").astype(np.uint8))
