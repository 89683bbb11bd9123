"""The front end at the full benchmark size (2,000 frames of 256 x 256, emitters
at 0.3 /um^2 with 100 nm pixels, Gamma(5, 200) photons, background 10,
sigma 1.3; frames from the device forward model):
  - FindROI (sigma 1.3 / 2.6, box 7, the default minimum) bit-exact against the
    oracle: every detection's linear index, corner and frame, in find() order;
  - gaussMLE XYNB on the first 100,000 FindROI crops (7 x 7, sigma 1.3, 20
    iterations) against the oracle: x, y, I, bg within 1e-4 relative for
    99.9 % of fits and 1e-3 for all; CRLB and log-likelihood within 1e-3 for
    99.9 %.
The oracle runs over frame chunks / fit chunks in parallel worker processes."""
import multiprocessing as mp
import os

import numpy as np
import pytest

from bamf_amd import _hip
from bamf_amd._lib import check, lib
from bamf_amd.find_roi import FindRoiDevice, default_minval
from bamf_amd.gauss_mle import gauss_mle_device

pytestmark = pytest.mark.gpu

NF, SIDE = 2000, 256
_W = {}


def _orc_findroi(rng):
    from oracle import pyoracle as po

    f0, f1 = rng
    st = _W["frames"][f0:f1]
    res, lin, row, col, fr = po.find_roi(st.ravel(), SIDE, SIDE, f1 - f0, 1.3, 2.6, 7, default_minval())
    return f0, lin + f0 * SIDE * SIDE, row, col, fr + f0


def _orc_mle(rng):
    from oracle import pyoracle as po

    a, b = rng
    return a, po.mle_xynb(_W["boxes"][a:b], 1.3, 20)


def _pool_map(fn, chunks):
    workers = min(16, os.cpu_count() or 1)
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(fn, chunks)


def rel(a, b, floor):
    return np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b.astype(np.float64)), floor)


def test_front_end_full_size(oracle):
    ncell = lib().bamf_synth_cells(SIDE, SIDE)
    d_frames = _hip.DeviceBuffer(NF * SIDE * SIDE * 4)
    d_truth = _hip.DeviceBuffer(NF * ncell * 128 * 12)
    d_cnt = _hip.DeviceBuffer(NF * ncell * 4)
    print(f"synth {NF} x {SIDE} x {SIDE}", flush=True)
    check(lib().bamf_synth_frames(NF, SIDE, SIDE, 0.003, 5.0, 200.0, 10.0, 1.3, 20261205, 128, 1, d_frames.vp,
                                  d_truth.vp, d_cnt.vp, None), "bamf_synth_frames")
    _hip.synchronize()
    frames = d_frames.download(np.float32).reshape(NF, SIDE, SIDE)
    del d_truth, d_cnt
    # FindROI on the device, the whole stack in one call
    n_px = NF * SIDE * SIDE
    cap = n_px // 64
    dev = FindRoiDevice(SIDE, SIDE, NF, capacity=cap)
    print("findroi", flush=True)
    dev.run(d_frames, 1.3, 2.6, 7, default_minval())
    _hip.synchronize()
    n, lin, row, col, fr = dev.results()
    assert 0 < n <= cap
    # the oracle over chunks of 50 frames (the filters act frame by frame)
    _W["frames"] = frames
    parts = sorted(_pool_map(_orc_findroi, [(f, min(NF, f + 50)) for f in range(0, NF, 50)]))
    lin_o = np.concatenate([p[1] for p in parts])
    row_o = np.concatenate([p[2] for p in parts])
    col_o = np.concatenate([p[3] for p in parts])
    fr_o = np.concatenate([p[4] for p in parts])
    print(f"detections gpu {n}, oracle {len(lin_o)}", flush=True)
    assert n == len(lin_o)
    assert np.array_equal(lin, lin_o) and np.array_equal(row, row_o)
    assert np.array_equal(col, col_o) and np.array_equal(fr, fr_o)
    # gaussMLE on the first 100,000 crops
    nfit = min(100_000, n)
    assert nfit == 100_000
    d_boxes = _hip.DeviceBuffer(nfit * 49 * 4)
    check(lib().bamf_find_roi_crop(d_frames.vp, SIDE, SIDE, dev.row.vp, dev.col.vp, dev.frame.vp, nfit, 7,
                                   d_boxes.vp, None), "bamf_find_roi_crop")
    dp, dc, dl = _hip.DeviceBuffer(16 * nfit), _hip.DeviceBuffer(16 * nfit), _hip.DeviceBuffer(4 * nfit)
    gauss_mle_device(d_boxes, 1.3, 7, 20, nfit, dp, dc, dl)
    _hip.synchronize()
    boxes = d_boxes.download(np.float32).reshape(nfit, 7, 7)
    # the crop is the MATLAB window Image(r0:r0+6, c0:c0+6, f), row fastest; a MATLAB
    # row runs along the frames' contiguous axis (findROI.m passes Xsize = size(Image, 2))
    for b in (0, nfit // 2, nfit - 1):
        r0, c0, f = row[b] - 1, col[b] - 1, fr[b] - 1
        assert np.array_equal(boxes[b], frames[f, c0:c0 + 7, r0:r0 + 7])
    pg = dp.download(np.float32).reshape(4, nfit)
    cg = dc.download(np.float32).reshape(4, nfit)
    lg = dl.download(np.float32)
    _W["boxes"] = boxes
    step = 6250
    parts = sorted(_pool_map(_orc_mle, [(a, min(nfit, a + step)) for a in range(0, nfit, step)]), key=lambda t: t[0])
    po_ = np.concatenate([p[1][0] for p in parts], axis=1)
    co_ = np.concatenate([p[1][1] for p in parts], axis=1)
    lo_ = np.concatenate([p[1][2] for p in parts])
    for k in range(4):
        r = rel(pg[k], po_[k], 1.0)
        print(f"param {k}: p99.9 rel {np.quantile(r, 0.999):.2e}, max {r.max():.2e}", flush=True)
        assert np.quantile(r, 0.999) < 1e-4 and r.max() < 1e-3
    fin = np.isfinite(co_).all(0) & np.isfinite(cg).all(0)
    assert fin.mean() > 0.999
    assert np.quantile(rel(cg[:, fin], co_[:, fin], 1e-6), 0.999) < 1e-3
    assert np.quantile(rel(lg, lo_, 1.0), 0.999) < 1e-3
