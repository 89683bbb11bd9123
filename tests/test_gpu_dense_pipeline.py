"""The dense low-SNR pipeline at reduced size against the oracle.

Frames of 512 x 512 from the device forward model at the dense workload's
settings (8 emitters per um^2 at 100 nm pixels, Gamma(4, 125) photons,
background 50, sigma 1.3):
  - FindROI with Boxsize 32 (sigma 1.3 / 2.6, Minval by the findPriors.m rule)
    bit-exact against the oracle: every detection's index, corner and frame;
  - the 32 x 32 crops equal their frame windows;
  - plain chains on the first crops (30 and more emitters each) replayed
    against the oracle's recorded draws with the dense workload's jump
    probabilities and priors: every decision, state, counter and final
    emitter equal, and each step's log-likelihood ratio within twice the
    oracle's own float32 error of the float64 re-evaluation."""
import math

import numpy as np
import pytest

from bamf_amd import _hip, synth
from bamf_amd._lib import check, lib
from bamf_amd.find_roi import FindRoiDevice
from bamf_amd.rjmcmc import REPLAY, Plan, RoiSet, make_params, make_replay, run_batch
from rj_common import run_oracle_chains
from test_gpu_rjmcmc import check_llr_against_float64, check_replay

pytestmark = pytest.mark.gpu

NF, SIDE, BOX = 4, 512, 32
DENS, GK, GT, BG, SIGMA = 0.08, 4.0, 125.0, 50.0, 1.3
CCAP = 64


def test_dense_pipeline_against_oracle(oracle):
    ncell = int(lib().bamf_synth_cells(SIDE, SIDE))
    d_frames = _hip.DeviceBuffer(NF * SIDE * SIDE * 4)
    d_truth = _hip.DeviceBuffer(NF * ncell * CCAP * 12)
    d_cnt = _hip.DeviceBuffer(NF * ncell * 4)
    print(f"synth {NF} x {SIDE} x {SIDE} dense", flush=True)
    check(lib().bamf_synth_frames(NF, SIDE, SIDE, DENS, GK, GT, BG, SIGMA, 20261206, CCAP, 1, d_frames.vp,
                                  d_truth.vp, d_cnt.vp, None), "bamf_synth_frames")
    _hip.synchronize()
    frames = d_frames.download(np.float32).reshape(NF, SIDE, SIDE)
    minval = (GK * GT) / (4 * math.pi * SIGMA ** 2) / 4
    cap = NF * 1024
    fr = FindRoiDevice(SIDE, SIDE, NF, capacity=cap)
    print("findroi box 32", flush=True)
    fr.run(d_frames, SIGMA, 2 * SIGMA, BOX, minval)
    _hip.synchronize()
    n, lin, row, col, frm = fr.results()
    _, lin_o, row_o, col_o, fr_o = oracle.find_roi(frames.ravel(), SIDE, SIDE, NF, SIGMA, 2 * SIGMA, BOX, minval)
    print(f"detections gpu {n}, oracle {len(lin_o)}", flush=True)
    assert n > 100 and n == len(lin_o)
    assert np.array_equal(lin, lin_o) and np.array_equal(row, row_o)
    assert np.array_equal(col, col_o) and np.array_equal(frm, fr_o)

    nroi = 8
    d_boxes = _hip.DeviceBuffer(nroi * BOX * BOX * 4)
    check(lib().bamf_find_roi_crop(d_frames.vp, SIDE, SIDE, fr.row.vp, fr.col.vp, fr.frame.vp, nroi, BOX,
                                   d_boxes.vp, None), "bamf_find_roi_crop")
    _hip.synchronize()
    boxes = d_boxes.download(np.float32).reshape(nroi, BOX, BOX)
    for b in range(nroi):
        r0, c0, f = row[b] - 1, col[b] - 1, frm[b] - 1
        assert np.array_equal(boxes[b], frames[f, c0:c0 + BOX, r0:r0 + BOX]), b

    # emitters inside each crop (the dense ROIs this workload chains)
    tc = d_cnt.download(np.int32).reshape(NF, ncell)
    tr = d_truth.download(np.float32).reshape(NF, ncell, CCAP, 3)
    n_in = []
    for b in range(nroi):
        f = frm[b] - 1
        em = np.concatenate([tr[f, c, :tc[f, c]] for c in range(ncell)])
        r0, c0 = row[b] - 1, col[b] - 1
        n_in.append(int(((em[:, 0] >= r0) & (em[:, 0] < r0 + BOX) & (em[:, 1] >= c0) & (em[:, 1] < c0 + BOX)).sum()))
    print("emitters per crop", n_in, flush=True)
    assert min(n_in) >= 30

    offs = np.arange(nroi + 1, dtype=np.int64) * BOX * BOX
    rois = {"data": boxes.ravel().copy(), "offset": offs, "side": np.full(nroi, BOX)}
    p = make_params(n_burnin=2000, n_trials=2000, rho=DENS, p_burnin=(0.4, 0.15, 0.15, 0.15, 0.15, 0, 0, 0),
                    p_trials=(0.6, 0.1, 0.1, 0.1, 0.1, 0, 0, 0))
    sig = synth.gamma_pdf_table(GK, GT, 10.0, 1001)
    off = synth.gamma_pdf_table(25, 2, 10.0, 1001)
    back = np.zeros(1001, np.float32)
    print(f"oracle chains: {nroi} dense 32 x 32 crops", flush=True)
    ores, tapes = run_oracle_chains(oracle, rois, p, seed=31, bg0=BG, sig=sig, off=off, back=back, chain=True)
    assert max(o["n"] for o in ores) >= 20
    plan = Plan(p, off, sig, back)
    rs = RoiSet(rois["data"], offs, rois["side"], rois["side"], bg0=np.full(nroi, BG, np.float32))
    rp = make_replay(tapes, np.stack([o["steps"] for o in ores]))
    print("gpu replay", flush=True)
    g = run_batch(plan, rs, mode=REPLAY, replay=rp, record_steps=True, hist_len=129)
    data = rois["data"]
    check_replay(rois, g, ores, tapes, p, lambda r: data[offs[r]:offs[r + 1]].sum())
    check_llr_against_float64(oracle, g, ores, data, offs, rois["side"], rois["side"], SIGMA, BG)
