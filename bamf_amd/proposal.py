"""Learned birth proposal: the network's parameters, training on synthetic
ROIs from the forward model, and per-ROI proposal maps for the chain.

The network and its training run on the device through the C ABI
(bamf_net_*); this module only prepares host arrays (initial parameters,
training batches) and drives the calls.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
from scipy.special import erf

from . import _hip
from ._lib import BamfError, check, lib

C = 32


def param_layout() -> dict:
    """Offsets of each tensor in the flat fp32 parameter array."""
    shapes = [("w1", (C, 9)), ("b1", (C,))]
    for l in (2, 3, 4):
        shapes += [(f"w{l}", (C, 9 * C)), (f"b{l}", (C,))]
    shapes += [("wh", (2, C)), ("bh", (2,))]
    out, at = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = (at, shp)
        at += n
    out["total"] = at
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal convolutions, small head, head biases at a sparse occupancy
    prior (logit -4) and a typical log-intensity (log 1000)."""
    lay = param_layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(lay["total"], np.float32)

    def put(name, arr):
        o, shp = lay[name]
        a = np.asarray(arr, np.float32).reshape(-1)
        p[o:o + a.size] = a

    put("w1", rng.normal(0, math.sqrt(2 / 9), (C, 9)))
    for l in (2, 3, 4):
        put(f"w{l}", rng.normal(0, math.sqrt(2 / (9 * C)), (C, 9 * C)))
    put("wh", rng.normal(0, 0.05, (2, C)))
    put("bh", [-4.0, math.log(1000.0)])
    return p


def training_batch(rng: np.random.Generator, n: int, side: int = 16, density=(0.5, 5.0), bg=(5.0, 30.0),
                   photons=(5.0, 200.0), sigma: float = 1.3, pixel_um: float = 0.1):
    """n synthetic side x side ROIs from the forward model (the batched
    workload's distribution).  Returns images [n][x][y] (ROI layout), the
    occupancy target (1 at pixel (floor(x + 1/2), floor(y + 1/2)) of each
    emitter: the pixel a birth proposes it from) and the log of the
    intensity summed per occupied pixel."""
    lam = rng.uniform(*density, n) * (side * pixel_um) ** 2
    k = rng.poisson(lam)
    kmax = max(int(k.max()), 1)
    live = np.arange(kmax)[None, :] < k[:, None]
    X = rng.uniform(0, side, (n, kmax))
    Y = rng.uniform(0, side, (n, kmax))
    I = rng.gamma(photons[0], photons[1], (n, kmax)) * live
    b = rng.uniform(*bg, n)
    c = 1.0 / (sigma * math.sqrt(2.0))
    e = np.arange(side + 1, dtype=np.float64)
    gx = 0.5 * np.diff(erf((e[None, None, :] - X[:, :, None]) * c), axis=2)  # (n, kmax, side)
    gy = 0.5 * np.diff(erf((e[None, None, :] - Y[:, :, None]) * c), axis=2)
    mu = b[:, None, None] + np.einsum("nk,nkx,nky->nxy", I, gx, gy)
    img = rng.poisson(mu).astype(np.float32)
    occ = np.zeros((n, side, side), np.float32)
    isum = np.zeros((n, side, side), np.float64)
    px = np.minimum(np.floor(X + 0.5).astype(np.int64), side - 1)
    py = np.minimum(np.floor(Y + 0.5).astype(np.int64), side - 1)
    rows = np.repeat(np.arange(n), kmax).reshape(n, kmax)
    sel = live
    np.add.at(isum, (rows[sel], px[sel], py[sel]), I[sel])
    occ[isum > 0] = 1.0
    logi = np.where(isum > 0, np.log(np.maximum(isum, 1e-30)), 0.0).astype(np.float32)
    return img, occ, logi


class Network:
    """A device-resident network (bamf_net) of fixed canvas and batch size."""

    def __init__(self, params: np.ndarray, max_batch: int, h: int, w: int):
        L = lib()
        params = np.ascontiguousarray(params, np.float32)
        if params.size != L.bamf_net_param_count():
            raise BamfError("parameter array of the wrong size")
        self.h, self.w, self.max_batch = h, w, max_batch
        self.ptr = ctypes.c_void_p()
        check(L.bamf_net_create(max_batch, h, w, params.ctypes.data_as(ctypes.c_void_p), ctypes.byref(self.ptr)),
              "bamf_net_create")

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib().bamf_net_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(lib().bamf_net_param_count(), np.float32)
        check(lib().bamf_net_get_params(self.ptr, out.ctypes.data_as(ctypes.c_void_p)), "bamf_net_get_params")
        return out

    def gradients(self, img, occ, logi):
        n = img.shape[0]
        P = lib().bamf_net_param_count()
        d = [_hip.DeviceBuffer.from_array(np.ascontiguousarray(a, np.float32)) for a in (img, occ, logi)]
        g, l = _hip.DeviceBuffer(P * 4), _hip.DeviceBuffer(8)
        check(lib().bamf_net_gradients(self.ptr, d[0].vp, d[1].vp, d[2].vp, n, g.vp, l.vp, None),
              "bamf_net_gradients")
        _hip.synchronize()
        return g.download(np.float32), l.download(np.float32)

    def forward(self, img):
        n = img.shape[0]
        d = _hip.DeviceBuffer.from_array(np.ascontiguousarray(img, np.float32))
        out = _hip.DeviceBuffer(n * self.h * self.w * 8)
        check(lib().bamf_net_forward(self.ptr, d.vp, n, out.vp, None), "bamf_net_forward")
        _hip.synchronize()
        return out.download(np.float32).reshape(n, self.h, self.w, 2)


def train(steps: int, batch: int = 256, lr: float = 2e-3, seed: int = 0, side: int = 16, params=None,
          log_every: int = 0, device_data: bool = True, chain_labels: float = 0.0, chain_jumps: int = 2000):
    """Adam training on fresh synthetic batches; returns (params, loss history).

    device_data: the batches come from the device forward model
    (bamf_synth_roi_batch: the batched workload's density, background and
    intensity distributions) straight into the training buffers; otherwise
    from training_batch on the host.  chain_labels: the share of steps whose
    targets are a plain chain's posterior sample on the same ROIs (the final
    state of a chain of chain_jumps jumps: occupancy and log-intensity of its
    emitters) instead of the ground truth."""
    rng = np.random.default_rng(seed)
    net = Network(init_params(seed) if params is None else params, batch, side, side)
    dimg, docc, dli = (_hip.DeviceBuffer(batch * side * side * 4) for _ in range(3))
    dloss = _hip.DeviceBuffer(8)
    stream = _hip.Stream()
    labeller = _ChainLabeller(batch, side, chain_jumps, seed) if chain_labels > 0 else None
    hist = []
    for it in range(steps):
        use_chain = labeller is not None and (it % max(1, int(round(1.0 / chain_labels)))) == 0
        if device_data:
            check(lib().bamf_synth_roi_batch(batch, side, 0.005, 0.05, 5.0, 30.0, 5.0, 200.0, 1.3,
                                             int(seed) * 1000003 + it, 1, dimg.vp, docc.vp, dli.vp, None, None,
                                             stream.vp), "bamf_synth_roi_batch")
        else:
            img, occ, logi = training_batch(rng, batch, side)
            stream.synchronize()  # the previous step has consumed the buffers
            dimg.upload(img)
            docc.upload(occ)
            dli.upload(logi)
        if use_chain:
            stream.synchronize()
            labeller.label(dimg, docc, dli, it)
        rate = lr * (0.1 ** (it / max(steps, 1)))  # exponential decay to lr/10
        check(lib().bamf_net_train_step(net.ptr, dimg.vp, docc.vp, dli.vp, batch, ctypes.c_float(rate), dloss.vp,
                                        stream.vp), "bamf_net_train_step")
        if log_every and (it % log_every == 0 or it == steps - 1):
            stream.synchronize()
            hist.append((it, *dloss.download(np.float32)))
    stream.synchronize()
    p = net.params()
    net.close()
    return p, hist


class _ChainLabeller:
    """Posterior-sample targets: plain chains (the batched workload's parameters)
    on the device images; each chain's final state gives the occupancy and
    log-intensity targets by the rule of training_batch."""

    def __init__(self, batch: int, side: int, jumps: int, seed: int):
        from . import synth
        from .rjmcmc import Plan, make_params

        self.batch, self.side, self.seed = batch, side, seed
        p = make_params(n_burnin=jumps // 2, n_trials=jumps - jumps // 2, rho=2.75 * 0.01)
        self.plan = Plan(p, synth.gamma_pdf_table(4, 5, 10.0, 1001), synth.gamma_pdf_table(5, 200, 10.0, 1001),
                         np.zeros(1001, np.float32))
        self.cap = 96

    def label(self, dimg, docc, dli, it: int) -> None:
        from .rjmcmc import PLAIN, RoiSet, run_batch

        n, s = self.batch, self.side
        offs = np.arange(n + 1, dtype=np.int64) * s * s
        xs = np.full(n, s, np.int32)
        rois = RoiSet(None, offs, xs, xs, bg0=np.full(n, 20.0, np.float32),
                      device=(dimg, _hip.DeviceBuffer.from_array(offs), _hip.DeviceBuffer.from_array(xs),
                              _hip.DeviceBuffer.from_array(xs)))
        res = run_batch(self.plan, rois, seed=self.seed * 7919 + it, mode=PLAIN, final_cap=self.cap)
        occ, logi = chain_targets(res.final, s, self.cap)
        docc.upload(occ)
        dli.upload(logi)


def chain_targets(final: np.ndarray, side: int, cap: int):
    """Occupancy and log-intensity targets of chains' final states (rows of
    [N, bg, abg, bbg, I[cap], X[cap], Y[cap]]) by the rule of training_batch:
    pixel (floor(X + 1/2), floor(Y + 1/2)) clipped to the ROI holds the summed
    intensity of its emitters; occupied where that sum is positive."""
    n = final.shape[0]
    isum = np.zeros((n, side, side), np.float64)
    for r in range(n):
        k = int(final[r, 0])
        I = final[r, 4:4 + k].astype(np.float64)
        X = final[r, 4 + cap:4 + cap + k]
        Y = final[r, 4 + 2 * cap:4 + 2 * cap + k]
        px = np.clip(np.floor(X + 0.5).astype(np.int64), 0, side - 1)
        py = np.clip(np.floor(Y + 0.5).astype(np.int64), 0, side - 1)
        np.add.at(isum[r], (px, py), np.maximum(I, 0.0))
    occ = (isum > 0).astype(np.float32)
    logi = np.where(isum > 0, np.log(np.maximum(isum, 1e-30)), 0.0).astype(np.float32)
    return occ, logi


# Multiply-adds of one 32 x 32 canvas through the network (x 2 = FLOP):
# 1024 pixels x (9 x 32 + 3 x 9 x 32 x 32 + 2 x 32).
NET_FLOP_PER_CANVAS = 2 * 1024 * (9 * C + 3 * 9 * C * C + 2 * C)


class MapMaker:
    """Proposal maps for ROI batches with a network built once (32 x 32 canvas)."""

    def __init__(self, params: np.ndarray, max_batch: int = 1024):
        self.net = Network(params, max_batch, 32, 32)

    def run(self, rois, dq, dmu, eps: float = 0.05, stream=None) -> None:
        check(lib().bamf_net_proposal_maps(self.net.ptr, ctypes.byref(rois.batch), ctypes.c_float(eps), dq.vp,
                                           dmu.vp, stream.vp if stream else None), "bamf_net_proposal_maps")

    def close(self):
        self.net.close()


def proposal_maps(params: np.ndarray, rois, eps: float = 0.05, max_batch: int = 1024):
    """Birth-proposal maps (q, mu) for a device ROI batch (rjmcmc.RoiSet),
    as device buffers in the batch's pixel layout."""
    mm = MapMaker(params, max_batch)
    total = int(rois.total_pixels)
    dq, dmu = _hip.DeviceBuffer(total * 4), _hip.DeviceBuffer(total * 4)
    mm.run(rois, dq, dmu, eps)
    _hip.synchronize()
    mm.close()
    return dq, dmu
