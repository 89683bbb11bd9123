"""CPU reference of the birth-proposal network.  TEST INFRASTRUCTURE ONLY.

A numpy restatement of the network that bamf_amd/csrc/net.hip runs on the
matrix cores (3x3 convolutions 1->32, 32->32 x 3 with ReLU, 1x1 head 32->2),
its loss and its gradients.  It has no counterpart in the reference (the
reference's birth draws from the data residual, LibRJMCMC.cpp:1688-1736).

Two precisions:
  bf16=True   the device's rounding points: the normalised input, the
              weights of the 3x3 layers, every activation and every
              back-propagated pre-activation gradient are rounded to
              bfloat16 (round to nearest even); sums are float32;
  bf16=False  float64 throughout (the finite-difference check of the
              gradient formulas uses this).
Parameter layout: as bamf_amd.proposal.param_layout().
"""
from __future__ import annotations

import numpy as np

C = 32


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest even) and back."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def layout():
    shapes = [("w1", (C, 9)), ("b1", (C,))]
    for l in (2, 3, 4):
        shapes += [(f"w{l}", (C, 9 * C)), (f"b{l}", (C,))]
    shapes += [("wh", (2, C)), ("bh", (2,))]
    out, at = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = (at, shp)
        at += n
    return out, at


def unpack(p: np.ndarray, dt) -> dict:
    lay, _ = layout()
    return {k: np.asarray(p[o:o + int(np.prod(s))], dt).reshape(s) for k, (o, s) in lay.items()}


def pack(d: dict) -> np.ndarray:
    lay, tot = layout()
    out = np.zeros(tot, np.float64)
    for k, (o, s) in lay.items():
        out[o:o + int(np.prod(s))] = d[k].reshape(-1)
    return out


def prep(img: np.ndarray, dt) -> np.ndarray:
    """(sqrt(max(D, 0) + 3/8) - mean over the image) / 2."""
    s = np.sqrt(np.maximum(np.asarray(img, dt), 0) + dt(0.375))
    return ((s - s.mean(axis=(1, 2), keepdims=True, dtype=dt).astype(dt)) * dt(0.5)).astype(dt)


def im2col(a: np.ndarray) -> np.ndarray:
    """(B, H, W, Cin) -> (B, H, W, 9, Cin) with zero padding; tap = 3 (dh + 1) + (dw + 1)."""
    B, H, W, Ci = a.shape
    pad = np.zeros((B, H + 2, W + 2, Ci), a.dtype)
    pad[:, 1:-1, 1:-1] = a
    cols = np.empty((B, H, W, 9, Ci), a.dtype)
    for t in range(9):
        dh, dw = t // 3 - 1, t % 3 - 1
        cols[:, :, :, t] = pad[:, 1 + dh:1 + dh + H, 1 + dw:1 + dw + W]
    return cols


def col2im(g: np.ndarray) -> np.ndarray:
    """Adjoint of im2col: (B, H, W, 9, Cin) -> (B, H, W, Cin)."""
    B, H, W, _, Ci = g.shape
    pad = np.zeros((B, H + 2, W + 2, Ci), g.dtype)
    for t in range(9):
        dh, dw = t // 3 - 1, t % 3 - 1
        pad[:, 1 + dh:1 + dh + H, 1 + dw:1 + dw + W] += g[:, :, :, t]
    return pad[:, 1:-1, 1:-1]


def loss_and_grad(params: np.ndarray, img: np.ndarray, occ: np.ndarray, logi: np.ndarray, bf16_mode: bool = True):
    """Returns (loss terms (2,), flat gradient, head outputs (B, H, W, 2))."""
    dt = np.float32 if bf16_mode else np.float64
    rnd = bf16 if bf16_mode else (lambda x: x)
    P = unpack(params, dt)
    x0 = rnd(prep(img, dt))
    B, H, W = x0.shape
    acts = [x0[..., None]]
    cols = []
    for l in range(4):
        w = rnd(P[f"w{l + 1}"]) if l > 0 else rnd(P["w1"])
        ci = acts[-1].shape[-1]
        col = im2col(acts[-1])  # (B, H, W, 9, ci)
        cols.append(col)
        z = col.reshape(B, H, W, 9 * ci) @ w.reshape(C, 9 * ci).T.astype(dt) + P[f"b{l + 1}"]
        acts.append(rnd(np.maximum(z, 0).astype(dt)))
    a4 = acts[-1]
    h = a4 @ P["wh"].T + P["bh"]
    logit, mu = h[..., 0], h[..., 1]
    n = dt(B * H * W)
    s = 1 / (1 + np.exp(-logit))
    l0 = (np.maximum(logit, 0) - logit * occ + np.log1p(np.exp(-np.abs(logit)))).sum() / n
    e = mu - logi
    l1 = (occ * e * e).sum() / n
    g0 = (s - occ) / n
    g1 = 2 * occ * e / n
    G = {k: np.zeros_like(v) for k, v in P.items()}
    G["wh"][0] = (a4 * g0[..., None]).sum(axis=(0, 1, 2))
    G["wh"][1] = (a4 * g1[..., None]).sum(axis=(0, 1, 2))
    G["bh"][0], G["bh"][1] = g0.sum(), g1.sum()
    dz = rnd(((g0[..., None] * P["wh"][0] + g1[..., None] * P["wh"][1]) * (a4 > 0)).astype(dt))
    for l in range(3, -1, -1):
        ci = acts[l].shape[-1]
        w = rnd(P[f"w{l + 1}"]).reshape(C, 9 * ci)
        G[f"w{l + 1}"] = (dz.reshape(-1, C).T @ cols[l].reshape(-1, 9 * ci)).reshape(G[f"w{l + 1}"].shape)
        G[f"b{l + 1}"] = dz.reshape(-1, C).sum(axis=0)
        if l > 0:
            da = col2im((dz.reshape(B, H, W, C) @ w).reshape(B, H, W, 9, ci))
            dz = rnd((da * (acts[l] > 0)).astype(dt))
    return np.array([l0, l1]), pack(G), h
