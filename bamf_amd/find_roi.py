"""Host mirror of @RJ/findROI.m.

``find_roi(image, sigma1, sigma2, boxsize, minval)`` returns
``(SMD, ROIStack, Res)`` like findROI.m:1: SMD has XBoxCorner, YBoxCorner
and FrameNum (1-based), ROIStack is (Boxsize, Boxsize, N) and Res the 0/1
map of the input's shape.  Everything runs on the device through
bamf_find_roi and bamf_find_roi_crop.
"""
from __future__ import annotations

import math

import numpy as np

from . import _hip
from ._lib import check, lib


def default_minval() -> float:
    # findROI.m:74
    return (1 / 4) * 2000 * (1 / (2 * math.pi * 2 * 1.3 ** 2))


class FindRoiDevice:
    """Device-resident detection over a stack laid out as the kernels expect:
    nframes frames of ysize runs of xsize contiguous floats."""

    def __init__(self, xsize: int, ysize: int, nframes: int, capacity: int):
        self.xsize, self.ysize, self.nframes, self.capacity = xsize, ysize, nframes, capacity
        self.ws_bytes = lib().bamf_find_roi_workspace_bytes(xsize, ysize, nframes)
        self.ws = _hip.DeviceBuffer(self.ws_bytes)
        n = xsize * ysize * nframes
        self.bits = _hip.DeviceBuffer(((n + 31) // 32) * 4)
        cap = max(capacity, 1)
        self.lin = _hip.DeviceBuffer(cap * 8)
        self.row = _hip.DeviceBuffer(cap * 4)
        self.col = _hip.DeviceBuffer(cap * 4)
        self.frame = _hip.DeviceBuffer(cap * 4)
        self.count = _hip.DeviceBuffer(8)

    def run(self, d_stack, sigma1: float, sigma2: float, boxsize: int, minval: float, stream=None) -> None:
        check(lib().bamf_find_roi(d_stack.vp, self.xsize, self.ysize, self.nframes, float(sigma1), float(sigma2),
                                  int(boxsize), float(minval), self.ws.vp, self.ws_bytes, self.bits.vp, self.lin.vp,
                                  self.row.vp, self.col.vp, self.frame.vp, self.capacity, self.count.vp,
                                  stream.vp if stream else None), "bamf_find_roi")

    def results(self):
        n = int(self.count.download(np.int64, 1)[0])
        m = min(n, self.capacity)
        return (n, self.lin.download(np.int64, m), self.row.download(np.int32, m), self.col.download(np.int32, m),
                self.frame.download(np.int32, m))

    def res_bits(self) -> np.ndarray:
        n = self.xsize * self.ysize * self.nframes
        w = self.bits.download(np.uint32, (n + 31) // 32)
        return np.unpackbits(w.view(np.uint8), bitorder="little")[:n]


MAX_FRAMES_PER_CALL = 65535  # bamf_find_roi's per-call limit


def find_roi(image: np.ndarray, sigma1: float = 1.3, sigma2: float = 2 * 1.3, boxsize: int = 7,
             minval: float | None = None, capacity: int | None = None, frames_per_call: int = MAX_FRAMES_PER_CALL):
    """image: MATLAB array (Ysize, Xsize[, Z]); any numpy order (it is read as MATLAB column-major).
    Frames go to the device in consecutive runs of at most frames_per_call
    (findROI.m:98-123 chunks the stack the same way); the filters act frame by
    frame, so the chunking changes no result and detections stay in find() order."""
    if minval is None:
        minval = default_minval()
    img = np.asarray(image, dtype=np.float32)
    if img.ndim == 2:
        img = img[:, :, None]
    ysize, xsize, nframes = img.shape
    flat = np.asarray(img, order="F").ravel(order="F")  # MATLAB memory order
    fpc = max(1, min(int(frames_per_call), MAX_FRAMES_PER_CALL, max(nframes, 1)))
    npf = xsize * ysize
    smd_parts, stacks, res_parts = [], [], []
    for f0 in range(0, max(nframes, 1), fpc):
        nf = min(fpc, nframes - f0)
        if nf <= 0:
            break
        smd, stack, res = _find_roi_call(flat[f0 * npf:(f0 + nf) * npf], ysize, xsize, nf, sigma1, sigma2, boxsize,
                                         minval, capacity)
        smd["FrameNum"] += f0
        smd_parts.append(smd)
        stacks.append(stack)
        res_parts.append(res)
    if not smd_parts:
        return ({"XBoxCorner": np.zeros(0), "YBoxCorner": np.zeros(0), "FrameNum": np.zeros(0)},
                np.zeros((boxsize, boxsize, 0), np.float32), np.zeros((ysize, xsize, 0)))
    smd = {k: np.concatenate([p[k] for p in smd_parts]) for k in ("XBoxCorner", "YBoxCorner", "FrameNum")}
    return smd, np.concatenate(stacks, axis=2), np.concatenate(res_parts, axis=2)


def _find_roi_call(flat, ysize, xsize, nframes, sigma1, sigma2, boxsize, minval, capacity):
    if capacity is None:
        capacity = max(1024, flat.size // 16)
    d_stack = _hip.DeviceBuffer.from_array(flat)
    # findROI.m passes Xsize = size(Image,2) as the kernels' contiguous run length.
    dev = FindRoiDevice(xsize, ysize, nframes, capacity)
    dev.run(d_stack, sigma1, sigma2, boxsize, minval)
    _hip.synchronize()
    n, lin, row, col, frame = dev.results()
    if n > capacity:
        return _find_roi_call(flat, ysize, xsize, nframes, sigma1, sigma2, boxsize, minval, n)
    res = dev.res_bits().astype(np.float64).reshape((ysize, xsize, nframes), order="F")
    stack = np.zeros((boxsize, boxsize, n), dtype=np.float32)
    if n:
        d_out = _hip.DeviceBuffer(n * boxsize * boxsize * 4)
        check(lib().bamf_find_roi_crop(d_stack.vp, xsize, ysize, dev.row.vp, dev.col.vp, dev.frame.vp, n, boxsize,
                                       d_out.vp, None), "bamf_find_roi_crop")
        _hip.synchronize()
        stack = d_out.download(np.float32).reshape((n, boxsize, boxsize)).transpose(2, 1, 0).copy(order="F")
    smd = {"XBoxCorner": col.astype(np.float64), "YBoxCorner": row.astype(np.float64),
           "FrameNum": frame.astype(np.float64)}
    return smd, stack, res
