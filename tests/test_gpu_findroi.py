"""GPU FindROI against the oracle: bit-exact filter outputs, Res maps and
ROI indices (BASELINE: FindROI must be bit-exact in ROI indices)."""
import ctypes

import numpy as np
import pytest

from bamf_amd import _hip, synth
from bamf_amd._lib import check, lib
from bamf_amd.find_roi import FindRoiDevice, default_minval, find_roi

pytestmark = pytest.mark.gpu


def run_kernel(name, arr, *args):
    d = _hip.DeviceBuffer.from_array(arr)
    check(getattr(lib(), name)(d.vp, *args, None), name)
    _hip.synchronize()
    return d.download(np.float32).reshape(arr.shape)


@pytest.mark.parametrize("shape", [(3, 32, 32), (2, 17, 45), (1, 256, 256), (2, 5, 300)])
@pytest.mark.parametrize("sigma", [1.3, 2.6])
def test_gauss_x_y_bit_exact(oracle, shape, sigma):
    nz, ys, xs = shape
    rng = np.random.default_rng(nz * 1000 + ys)
    img = rng.poisson(30, shape).astype(np.float32)
    c = oracle.gauss_coeffs(sigma)
    print(f"shape={shape} sigma={sigma}", flush=True)
    gx = run_kernel("bamf_find_roi_gauss_x", img, xs, ys, nz, *c)
    ox = img.copy()
    oracle.lib().orc_gauss_x(ox.ctypes.data, xs, ys, nz, *c)
    assert np.array_equal(gx.view(np.uint32), ox.view(np.uint32))
    gy = run_kernel("bamf_find_roi_gauss_y", ox, xs, ys, nz, *c)
    oracle.lib().orc_gauss_y(ox.ctypes.data, xs, ys, nz, *c)
    assert np.array_equal(gy.view(np.uint32), ox.view(np.uint32))


@pytest.mark.parametrize("k", [1, 7, 20])
def test_max_kernels_bit_exact(oracle, k):
    rng = np.random.default_rng(k)
    a = rng.normal(0, 30, (2, 40, 70)).astype(np.float32)
    nz, ys, xs = a.shape
    da = _hip.DeviceBuffer.from_array(a)
    db = _hip.DeviceBuffer(a.nbytes)
    check(lib().bamf_find_roi_max_x(da.vp, db.vp, xs, ys, nz, k, 5.0, None), "max_x")
    check(lib().bamf_find_roi_max_y(da.vp, db.vp, xs, ys, nz, k, 5.0, None), "max_y")
    _hip.synchronize()
    ob = np.zeros_like(a)
    oracle.lib().orc_max_x(a.ctypes.data, ob.ctypes.data, xs, ys, nz, k, 5.0)
    assert np.array_equal(db.download(np.float32).reshape(a.shape).view(np.uint32), ob.view(np.uint32))
    oa = np.zeros_like(a)
    oracle.lib().orc_max_y(oa.ctypes.data, ob.ctypes.data, xs, ys, nz, k, 5.0)
    assert np.array_equal(da.download(np.float32).reshape(a.shape), oa)


@pytest.mark.parametrize("nframes,size,box,seed", [(6, 256, 7, 1), (3, 100, 7, 2), (2, 64, 18, 3), (1, 33, 3, 4)])
def test_find_roi_pipeline_bit_exact(oracle, nframes, size, box, seed):
    frames = synth.make_frames(nframes, size, seed=seed)
    mv = default_minval()
    print(f"nframes={nframes} size={size} box={box}", flush=True)
    res_o, lin_o, row_o, col_o, fr_o = oracle.find_roi(frames.ravel(), size, size, nframes, 1.3, 2.6, box, mv)
    d = _hip.DeviceBuffer.from_array(frames)
    dev = FindRoiDevice(size, size, nframes, capacity=size * size * nframes)
    dev.run(d, 1.3, 2.6, box, mv)
    _hip.synchronize()
    n, lin, row, col, fr = dev.results()
    assert n == len(lin_o) and n > 0
    assert np.array_equal(lin, lin_o)
    assert np.array_equal(row, row_o) and np.array_equal(col, col_o) and np.array_equal(fr, fr_o)
    assert np.array_equal(dev.res_bits().astype(np.float32), res_o)


@pytest.mark.parametrize("nframes,ysize,xsize,box", [
    (2, 48, 200, 5),    # rectangular, partial last row chunk
    (3, 37, 29, 0),     # one partial chunk per row, 1 x 1 window
    (2, 33, 65, 1),
    (2, 70, 70, 16),    # widest window of the fused local max
    (1, 64, 96, 17),    # wide window: the reference's two max passes
    (1, 400, 50, 4),    # column checkpoints in 64-column work-groups
    (1, 1100, 40, 2),   # more column chunks than the checkpointed column pass holds
    (1, 40, 96, 5),     # whole row blocks unguarded, then a guarded partial block (row_begin > 0)
    (3, 50, 64, 7),     # the same over several frames
])
def test_find_roi_pipeline_shapes(oracle, nframes, ysize, xsize, box):
    side = max(ysize, xsize)
    frames = np.ascontiguousarray(synth.make_frames(nframes, side, seed=ysize + xsize)[:, :ysize, :xsize])
    mv = default_minval()
    print(f"nframes={nframes} ysize={ysize} xsize={xsize} box={box}", flush=True)
    res_o, lin_o, row_o, col_o, fr_o = oracle.find_roi(frames.ravel(), xsize, ysize, nframes, 1.3, 2.6, box, mv)
    d = _hip.DeviceBuffer.from_array(frames)
    dev = FindRoiDevice(xsize, ysize, nframes, capacity=xsize * ysize * nframes)
    dev.run(d, 1.3, 2.6, box, mv)
    _hip.synchronize()
    n, lin, row, col, fr = dev.results()
    assert n == len(lin_o) and n > 0
    assert np.array_equal(lin, lin_o)
    assert np.array_equal(row, row_o) and np.array_equal(col, col_o) and np.array_equal(fr, fr_o)
    assert np.array_equal(dev.res_bits().astype(np.float32), res_o)


def test_find_roi_capacity_and_empty():
    frames = np.full((2, 64, 64), 10.0, np.float32)  # flat: nothing above minval
    d = _hip.DeviceBuffer.from_array(frames)
    dev = FindRoiDevice(64, 64, 2, capacity=4)
    dev.run(d, 1.3, 2.6, 7, default_minval())
    _hip.synchronize()
    assert dev.results()[0] == 0
    frames = synth.make_frames(2, 64, seed=9)
    d = _hip.DeviceBuffer.from_array(frames)
    dev.run(d, 1.3, 2.6, 7, default_minval())
    _hip.synchronize()
    n, lin, *_ = dev.results()
    assert n > 4 and len(lin) == 4  # count reports all, only capacity entries written


def test_find_roi_matlab_mirror(oracle):
    frames = synth.make_frames(2, 64, seed=10)
    img = np.transpose(frames, (2, 1, 0))  # MATLAB (Ysize, Xsize, Z) view of the same memory
    smd, stack, res = find_roi(np.asfortranarray(img), 1.3, 2.6, 7)
    _, lin_o, row_o, col_o, fr_o = oracle.find_roi(frames.ravel(), 64, 64, 2, 1.3, 2.6, 7, default_minval())
    assert np.array_equal(smd["YBoxCorner"], row_o) and np.array_equal(smd["XBoxCorner"], col_o)
    for b in range(len(lin_o)):
        r, c, f = row_o[b] - 1, col_o[b] - 1, fr_o[b] - 1
        assert np.array_equal(stack[:, :, b], img[r:r + 7, c:c + 7, f])


def test_find_roi_mirror_frame_chunks(oracle):
    """The MATLAB mirror in runs of 3 frames equals one call over all 7 frames."""
    frames = synth.make_frames(7, 48, seed=21)
    img = np.asfortranarray(np.transpose(frames, (2, 1, 0)))
    one = find_roi(img, 1.3, 2.6, 7)
    runs = find_roi(img, 1.3, 2.6, 7, frames_per_call=3)
    for k in ("XBoxCorner", "YBoxCorner", "FrameNum"):
        assert np.array_equal(one[0][k], runs[0][k])
    assert np.array_equal(one[1], runs[1]) and np.array_equal(one[2], runs[2])
    _, lin_o, row_o, col_o, fr_o = oracle.find_roi(frames.ravel(), 48, 48, 7, 1.3, 2.6, 7, default_minval())
    assert np.array_equal(runs[0]["FrameNum"], fr_o) and np.array_equal(runs[0]["YBoxCorner"], row_o)


def test_find_roi_more_frames_than_one_call(oracle):
    """65,600 frames of 16 x 16: more than one bamf_find_roi call holds, through
    the mirror and the host gateway (findROI.m handles any frame count)."""
    nf, side = 65600, 16
    rng = np.random.default_rng(31)
    frames = rng.poisson(10.0, (nf, side, side)).astype(np.float32)
    hot = rng.integers(0, nf, 4000)
    frames[hot, rng.integers(3, 13, 4000), rng.integers(3, 13, 4000)] += 400.0
    mv = default_minval()
    print(f"frames={nf} side={side}", flush=True)
    _, lin_o, row_o, col_o, fr_o = oracle.find_roi(frames.ravel(), side, side, nf, 1.3, 2.6, 7, mv)
    assert len(fr_o) > 0 and fr_o.max() > 65535
    img = np.transpose(frames, (2, 1, 0))  # MATLAB (Y, X, Z) view of the same memory
    smd, _, _ = find_roi(img, 1.3, 2.6, 7, mv)
    assert np.array_equal(smd["FrameNum"], fr_o) and np.array_equal(smd["YBoxCorner"], row_o)
    assert np.array_equal(smd["XBoxCorner"], col_o)
    cap = len(fr_o) + 10
    xc, yc, fn = (np.zeros(cap, np.int32) for _ in range(3))
    count = np.zeros(1, np.int64)
    check(lib().bamf_find_roi_host(np.ascontiguousarray(frames).ctypes.data, side, side, nf, 1.3, 2.6, 7, mv, cap,
                                   xc.ctypes.data, yc.ctypes.data, fn.ctypes.data, None, None, count.ctypes.data),
          "bamf_find_roi_host")
    n = int(count[0])
    assert n == len(fr_o)
    assert np.array_equal(fn[:n], fr_o) and np.array_equal(yc[:n], row_o) and np.array_equal(xc[:n], col_o)
