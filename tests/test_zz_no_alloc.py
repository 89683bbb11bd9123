"""bamf_find_roi_noalloc (the FindROI pipeline on the library's per-device
scratch array, allocating nothing) against bamf_find_roi with a caller
workspace: equal detections, corners, frames and Res bitmaps, bit for bit, on
consecutive calls (the second smaller than the first, then a larger one again,
so stale scratch contents from an earlier call cannot leak into a result), and
against the oracle.  Oversized stacks are refused before anything is launched."""
import numpy as np
import pytest

from bamf_amd import _hip, synth
from bamf_amd._lib import BamfError, check, lib
from bamf_amd.find_roi import FindRoiDevice, default_minval

pytestmark = pytest.mark.gpu

BAMF_E_INVALID_ARG = -1
SCRATCH_BYTES = 64 << 20


def _run_noalloc(d_stack, xsize, ysize, nframes, box, mv, capacity):
    n = xsize * ysize * nframes
    bits = _hip.DeviceBuffer(((n + 31) // 32) * 4)
    cap = max(capacity, 1)
    lin, row, col, fr = (_hip.DeviceBuffer(cap * 8), _hip.DeviceBuffer(cap * 4), _hip.DeviceBuffer(cap * 4),
                         _hip.DeviceBuffer(cap * 4))
    count = _hip.DeviceBuffer(8)
    check(lib().bamf_find_roi_noalloc(d_stack.vp, xsize, ysize, nframes, 1.3, 2.6, box, mv, bits.vp, lin.vp, row.vp,
                                      col.vp, fr.vp, capacity, count.vp, None), "bamf_find_roi_noalloc")
    _hip.synchronize()
    k = int(count.download(np.int64, 1)[0])
    m = min(k, capacity)
    res = np.unpackbits(bits.download(np.uint32, (n + 31) // 32).view(np.uint8), bitorder="little")[:n]
    return (k, lin.download(np.int64, m), row.download(np.int32, m), col.download(np.int32, m),
            fr.download(np.int32, m), res)


def _run_workspace(d_stack, xsize, ysize, nframes, box, mv, capacity):
    dev = FindRoiDevice(xsize, ysize, nframes, capacity=capacity)
    dev.run(d_stack, 1.3, 2.6, box, mv)
    _hip.synchronize()
    return (*dev.results(), dev.res_bits())


def test_workspace_limit_is_the_declared_one():
    assert lib().bamf_find_roi_workspace_bytes(256, 256, 115) <= SCRATCH_BYTES
    assert lib().bamf_find_roi_workspace_bytes(256, 256, 116) > SCRATCH_BYTES


def test_noalloc_equals_workspace_entry_on_consecutive_calls(oracle):
    mv = default_minval()
    # large (fills most of the scratch), then smaller, then a different shape
    cases = [(100, 256, 256, 7, 11), (3, 48, 200, 5, 12), (2, 70, 70, 16, 13)]
    for nframes, ysize, xsize, box, seed in cases:
        print(f"noalloc case nframes={nframes} ysize={ysize} xsize={xsize} box={box}", flush=True)
        side = max(ysize, xsize)
        frames = np.ascontiguousarray(synth.make_frames(nframes, side, seed=seed)[:, :ysize, :xsize])
        d = _hip.DeviceBuffer.from_array(frames)
        cap = xsize * ysize * nframes // 8 + 16
        got = _run_noalloc(d, xsize, ysize, nframes, box, mv, cap)
        ref = _run_workspace(d, xsize, ysize, nframes, box, mv, cap)
        assert got[0] == ref[0] and got[0] > 0
        for g, r in zip(got[1:], ref[1:]):
            assert np.array_equal(g, r)
        if nframes <= 3:
            res_o, lin_o, row_o, col_o, fr_o = oracle.find_roi(frames.ravel(), xsize, ysize, nframes, 1.3, 2.6, box,
                                                               mv)
            assert np.array_equal(got[1], lin_o) and np.array_equal(got[2], row_o)
            assert np.array_equal(got[3], col_o) and np.array_equal(got[4], fr_o)
            assert np.array_equal(got[5].astype(np.float32), res_o)


def test_noalloc_refuses_oversized_stack():
    # refused by the host-side size check before any launch: the buffers are never touched
    d = _hip.DeviceBuffer(64)
    count = _hip.DeviceBuffer(8)
    rc = lib().bamf_find_roi_noalloc(d.vp, 256, 256, 116, 1.3, 2.6, 7, default_minval(), None, None, None, None,
                                     None, 0, count.vp, None)
    assert rc == BAMF_E_INVALID_ARG
    assert b"too large" in lib().bamf_last_error()
    with pytest.raises(BamfError):
        check(rc, "bamf_find_roi_noalloc")
