"""Device tiling and the stitched posterior image against the oracle:
bamf_make_subregions equals makeSubRegions.m (ROI_Data, ROI_Struct, the
compact chain batch, the placement), and chains run with the placement add
into one image exactly what removeOverlapping.m + stitchROIs.m (with the
per-position sum over frames of analyzeROIs.m:153) make of the per-ROI grids
of the same chains."""
import numpy as np
import pytest

from bamf_amd import _hip, synth
from bamf_amd.rjmcmc import Plan, run_batch
from bamf_amd.tiles import Tiling, make_subregions
from oracle import tiles as ot
from rj_common import BACK_PDF, OFF_PDF, SIG_PDF, params

pytestmark = pytest.mark.gpu

GEOS = [(64, 64, 2, 16, 3), (53, 37, 2, 16, 3), (16, 40, 1, 16, 3), (40, 16, 2, 16, 0), (35, 70, 3, 8, 2),
        (20, 20, 1, 16, 3)]


def matlab_flat(data):
    return np.asfortranarray(data).ravel(order="F")


@pytest.mark.parametrize("geo", GEOS)
def test_make_subregions_matches_oracle(geo):
    ypix, xpix, nf, box, ov = geo
    data = np.random.default_rng(sum(geo)).normal(size=(ypix, xpix, nf)).astype(np.float32)
    print(f"tiling {geo}", flush=True)
    roi_o, st_o = ot.make_subregions(data, box, ov)
    roi, st = make_subregions(data, box, ov)
    assert np.array_equal(roi, roi_o)
    for k in ("Xsize", "Ysize", "FrameNum", "ROI_Ind"):
        assert np.array_equal(st[k], st_o[k]), k
    zoom = 2
    t = Tiling(ypix, xpix, nf, box, ov).run(_hip.DeviceBuffer.from_array(matlab_flat(data)), zoom=zoom)
    _hip.synchronize()
    n = t.n
    off = t.d_off.download(np.int64, n)
    pix = t.d_pixels.download(np.float32, t.npix)
    w, h = st_o["Xsize"], st_o["Ysize"]
    assert off[0] == 0 and np.array_equal(off[1:], np.cumsum(w * h)[:-1]) and (w * h).sum() == t.npix
    for b in range(n):
        tile = pix[off[b]:off[b] + w[b] * h[b]].reshape(w[b], h[b]).T
        assert np.array_equal(tile, roi_o[:h[b], :w[b], b]), b
    # placement: the zoomed tile lands where stitchROIs.m puts it, core = what removeOverlapping.m keeps
    goff = t.d_goff.download(np.int64, n)
    place = t.d_place.download(np.int32, 5 * n).reshape(n, 5)
    ld = zoom * ypix
    assert np.all(place[:, 0] == ld)
    ones = np.zeros((zoom * (box + 2 * ov), zoom * (box + 2 * ov), n), np.float32)
    for b in range(n):
        ones[:zoom * h[b], :zoom * w[b], b] = 1
    kept = ot.remove_overlapping(st_o, ov, zoom, ones)
    for b in range(n):
        g = kept[:zoom * h[b], :zoom * w[b], b]
        ys, xs = np.nonzero(g)
        assert (xs.min(), xs.max() + 1, ys.min(), ys.max() + 1) == tuple(place[b, 1:]), b
        assert g.sum() == (place[b, 2] - place[b, 1]) * (place[b, 4] - place[b, 3])
    # one marked pixel per tile, placed by goff, lands where the oracle stitch puts it
    for b in range(n):
        one = np.zeros_like(ones[:, :, :1])
        y0, x0 = place[b, 3], place[b, 1]
        one[y0, x0, 0] = 1
        st1 = {k: (v[b:b + 1] if isinstance(v, np.ndarray) else v) for k, v in st_o.items()}
        st1["FrameNum"] = np.array([1])
        im = ot.stitch_rois(one, st1, ov, (ypix, xpix, 1), box, zoom)[:, :, 0]
        r, c = np.argwhere(im)[0]
        assert goff[b] + x0 * ld + y0 == c * ld + r, b


def test_stitched_posterior_image_equals_remove_and_stitch():
    ypix = xpix = 48
    nf, box, ov, zoom = 2, 16, 3, 4
    frames = synth.make_frames(nf, 48, seed=71)          # [f, y, x]
    data = np.ascontiguousarray(frames.transpose(1, 2, 0))  # (y, x, f)
    t = Tiling(ypix, xpix, nf, box, ov).run(_hip.DeviceBuffer.from_array(matlab_flat(data)), zoom=zoom)
    _hip.synchronize()
    rs = t.roi_set(bg0=np.full(t.n, 10.0, np.float32))
    p = params(nb=400, nt=400, rho=0.02, grid_zoom=zoom)
    plan = Plan(p, OFF_PDF, SIG_PDF, BACK_PDF)
    print(f"tiles: {t.n}, sides {sorted(set(rs.xsize.tolist()))}", flush=True)
    per_roi = run_batch(plan, rs, seed=5, grids=True)
    stitched = run_batch(plan, rs, seed=5, grids=t.placement())
    assert np.all(per_roi.status == 0)
    assert per_roi.final.tobytes() == stitched.final.tobytes()
    st = t.struct()
    side = zoom * (box + 2 * ov)
    for name in ("pgrid", "pback"):
        flat = getattr(per_roi, name)
        stack = np.zeros((side, side, t.n), np.float32)
        o = 0
        for b in range(t.n):
            hz, wz = zoom * st["Ysize"][b], zoom * st["Xsize"][b]
            stack[:hz, :wz, b] = flat[o:o + hz * wz].reshape(wz, hz).T
            o += hz * wz
        want = ot.stitch_rois(ot.remove_overlapping(st, ov, zoom, stack), st, ov, (ypix, xpix, nf), box, zoom)
        want = want.sum(axis=2, dtype=np.float32)
        got = t.image(getattr(stitched, name))
        assert np.array_equal(got, want), name
    assert t.image(stitched.pgrid).sum() > 0
