"""The tiling/stitching oracle (oracle/tiles.py) pinned by what the geometry
fixes: tiles are the frame crops [max(1, B(i-1)+1-O), min(P, B i + O)] per axis
(an independent form of makeSubRegions.m's ranges), zero padded, numbered
frame-major then column-major; the tiles' cores partition the frame, so the
overlap-removed, zoomed crops stitch back into the zoomed frames exactly; and
planted off-by-one mistakes in a band or a placement break that rebuild.
Also: the library rejects the tilings MATLAB cannot index (host-side check)."""
import numpy as np
import pytest

from oracle import tiles as ot

GEOS = [  # ypix, xpix, nframes, box, overlap
    (64, 64, 2, 16, 3), (53, 37, 2, 16, 3), (16, 40, 1, 16, 3), (40, 16, 2, 16, 0), (35, 70, 3, 8, 2),
    (20, 20, 1, 16, 3), (48, 48, 2, 16, 3),
]


def crop_ranges(i, pix, box, ov):
    return max(1, box * (i - 1) + 1 - ov), min(pix, box * i + ov)


@pytest.mark.parametrize("geo", GEOS)
def test_tiles_are_frame_crops(geo):
    ypix, xpix, nf, box, ov = geo
    rng = np.random.default_rng(sum(geo))
    data = rng.normal(size=(ypix, xpix, nf)).astype(np.float32)
    roi, st = ot.make_subregions(data, box, ov)
    xsub, ysub = -(-xpix // box), -(-ypix // box)
    assert roi.shape == (box + 2 * ov, box + 2 * ov, xsub * ysub * nf)
    b = 0
    for f in range(nf):
        for ix in range(1, xsub + 1):
            for iy in range(1, ysub + 1):
                sx, ex = crop_ranges(ix, xpix, box, ov)
                sy, ey = crop_ranges(iy, ypix, box, ov)
                w, h = ex - sx + 1, ey - sy + 1
                assert (st["Xsize"][b], st["Ysize"][b], st["FrameNum"][b], st["ROI_Ind"][b]) == \
                    (w, h, f + 1, ysub * (ix - 1) + iy)
                assert np.array_equal(roi[:h, :w, b], data[sy - 1:ey, sx - 1:ex, f])
                assert not roi[h:, :, b].any() and not roi[:, w:, b].any()
                b += 1
    # the widths across (and heights down) add up to the frame plus two overlaps per interior seam
    assert st["Xsize"][0:xsub * ysub:ysub].sum() == xpix + (2 * ov * (xsub - 1) if xsub > 1 else 0)
    assert st["Ysize"][0:ysub].sum() == ypix + (2 * ov * (ysub - 1) if ysub > 1 else 0)


def zoomed_tile_stack(roi, st, zoom):
    side = roi.shape[0]
    out = np.zeros((zoom * side, zoom * side, roi.shape[2]), np.float32)
    for b in range(roi.shape[2]):
        h, w = st["Ysize"][b], st["Xsize"][b]
        out[:zoom * h, :zoom * w, b] = np.kron(roi[:h, :w, b], np.ones((zoom, zoom), np.float32))
    return out


def rebuilds(data, box, ov, zoom, remove=ot.remove_overlapping, stitch=ot.stitch_rois):
    roi, st = ot.make_subregions(data, box, ov)
    stack = remove(st, ov, zoom, zoomed_tile_stack(roi, st, zoom))
    im = stitch(stack, st, ov, data.shape, box, zoom)
    want = np.stack([np.kron(data[:, :, f], np.ones((zoom, zoom), np.float32)) for f in range(data.shape[2])], 2)
    return np.array_equal(im, want)


@pytest.mark.parametrize("geo", GEOS)
@pytest.mark.parametrize("zoom", [1, 3])
def test_cores_partition_the_frame(geo, zoom):
    ypix, xpix, nf, box, ov = geo
    data = np.random.default_rng(7).uniform(1, 2, size=(ypix, xpix, nf)).astype(np.float32)
    assert rebuilds(data, box, ov, zoom)


def test_planted_mistakes_break_the_rebuild():
    data = np.random.default_rng(8).uniform(1, 2, size=(48, 48, 2)).astype(np.float32)

    def remove_short_band(st, ov, zoom, stack):  # zeroes one column too few on the right
        out = ot.remove_overlapping(st, ov, zoom, stack)
        for nn in range(stack.shape[2]):
            ix = (int(st["ROI_Ind"][nn]) - 1) // st["Ysub"] + 1
            if ix != st["Xsub"]:
                c = st["Xsize"][nn] * zoom - zoom * ov
                out[:, c, nn] = stack[:, c, nn]
        return out

    def stitch_shifted(rois, st, ov, imsize, box, zoom):  # placement one zoomed row too low
        im = ot.stitch_rois(rois, st, ov, imsize, box, zoom)
        return np.roll(im, 1, axis=0)

    assert rebuilds(data, 16, 3, 2)
    assert not rebuilds(data, 16, 3, 2, remove=remove_short_band)
    assert not rebuilds(data, 16, 3, 2, stitch=stitch_shifted)


def test_unindexable_tilings_are_rejected():
    from bamf_amd._lib import lib
    data = np.zeros((33, 33, 1), np.float32)  # box 16: the middle tile would reach row/column 35
    with pytest.raises(ValueError):
        ot.make_subregions(data, 16, 3)
    L = lib()
    assert L.bamf_subregion_count(33, 33, 1, 16, 3) == -1
    assert L.bamf_subregion_count(64, 64, 2, 16, 3) == 32
    assert L.bamf_subregion_pixels(64, 64, 2, 16, 3) == 2 * (64 + 2 * 3 * 3) ** 2
    rc = L.bamf_make_subregions(None, 33, 33, 1, 16, 3, None, None, None, None, None, None, None, 0, None, None,
                                None)
    assert rc != 0 and b"bad tiling" in L.bamf_last_error()
