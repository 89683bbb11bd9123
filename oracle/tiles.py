"""Oracle (test infrastructure only): numpy restatement of the reference's
frame tiling and stitching, written as the MATLAB loops are, with MATLAB's
1-based ranges converted at each indexing site.

  make_subregions     @RJ/makeSubRegions.m:42-102
  remove_overlapping  @RJ/removeOverlapping.m:29-54
  stitch_rois         @RJ/stitchROIs.m:28-65

Arrays use MATLAB's axis order (rows, cols, slices).  Pinned by
tests/test_oracle_tiles.py against what the geometry fixes: every tile equals
its frame crop and is zero padded; tile sizes and counts; the tiles' cores
partition the frame, so stitching the overlap-removed, zoomed crops of a frame
rebuilds the zoomed frame exactly; and a planted off-by-one in a band or a
placement breaks that rebuild.
"""
from __future__ import annotations

import numpy as np


def _ranges(i, nsub, pix, box, ov):
    """1-based (start, end, size) of tile i along one axis (makeSubRegions.m:62-87)."""
    s = box * (i - 1) - (ov - 1)
    e = box * i + ov
    size = box + 2 * ov
    if i == 1:
        size = box + ov
        s = 1
        e = box + ov
    if i == nsub:
        e = pix
        size = pix - box * (nsub - 1) + ov
    if nsub == 1:
        size = pix
    return s, e, size


def make_subregions(data: np.ndarray, box: int, overlap: int):
    """makeSubRegions.m:42-102.  data (ypix, xpix, nframes) -> (ROI_Data, ROI_Struct)."""
    ypix, xpix, zpix = data.shape
    xsub = -(-xpix // box)
    ysub = -(-ypix // box)
    n = xsub * ysub * zpix
    side = box + 2 * overlap
    roi = np.zeros((side, side, n), np.float32)
    st = {k: np.zeros(n, np.int64) for k in ("Xsize", "Ysize", "FrameNum", "ROI_Ind")}
    for frame in range(1, zpix + 1):
        for ix in range(1, xsub + 1):
            for iy in range(1, ysub + 1):
                sx, ex, roiex = _ranges(ix, xsub, xpix, box, overlap)
                sy, ey, roiey = _ranges(iy, ysub, ypix, box, overlap)
                ind = xsub * ysub * (frame - 1) + ysub * (ix - 1) + iy
                st["Xsize"][ind - 1] = roiex
                st["Ysize"][ind - 1] = roiey
                st["FrameNum"][ind - 1] = frame
                st["ROI_Ind"][ind - 1] = ysub * (ix - 1) + iy
                crop = data[sy - 1:ey, sx - 1:ex, frame - 1]
                if crop.shape != (roiey, roiex):
                    raise ValueError("tile reaches outside the frame")  # MATLAB raises an index error here
                roi[:roiey, :roiex, ind - 1] = crop
    st.update(SizeData=[ypix, xpix, zpix], Xsub=xsub, Ysub=ysub, Overlap=overlap)
    return roi, st


def remove_overlapping(st: dict, overlap: int, zoom: int, stack: np.ndarray) -> np.ndarray:
    """removeOverlapping.m:29-54: zero each tile's overlap bands in its zoomed grid."""
    out = stack.copy()
    xsub, ysub = st["Xsub"], st["Ysub"]
    xs = st["Xsize"] * zoom
    ys = st["Ysize"] * zoom
    zo = zoom * overlap
    for nn in range(stack.shape[2]):
        ind = int(st["ROI_Ind"][nn])
        ix = (ind - 1) // ysub + 1
        iy = ind - (ix - 1) * ysub
        if ix != xsub:
            out[:, xs[nn] - (zo - 1) - 1:xs[nn], nn] = 0
        if iy != ysub:
            out[ys[nn] - (zo - 1) - 1:ys[nn], :, nn] = 0
        if ix != 1:
            out[:, :zo, nn] = 0
        if iy != 1:
            out[:zo, :, nn] = 0
    return out


def stitch_rois(rois: np.ndarray, st: dict, overlap: int, imsize, box: int, zoom: int) -> np.ndarray:
    """stitchROIs.m:28-65: add each tile's (ysize zoom) x (xsize zoom) grid at
    its place in a (zoom Y, zoom X, Z) image, slice FrameNum."""
    im = np.zeros((zoom * imsize[0], zoom * imsize[1], imsize[2]), np.float32)
    xsub, ysub = st["Xsub"], st["Ysub"]
    xs = st["Xsize"] * zoom
    ys = st["Ysize"] * zoom
    for nn in range(rois.shape[2]):
        frame = int(st["FrameNum"][nn])
        ind = int(st["ROI_Ind"][nn])
        ix = (ind - 1) // ysub + 1
        iy = ind - (ix - 1) * ysub
        sy = box * (iy - 1) - (overlap - 1)
        ey = box * iy + overlap
        sx = box * (ix - 1) - (overlap - 1)
        ex = box * ix + overlap
        if iy == 1:
            sy, ey = 1, box + overlap
        if ix == 1:
            sx, ex = 1, box + overlap
        if iy == ysub:
            ey = imsize[0]
        if ix == xsub:
            ex = imsize[1]
        im[zoom * sy - (zoom - 1) - 1:zoom * ey, zoom * sx - (zoom - 1) - 1:zoom * ex, frame - 1] += \
            rois[:ys[nn], :xs[nn], nn]
    return im
