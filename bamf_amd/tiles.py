"""Frame tiling and the stitched posterior image on the device.

Mirrors @RJ/makeSubRegions.m:42-102 (the chain batch cut from the frames),
@RJ/removeOverlapping.m:29-54 and @RJ/stitchROIs.m:28-65 (each tile's
posterior grid, its overlap bands removed, placed in one frame-size image) over
bamf_make_subregions.  Frames are in MATLAB layout: Data(y, x, f) at
y + ypix * (x + xpix * f), i.e. a Fortran-ordered (ypix, xpix, nframes) array.
"""
from __future__ import annotations

import numpy as np

from . import _hip
from ._lib import BamfError, check, lib
from .rjmcmc import RoiSet


class Tiling:
    """Tiles of nframes frames of ypix x xpix pixels, box + 2 overlap per side
    (less at the frame edges), numbered as makeSubRegions.m:90 numbers them."""

    def __init__(self, ypix: int, xpix: int, nframes: int, box: int = 16, overlap: int = 3):
        L = lib()
        self.ypix, self.xpix, self.nframes, self.box, self.overlap = ypix, xpix, nframes, box, overlap
        self.n = int(L.bamf_subregion_count(ypix, xpix, nframes, box, overlap))
        if self.n < 0:
            raise BamfError(f"bad tiling: {ypix} x {xpix} frames, box {box}, overlap {overlap}")
        self.npix = int(L.bamf_subregion_pixels(ypix, xpix, nframes, box, overlap))
        self.side = box + 2 * overlap
        self.zoom = 0

    def run(self, d_data, zoom: int = 0, roi_data: bool = False, stream=None) -> "Tiling":
        """Cuts the tiles of the frames in d_data (a device buffer).  zoom > 0
        also places each tile's zoom x zoom grid in the stitched image."""
        B = _hip.DeviceBuffer
        n = max(self.n, 1)
        if getattr(self, "d_pixels", None) is None:  # outputs allocated once, reused by later runs
            self.d_pixels = B(4 * max(self.npix, 1))
            self.d_off, self.d_xs, self.d_ys = B(8 * n), B(4 * n), B(4 * n)
            self.d_frame, self.d_ind = B(4 * n), B(4 * n)
            self.d_roi_data = self.d_goff = self.d_place = None
        if roi_data and self.d_roi_data is None:
            self.d_roi_data = B(4 * self.side * self.side * n)
        self.zoom = int(zoom)
        if zoom and self.d_goff is None:
            self.d_goff, self.d_place = B(8 * n), B(4 * 5 * n)
        vp = lambda b, on=True: b.vp if (b is not None and on) else None  # noqa: E731
        check(lib().bamf_make_subregions(d_data.vp, self.ypix, self.xpix, self.nframes, self.box, self.overlap,
                                         self.d_pixels.vp, self.d_off.vp, self.d_xs.vp, self.d_ys.vp,
                                         self.d_frame.vp, self.d_ind.vp, vp(self.d_roi_data, roi_data), self.zoom,
                                         vp(self.d_goff, zoom), vp(self.d_place, zoom), stream.vp if stream else None),
              "bamf_make_subregions")
        return self

    def struct(self) -> dict:
        """ROI_Struct (makeSubRegions.m:52-101) as host arrays; sync first."""
        n = self.n
        return {"Xsize": self.d_xs.download(np.int32, n), "Ysize": self.d_ys.download(np.int32, n),
                "FrameNum": self.d_frame.download(np.int32, n), "ROI_Ind": self.d_ind.download(np.int32, n),
                "SizeData": [self.ypix, self.xpix, self.nframes], "Xsub": -(-self.xpix // self.box),
                "Ysub": -(-self.ypix // self.box), "Overlap": self.overlap}

    def roi_set(self, **kw) -> RoiSet:
        """The tiles as a device-resident chain batch (no host copy of the pixels); sync first."""
        st = self.struct()
        offs = self.d_off.download(np.int64, self.n)
        return RoiSet(None, offs, st["Xsize"], st["Ysize"], device=(self.d_pixels, self.d_off, self.d_xs, self.d_ys),
                      **kw)

    def placement(self) -> dict:
        """The grids argument of rjmcmc.run_batch for one stitched
        (zoom ypix) x (zoom xpix) column-major posterior image."""
        if not self.zoom:
            raise BamfError("run the tiling with zoom > 0 for a placement")
        return {"offset": self.d_goff, "place": self.d_place, "pixels": self.zoom * self.zoom * self.ypix * self.xpix}

    def image(self, flat: np.ndarray) -> np.ndarray:
        """A downloaded stitched image as the (zoom ypix, zoom xpix) MATLAB array."""
        z = self.zoom
        return np.asarray(flat, np.float32).reshape(z * self.xpix, z * self.ypix).T


def make_subregions(data: np.ndarray, box: int = 16, overlap: int = 3):
    """makeSubRegions(Data, BoxSize, Overlap) with MATLAB's argument meaning:
    data (ypix, xpix[, nframes]); returns ROI_Data as a (box + 2 overlap,
    box + 2 overlap, n) array and ROI_Struct."""
    a = np.asarray(data, np.float32)
    if a.ndim == 2:
        a = a[:, :, None]
    ypix, xpix, nframes = a.shape
    t = Tiling(ypix, xpix, nframes, box, overlap)
    d = _hip.DeviceBuffer.from_array(np.asfortranarray(a).ravel(order="F"))
    t.run(d, roi_data=True)
    _hip.synchronize()
    s = t.side
    roi = t.d_roi_data.download(np.float32, s * s * t.n).reshape(t.n, s, s).transpose(2, 1, 0)
    return np.asfortranarray(roi), t.struct()
