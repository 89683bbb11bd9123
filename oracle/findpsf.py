"""Oracle (test infrastructure only): numpy restatement of the sampled-PSF
construction of @RJ/findPSF.m:30-119 after its fits.

  shift_average   findPSF.m:38-47: each ROI's centred spectrum
                  fftshift(fft2(fftshift(R / 10))) with the phase ramp that
                  moves its fitted emitter to a common position, summed over
                  the first min(N, 1000) ROIs
  sampled_image   findPSF.m:50-68: the spectrum zero-padded by 12 per side,
                  |fftshift(ifft2(.))|, the mean of the four edge means
                  subtracted, negatives zeroed, scaled to peak 1
  phase_images    findPSF.m:70-117: padded by 112 per side and split into
                  the 4 x 4 sub-sampled 64 x 64 images (the middle third of
                  rows / columns moved down / right by one for phases > 0),
                  each normalised to sum 1: SampledPSF(y, x, phase_y, phase_x)
  find_psf        all three from the (preprocessed) ROIs and fitted X, Y

Uses numpy's FFT (the textbook DFT definitions that MATLAB's fft2/ifft2 and
fftshift share) in double precision; the reference runs in single.  The fit
itself (SMA_Core.gaussMLE 'Sigma', findPSF.m:31) belongs to an un-vendored
package: the positions are inputs here.  Pinned by tests/test_oracle_findpsf.py
(shift theorem, permutation and normalisation invariants, a centred
symmetric emitter).
"""
from __future__ import annotations

import numpy as np

BOX = 8          # the box size for which the fixed 64 x 64 x 4 x 4 split is consistent
PAD_F = 12       # findPSF.m:55
PAD_R = 112      # findPSF.m:73
OS = 4           # findPSF.m:51
SUB = 64         # 4 x FitBoxSize, findPSF.m:52,76


def _axis(b):
    return np.arange(-int(np.ceil((b - 1) / 2)), int(np.floor((b - 1) / 2)) + 1)


def shift_average(rois, x, y, box=BOX):
    """rois: (N, box, box) as R(:, :, n) (row = y, column = x); x, y the fits."""
    xg, yg = np.meshgrid(_axis(box), _axis(box))  # XGm varies along columns
    fm = np.zeros((box, box), complex)
    for n in range(min(len(rois), 1000)):
        r = np.asarray(rois[n], np.float64) / 10.0
        fim = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(r)))
        fim = fim * np.exp(2j * np.pi * xg * (float(x[n]) - box / 2 - 0.5) / box) \
                  * np.exp(2j * np.pi * yg * (float(y[n]) - box / 2 - 0.5) / box)
        fm = fm + fim
    return fm


def sampled_image(fm):
    pad = np.pad(fm, PAD_F)
    img = np.abs(np.fft.fftshift(np.fft.ifft2(pad)))
    off = np.mean([img[0, :].mean(), img[-1, :].mean(), img[:, -1].mean(), img[:, 0].mean()])
    img = img - off
    img[img < 0] = 0
    return img / img.max()


def phase_images(img, moves=True, normalise=True):
    """moves / normalise off: the bare split (for the tests)."""
    p = np.pad(img, PAD_R)
    ys, xs = p.shape
    xg, yg = np.meshgrid(np.arange(-int(np.ceil(xs / 2)), int(np.floor((xs - 1) / 2)) + 1),
                         np.arange(-int(np.ceil(ys / 2)), int(np.floor((ys - 1) / 2)) + 1))
    xgm = OS - xg / OS
    xgm = (xgm - np.floor(xgm)) * OS
    ygm = OS - yg / OS
    ygm = (ygm - np.floor(ygm)) * OS
    out = np.zeros((SUB, SUB, OS, OS))
    for ii in range(OS):
        for jj in range(OS):
            mask = (xgm == ii) & (ygm == jj)
            vals = p.T[mask.T]  # column-major traversal: rows inner, columns outer
            t = vals.reshape(SUB, SUB, order="F")
            s1, s2 = t.shape
            if moves and jj > 0:
                a, b = int(np.round(s1 / 3)), int(np.round(2 * s1 / 3))
                t[a:b + 1, :] = t[a - 1:b, :].copy()
            if moves and ii > 0:
                a, b = int(np.round(s2 / 3)), int(np.round(2 * s2 / 3))
                t[:, a:b + 1] = t[:, a - 1:b].copy()
            out[:, :, jj, ii] = t / t.sum() if normalise else t
    return out


def find_psf(rois, x, y, box=BOX):
    return phase_images(sampled_image(shift_average(rois, x, y, box))).astype(np.float32)
