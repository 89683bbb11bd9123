"""ctypes wrapper of liboracle.so — TEST INFRASTRUCTURE ONLY (see oracle/__init__.py)."""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, Structure, c_double, c_float, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def _f32p(a):
    return a.ctypes.data_as(POINTER(c_float))


class McmcParams(Structure):
    """orc_mcmc_params / bamf_mcmc_params (cRJMCMC.cpp:14-50 fields)."""
    _fields_ = [("psf_sigma", c_float), ("n_trials", c_int32), ("n_burnin", c_int32),
                ("i_std_fg", c_float), ("i_std_bg", c_float), ("x_std_fg", c_float), ("x_std_bg", c_float),
                ("p_burnin", c_float * 8), ("p_trials", c_float * 8),
                ("rho", c_float), ("bnd_out", c_float), ("grid_zoom", c_int32), ("split_std", c_float),
                ("icutoff", c_float), ("bg_std", c_float), ("abbg_std", c_float), ("dx", c_float),
                ("drift_x", c_float), ("drift_y", c_float), ("is_backg", c_int32)]


STEP_DTYPE = np.dtype([("kind", "<i4"), ("jump_type", "<i4"), ("flags", "<i4"), ("n", "<i4"),
                       ("llr", "<f4"), ("pr", "<f4"), ("bg", "<f4"), ("photons", "<f4"),
                       ("log_pr", "<f8"), ("draws", "<i4"), ("n_prop", "<i4"), ("pix", "<i4"), ("pad", "<i4")])
assert STEP_DTYPE.itemsize == 56


class Src(Structure):
    _fields_ = [("kind", c_int32), ("roi", c_uint32), ("seed", c_uint64),
                ("tape", POINTER(c_uint32)), ("tape_len", c_int64),
                ("record", POINTER(c_uint32)), ("record_cap", c_int64),
                ("consumed", c_int64), ("step", c_uint32), ("j", c_uint32), ("exhausted", c_int32)]


class Out(Structure):
    _fields_ = [("steps", c_void_p), ("pgrid", POINTER(c_float)), ("pback", POINTER(c_float)),
                ("counters", c_int32 * 30), ("n", c_int32), ("bg", c_float), ("abg", c_float), ("bbg", c_float),
                ("fin_i", POINTER(c_float)), ("fin_x", POINTER(c_float)), ("fin_y", POINTER(c_float)),
                ("fin_sig", POINTER(c_int32)), ("fin_cap", c_int32),
                ("rec_cap", c_int64), ("em_cap", c_int64),
                ("acc_head", POINTER(c_int32)), ("acc_val", POINTER(c_float)), ("acc_em", POINTER(c_float)),
                ("prop_head", POINTER(c_int32)), ("prop_val", POINTER(c_float)), ("prop_em", POINTER(c_float)),
                ("n_rec", c_int64), ("n_acc_em", c_int64), ("n_prop_em", c_int64), ("status", c_int32),
                ("snap_thin", c_int32), ("snap_skip", c_int32), ("snap_cap", c_int64), ("snap_em_cap", c_int64),
                ("snap_n", POINTER(c_int32)), ("snap_em", POINTER(c_float)), ("n_snap", c_int64),
                ("n_snap_em", c_int64)]


SRC_PHILOX, SRC_TAPE, SRC_LIBC = 0, 1, 2


def _declare(L):
    L.orc_philox4x32_10.argtypes = [POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)]
    L.orc_chain_raw.argtypes = [c_uint64, c_uint32, c_uint32, c_uint32]
    L.orc_chain_raw.restype = c_uint32
    L.orc_mle_xynb.argtypes = [c_void_p, c_float, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int]
    L.orc_mle_xynbs.argtypes = [c_void_p, c_float, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int]
    L.orc_mle_xynbsxsy.argtypes = [c_void_p, c_float, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int]
    L.orc_mle_xynbz.argtypes = [c_void_p] + [c_float] * 8 + [c_int, c_int, c_void_p, c_void_p, c_void_p, c_int]
    L.orc_mle_scmos.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_float, c_void_p, c_int, c_int,
                                c_void_p, c_void_p, c_void_p, c_int]
    L.orc_int_gauss_1d.argtypes = [c_int, c_float, c_float]
    L.orc_int_gauss_1d.restype = c_float
    L.orc_center_of_mass_2d.argtypes = [c_int, c_void_p, POINTER(c_float), POINTER(c_float)]
    L.orc_gauss_fmaxmin_2d.argtypes = [c_int, c_float, c_void_p, POINTER(c_float), POINTER(c_float)]
    L.orc_mat_inv_n.argtypes = [c_void_p, c_void_p, c_void_p, c_int]
    L.orc_gauss_coeffs.argtypes = [c_float] + [POINTER(c_float)] * 5
    L.orc_gauss_x.argtypes = [c_void_p, c_int, c_int, c_int] + [c_float] * 5
    L.orc_gauss_y.argtypes = [c_void_p, c_int, c_int, c_int] + [c_float] * 5
    L.orc_subtract.argtypes = [c_void_p, c_void_p, ctypes.c_size_t]
    L.orc_max_x.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_float]
    L.orc_max_y.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_float]
    L.orc_find_roi.argtypes = [c_void_p, c_int, c_int, c_int, c_float, c_float, c_int, c_float, c_void_p, c_int64,
                               c_void_p, c_void_p, c_void_p, c_void_p]
    L.orc_find_roi.restype = c_int64
    L.orc_rjmcmc.argtypes = [c_void_p, c_int, c_int, c_void_p, POINTER(McmcParams),
                             c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                             c_int, c_void_p, c_void_p, c_void_p, c_float, POINTER(Src), POINTER(Out)]
    L.orc_urand_from_raw.argtypes = [c_uint32]
    L.orc_urand_from_raw.restype = c_float
    L.orc_stat_findpdf.argtypes = [c_void_p, c_int, c_float, c_float]
    L.orc_stat_findpdf.restype = c_float
    L.orc_stat_tables.argtypes = [c_void_p, c_int, c_float, c_void_p, c_void_p, c_void_p, c_void_p]
    L.orc_stat_genrand.argtypes = [c_void_p, c_void_p, c_int, c_float]
    L.orc_stat_genrand.restype = c_float
    L.orc_model.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float, c_int, c_int, c_float,
                            c_void_p, c_void_p]
    L.orc_num_model.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float, c_int, c_int,
                                c_void_p, c_void_p, c_void_p, c_void_p]
    L.orc_rjmcmc_psf.argtypes = [c_void_p, c_int, c_int, c_void_p, POINTER(McmcParams),
                                 c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                                 c_int, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_int,
                                 POINTER(Src), POINTER(Out)]
    L.orc_gsplit_map.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.orc_gmerge_map.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p]
    L.orc_split_map.argtypes = [c_float] * 6 + [c_void_p]
    L.orc_split_map.restype = None
    L.orc_merge_map.argtypes = [c_float] * 6 + [c_void_p]
    L.orc_merge_map.restype = None
    L.orc_set_slip.argtypes = [c_int]
    L.orc_set_slip.restype = None


# ---------------------------------------------------------------------------
def philox(ctr, key):
    L = lib()
    c = (c_uint32 * 4)(*ctr)
    k = (c_uint32 * 2)(*key)
    o = (c_uint32 * 4)()
    L.orc_philox4x32_10(c, k, o)
    return list(o)


def mle_xynb(boxes: np.ndarray, sigma: float, iterations: int = 20):
    """kernel_MLEFit_XYNB_ restated.  Returns params (4,N), crlb (4,N), ll (N,)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n, sz = boxes.shape[0], boxes.shape[1]
    p = np.zeros(4 * n, np.float32)
    c = np.zeros(4 * n, np.float32)
    ll = np.zeros(n, np.float32)
    lib().orc_mle_xynb(boxes.ctypes.data, float(sigma), sz, iterations, p.ctypes.data, c.ctypes.data,
                       ll.ctypes.data, n)
    return p.reshape(4, n), c.reshape(4, n), ll


def mle_xynbs(boxes: np.ndarray, sigma: float, iterations: int = 20):
    """kernel_MLEFit_XYNBS_ restated.  Returns params (5,N), crlb (5,N), ll (N,)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n, sz = boxes.shape[0], boxes.shape[1]
    p = np.zeros(5 * n, np.float32)
    c = np.zeros(5 * n, np.float32)
    ll = np.zeros(n, np.float32)
    lib().orc_mle_xynbs(boxes.ctypes.data, float(sigma), sz, iterations, p.ctypes.data, c.ctypes.data,
                        ll.ctypes.data, n)
    return p.reshape(5, n), c.reshape(5, n), ll


def mle_xynbsxsy(boxes: np.ndarray, sigma: float, iterations: int = 20):
    """kernel_MLEFit_XYNBSXSY_ restated.  Returns params (6,N), crlb (6,N), ll (N,)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n, sz = boxes.shape[0], boxes.shape[1]
    p = np.zeros(6 * n, np.float32)
    c = np.zeros(6 * n, np.float32)
    ll = np.zeros(n, np.float32)
    lib().orc_mle_xynbsxsy(boxes.ctypes.data, float(sigma), sz, iterations, p.ctypes.data, c.ctypes.data,
                           ll.ctypes.data, n)
    return p.reshape(6, n), c.reshape(6, n), ll


def mle_xynbz(boxes: np.ndarray, cal, iterations: int = 20):
    """kernel_MLEFit_XYNBZ_ restated; cal = (sx0, ax, ay, bx, by, gamma, d, sy0).
    Returns params (5,N), crlb (5,N), ll (N,)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n, sz = boxes.shape[0], boxes.shape[1]
    p = np.zeros(5 * n, np.float32)
    c = np.zeros(5 * n, np.float32)
    ll = np.zeros(n, np.float32)
    lib().orc_mle_xynbz(boxes.ctypes.data, *[float(v) for v in cal], sz, iterations, p.ctypes.data, c.ctypes.data,
                        ll.ctypes.data, n)
    return p.reshape(5, n), c.reshape(5, n), ll


def mle_scmos(fit_type: int, boxes: np.ndarray, coords: np.ndarray, gain: np.ndarray, sigma: float,
              iterations: int = 20, zcal=None, z0=None):
    """The sCMOS kernels restated (fit_type 0 XYNB, 1 XYNBS, 2 XYNBSXSY, 3 XYNBZ).
    coords (N, 2) box corners (x along ii, y along jj) in the square gain map
    gain (mapsz, mapsz) read as gain[(y + jj) * mapsz + x + ii]."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n, sz = boxes.shape[0], boxes.shape[1]
    npar = 6 if fit_type == 2 else 4 if fit_type == 0 else 5
    coords = np.ascontiguousarray(coords, dtype=np.float32)
    gain = np.ascontiguousarray(gain, dtype=np.float32)
    zc = np.ascontiguousarray(zcal if zcal is not None else np.zeros(8), np.float32)
    zz = np.ascontiguousarray(z0 if z0 is not None else np.zeros(n), np.float32)
    p = np.zeros(npar * n, np.float32)
    c = np.zeros(npar * n, np.float32)
    ll = np.zeros(n, np.float32)
    lib().orc_mle_scmos(int(fit_type), boxes.ctypes.data, coords.ctypes.data, gain.ctypes.data, gain.shape[0],
                        zz.ctypes.data, float(sigma), zc.ctypes.data, sz, iterations, p.ctypes.data, c.ctypes.data,
                        ll.ctypes.data, n)
    return p.reshape(npar, n), c.reshape(npar, n), ll


def gauss_coeffs(sigma: float):
    o = [c_float() for _ in range(5)]
    lib().orc_gauss_coeffs(float(sigma), *[ctypes.byref(v) for v in o])
    return [v.value for v in o]


def find_roi(stack: np.ndarray, xsize: int, ysize: int, nframes: int, sigma1: float, sigma2: float, boxsize: int,
             minval: float, cap: int | None = None):
    """findROI.m restated on a flat stack in kernel layout.  Returns (res, lin, row, col, frame)."""
    stack = np.ascontiguousarray(stack, dtype=np.float32).ravel()
    n = xsize * ysize * nframes
    res = np.zeros(n, np.float32)
    if cap is None:
        cap = n
    lin = np.zeros(cap, np.int64)
    row = np.zeros(cap, np.int32)
    col = np.zeros(cap, np.int32)
    fr = np.zeros(cap, np.int32)
    cnt = lib().orc_find_roi(stack.ctypes.data, xsize, ysize, nframes, sigma1, sigma2, boxsize, minval,
                             res.ctypes.data, cap, lin.ctypes.data, row.ctypes.data, col.ctypes.data, fr.ctypes.data)
    m = min(cnt, cap)
    return res, lin[:m], row[:m], col[:m], fr[:m]


def model(I, X, Y, bg, abg, bbg, xsize, ysize, sigma, scmos=None):
    I = np.ascontiguousarray(I, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    if scmos is None:
        scmos = np.zeros(xsize * ysize, np.float32)
    m = np.zeros(xsize * ysize, np.float32)
    lib().orc_model(len(I), I.ctypes.data, X.ctypes.data, Y.ctypes.data, bg, abg, bbg, xsize, ysize, sigma,
                    np.ascontiguousarray(scmos, np.float32).ctypes.data, m.ctypes.data)
    return m


def split_map(I, mux, muy, u1, u2, u3):
    """split's map: (I1, X1, Y1, I2, X2, Y2) as float32."""
    out = np.zeros(6, np.float32)
    lib().orc_split_map(I, mux, muy, u1, u2, u3, out.ctypes.data)
    return out


def merge_map(I1, X1, Y1, I2, X2, Y2):
    """merge's map: (I, mux, muy, u1, u2, u3) as float32."""
    out = np.zeros(6, np.float32)
    lib().orc_merge_map(I1, X1, Y1, I2, X2, Y2, out.ctypes.data)
    return out


SLIPS = {"split_jacob_pow1": 1, "split_p_inverted": 2, "split_prm_n": 3, "split_no_area": 4,
         "merge_jacob_pow3": 5, "merge_prm_n1": 6, "gsplit_jacob_np": 7, "gsplit_us_no_np": 8,
         "gmerge_p_inverted": 9, "backfore_swapped": 10}


class planted_slip:
    """Context manager that plants one named mistake in the oracle's ratios
    (for the oracle's own pinning tests: the check must catch it)."""

    def __init__(self, name):
        self.code = SLIPS[name]

    def __enter__(self):
        lib().orc_set_slip(self.code)
        return self

    def __exit__(self, *exc):
        lib().orc_set_slip(0)
        return False


def gsplit_map(I, X, Y, idx, u1, u2, u3):
    """gSplit's proposal map: (members' new I, X, Y in idx order, new emitter (I, X, Y), group I0)."""
    I, X, Y = (np.ascontiguousarray(v, np.float32) for v in (I, X, Y))
    idx = np.ascontiguousarray(idx, np.int32)
    n = len(idx)
    o = [np.zeros(n, np.float32) for _ in range(3)] + [np.zeros(1, np.float32) for _ in range(4)]
    lib().orc_gsplit_map(n, _f32p(I), _f32p(X), _f32p(Y), idx.ctypes.data, float(u1), float(u2), float(u3),
                         *[v.ctypes.data for v in o])
    return o[0], o[1], o[2], (float(o[3][0]), float(o[4][0]), float(o[5][0])), float(o[6][0])


def gmerge_map(I, X, Y, jj, idx):
    """gMerge's proposal map: (members' new I, X, Y in idx order, (U1, U2, U3), group I0)."""
    I, X, Y = (np.ascontiguousarray(v, np.float32) for v in (I, X, Y))
    idx = np.ascontiguousarray(idx, np.int32)
    n = len(idx)
    o = [np.zeros(n, np.float32) for _ in range(3)] + [np.zeros(1, np.float32) for _ in range(4)]
    lib().orc_gmerge_map(n, _f32p(I), _f32p(X), _f32p(Y), int(jj), idx.ctypes.data, *[v.ctypes.data for v in o])
    return o[0], o[1], o[2], (float(o[3][0]), float(o[4][0]), float(o[5][0])), float(o[6][0])


def stat_tables(pdf: np.ndarray, dx: float):
    pdf = np.ascontiguousarray(pdf, np.float32)
    n = len(pdf)
    x = np.zeros(n, np.float32)
    c = np.zeros(n, np.float32)
    ux = np.zeros(n, np.float32)
    uc = np.zeros(n, np.float32)
    nu = lib().orc_stat_tables(pdf.ctypes.data, n, dx, x.ctypes.data, c.ctypes.data, ux.ctypes.data, uc.ctypes.data)
    return x, c, ux[:nu], uc[:nu]


def _psf_flat(psf):
    if psf is None:
        return None, None
    a = np.asarray(psf, np.float32)
    if a.ndim != 4:
        raise ValueError("the sampled PSF is a 4-D array")
    return np.ascontiguousarray(a.ravel(order="F")), np.array(a.shape, np.int32)


def num_model(I, X, Y, bg, abg, bbg, xsize, ysize, psf, scmos=None):
    """NumModel (LibRJMCMC.cpp:67-169) restated: the model in the chain's pixel order (x * ysize + y)."""
    I, X, Y = (np.ascontiguousarray(v, np.float32) for v in (I, X, Y))
    flat, dims = _psf_flat(psf)
    sc = np.zeros(xsize * ysize, np.float32) if scmos is None else np.ascontiguousarray(scmos, np.float32)
    m = np.zeros(xsize * ysize, np.float32)
    lib().orc_num_model(len(I), I.ctypes.data, X.ctypes.data, Y.ctypes.data, float(bg), float(abg), float(bbg),
                        xsize, ysize, flat.ctypes.data, dims.ctypes.data, sc.ctypes.data, m.ctypes.data)
    return m


def rjmcmc(data: np.ndarray, xsize: int, ysize: int, params: McmcParams, off_pdf, sig_pdf, back_pdf,
           bg0: float, *, src_kind=SRC_PHILOX, seed=0, roi=0, tape=None, record=False, scmos=None,
           n0=0, i0=None, x0=None, y0=None, grids=False, chain=False, fin_cap=1024, rec_cap=None, em_cap=None,
           psf=None, snap_thin=0, snap_skip=0, snap_em_cap=None):
    """One chain, restated.  Returns a dict with steps (STEP_DTYPE), counters, final state,
    optional replay tape, grids and Chain / Chain_Test records.  psf: a 4-D sampled PSF
    (SampledPSF with OptModel = 1), indexed [y, x, phase_y, phase_x]."""
    L = lib()
    data = np.ascontiguousarray(data, np.float32).ravel()
    if scmos is None:
        scmos = np.zeros_like(data)
    scmos = np.ascontiguousarray(scmos, np.float32).ravel()
    off_pdf = np.ascontiguousarray(off_pdf, np.float32)
    sig_pdf = np.ascontiguousarray(sig_pdf, np.float32)
    back_pdf = np.ascontiguousarray(back_pdf, np.float32)
    total = params.n_burnin + params.n_trials
    steps = np.zeros(total, STEP_DTYPE)
    src = Src()
    src.kind = src_kind
    src.roi = roi
    src.seed = seed
    keep = []
    if tape is not None:
        tape = np.ascontiguousarray(tape, np.uint32)
        keep.append(tape)
        src.tape = tape.ctypes.data_as(POINTER(c_uint32))
        src.tape_len = len(tape)
    rec = None
    if record:
        rec = np.zeros(total * 64 + 1024, np.uint32)
        src.record = rec.ctypes.data_as(POINTER(c_uint32))
        src.record_cap = len(rec)
    out = Out()
    out.steps = steps.ctypes.data
    fin_i = np.zeros(fin_cap, np.float32)
    fin_x = np.zeros(fin_cap, np.float32)
    fin_y = np.zeros(fin_cap, np.float32)
    fin_s = np.zeros(fin_cap, np.int32)
    out.fin_i, out.fin_x, out.fin_y = _f32p(fin_i), _f32p(fin_x), _f32p(fin_y)
    out.fin_sig = fin_s.ctypes.data_as(POINTER(c_int32))
    out.fin_cap = fin_cap
    z = params.grid_zoom
    pg = pb = None
    if grids:
        pg = np.zeros(xsize * ysize * z * z, np.float32)
        pb = np.zeros_like(pg)
        out.pgrid, out.pback = _f32p(pg), _f32p(pb)
    heads = None
    if chain:
        rc = rec_cap or total
        ec = em_cap or total * 64
        heads = [np.zeros(rc * 4, np.int32), np.zeros(rc * 5, np.float32), np.zeros(ec * 4, np.float32),
                 np.zeros(rc * 4, np.int32), np.zeros(rc * 5, np.float32), np.zeros(ec * 4, np.float32)]
        out.rec_cap, out.em_cap = rc, ec
        out.acc_head = heads[0].ctypes.data_as(POINTER(c_int32))
        out.acc_val, out.acc_em = _f32p(heads[1]), _f32p(heads[2])
        out.prop_head = heads[3].ctypes.data_as(POINTER(c_int32))
        out.prop_val, out.prop_em = _f32p(heads[4]), _f32p(heads[5])
    snaps = None
    if snap_thin > 0:
        ns = max((params.n_trials - snap_skip) // snap_thin, 0) + 1
        ec = snap_em_cap or ns * 128
        snaps = (np.zeros(ns, np.int32), np.zeros(ec * 3, np.float32))
        out.snap_thin, out.snap_skip, out.snap_cap, out.snap_em_cap = snap_thin, snap_skip, ns, ec
        out.snap_n = snaps[0].ctypes.data_as(POINTER(c_int32))
        out.snap_em = _f32p(snaps[1])
    if i0 is None:
        i0 = x0 = y0 = np.zeros(1, np.float32)
    i0 = np.ascontiguousarray(i0, np.float32)
    x0 = np.ascontiguousarray(x0, np.float32)
    y0 = np.ascontiguousarray(y0, np.float32)
    flat, dims = _psf_flat(psf)
    st = L.orc_rjmcmc_psf(data.ctypes.data, xsize, ysize, scmos.ctypes.data, ctypes.byref(params),
                          off_pdf.ctypes.data, len(off_pdf), sig_pdf.ctypes.data, len(sig_pdf),
                          back_pdf.ctypes.data, len(back_pdf), n0, i0.ctypes.data, x0.ctypes.data, y0.ctypes.data,
                          float(bg0), None if flat is None else flat.ctypes.data,
                          None if dims is None else dims.ctypes.data, 0 if psf is None else 1,
                          ctypes.byref(src), ctypes.byref(out))
    r = {"status": st, "steps": steps, "counters": np.array(out.counters[:], np.int32), "n": out.n,
         "bg": out.bg, "abg": out.abg, "bbg": out.bbg,
         "I": fin_i[:out.n].copy(), "X": fin_x[:out.n].copy(), "Y": fin_y[:out.n].copy(), "sig": fin_s[:out.n].copy(),
         "consumed": src.consumed}
    if record:
        r["tape"] = rec[:src.consumed].copy()
    if grids:
        r["pgrid"], r["pback"] = pg, pb
    if snaps is not None:
        k, ke = int(out.n_snap), int(out.n_snap_em)
        if ke > len(snaps[1]) // 3:
            raise RuntimeError("snapshot emitter capacity exceeded")
        r["snap_n"] = snaps[0][:k].copy()
        r["snap_em"] = snaps[1][:ke * 3].reshape(ke, 3).copy()
    if chain:
        nr = out.n_rec
        r["chain"] = {"acc_head": heads[0][:nr * 4].reshape(nr, 4), "acc_val": heads[1][:nr * 5].reshape(nr, 5),
                      "acc_em": heads[2][:out.n_acc_em * 4].reshape(-1, 4),
                      "prop_head": heads[3][:nr * 4].reshape(nr, 4), "prop_val": heads[4][:nr * 5].reshape(nr, 5),
                      "prop_em": heads[5][:out.n_prop_em * 4].reshape(-1, 4)}
    return r
