"""ctypes mirror of the chain structs in include/bamf_gpu.h."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, Structure, c_double, c_float, c_int, c_int16, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np


class Tuning(Structure):
    """bamf_rjmcmc_tuning."""
    _fields_ = [("chains_per_simd", c_int32 * 2), ("concurrent", c_int32), ("ms", c_float),
                ("small_max_slots", c_int32)]


class McmcParams(Structure):
    """bamf_mcmc_params: the MCMC struct fields of cRJMCMC.cpp:14-50."""
    _fields_ = [("psf_sigma", c_float), ("n_trials", c_int32), ("n_burnin", c_int32),
                ("i_std_fg", c_float), ("i_std_bg", c_float), ("x_std_fg", c_float), ("x_std_bg", c_float),
                ("p_burnin", c_float * 8), ("p_trials", c_float * 8),
                ("rho", c_float), ("bnd_out", c_float), ("grid_zoom", c_int32), ("split_std", c_float),
                ("icutoff", c_float), ("bg_std", c_float), ("abbg_std", c_float), ("dx", c_float),
                ("drift_x", c_float), ("drift_y", c_float), ("is_backg", c_int32)]


STEP_DTYPE = np.dtype([("kind", "<i4"), ("jump_type", "<i4"), ("flags", "<i4"), ("n", "<i4"),
                       ("llr", "<f4"), ("pr", "<f4"), ("bg", "<f4"), ("photons", "<f4"),
                       ("log_pr", "<f8"), ("draws", "<i4"), ("n_prop", "<i4"), ("pix", "<i4"), ("aux", "<i4")])
assert STEP_DTYPE.itemsize == 56


class RoiBatch(Structure):
    _fields_ = [("n_roi", c_int32), ("d_pix_offset", c_void_p), ("d_xsize", c_void_p), ("d_ysize", c_void_p),
                ("d_data", c_void_p), ("d_scmos", c_void_p), ("d_rho", c_void_p), ("d_bg0", c_void_p),
                ("d_n0", c_void_p), ("d_em0_offset", c_void_p), ("d_i0", c_void_p), ("d_x0", c_void_p),
                ("d_y0", c_void_p), ("d_key", c_void_p), ("d_order", c_void_p)]


class Replay(Structure):
    _fields_ = [("d_tape", c_void_p), ("d_tape_offset", c_void_p), ("d_tape_len", c_void_p),
                ("d_ref_steps", c_void_p)]


class RjOut(Structure):
    _fields_ = [("d_status", c_void_p), ("d_counters", c_void_p), ("d_final", c_void_p), ("final_cap", c_int32),
                ("d_n_hist", c_void_p), ("hist_len", c_int32), ("d_logpost", c_void_p),
                ("d_trace_n", c_void_p), ("d_trace_photons", c_void_p), ("trace_len", c_int32),
                ("trace_thin", c_int32), ("d_steps", c_void_p), ("d_pgrid", c_void_p), ("d_pback", c_void_p),
                ("d_grid_offset", c_void_p), ("d_grid_place", c_void_p),
                ("d_acc_rec", c_void_p), ("d_prop_rec", c_void_p), ("d_acc_em", c_void_p), ("d_prop_em", c_void_p),
                ("d_rec_count", c_void_p), ("rec_cap", c_int32), ("em_cap", c_int32),
                ("d_alg_flops", c_void_p)]


# bamf_chain_rec
CHAIN_REC_DTYPE = np.dtype([("n", "<i4"), ("jump_type", "<i4"), ("accepted", "<i4"), ("em_index", "<i4"),
                            ("llr", "<f4"), ("pr", "<f4"), ("bg", "<f4"), ("abg", "<f4"), ("bbg", "<f4")])
assert CHAIN_REC_DTYPE.itemsize == 36

PLAIN, REPLAY, LEARNED = 0, 1, 2


def declare(L: ctypes.CDLL) -> None:
    L.bamf_rjmcmc_plan_create.argtypes = [POINTER(McmcParams), c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                                          POINTER(c_void_p)]
    L.bamf_rjmcmc_plan_destroy.argtypes = [c_void_p]
    L.bamf_rjmcmc_run.argtypes = [c_void_p, POINTER(RoiBatch), c_uint64, c_int, POINTER(Replay), POINTER(RjOut),
                                  c_void_p]
    L.bamf_rjmcmc_last_claimed.argtypes = [c_void_p, POINTER(c_int64)]
    L.bamf_rjmcmc_set_sampled_psf.argtypes = [c_void_p, c_void_p, c_void_p]
    L.bamf_math_erf.argtypes = [c_void_p, c_void_p, c_int64, c_void_p]
    L.bamf_crjmcmc.argtypes = [POINTER(McmcParams), c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                               c_float, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                               c_void_p, c_int, c_uint64,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32,
                               c_void_p, c_void_p]
    L.bamf_gauss_mle_host.argtypes = [c_void_p, c_int, c_int64, c_float, c_int] + [c_void_p] * 10
    L.bamf_find_roi_host.argtypes = [c_void_p, c_int, c_int, c_int, c_float, c_float, c_int, c_float, c_int64,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.bamf_rjmcmc_set_segments.argtypes = [c_void_p, c_int, c_int]
    L.bamf_find_map.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int, c_int, c_int, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p]
    L.bamf_chain_summary.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p]
    L.bamf_rjmcmc_set_intensity_proposal.argtypes = [c_void_p, c_int, c_float, c_float]
    L.bamf_rjmcmc_set_split_proposal.argtypes = [c_void_p, c_float, c_float]
    L.bamf_rjmcmc_autotune.argtypes = [c_void_p, c_void_p, c_int, POINTER(Tuning), c_void_p]
    L.bamf_rjmcmc_set_tuning.argtypes = [c_void_p, POINTER(Tuning)]
    L.bamf_rjmcmc_get_tuning.argtypes = [c_void_p, POINTER(Tuning)]
    L.bamf_rjmcmc_set_proposal.argtypes = [c_void_p, c_void_p, c_void_p, c_float, c_float, c_float]
