"""ctypes declarations of the proposal-network entry points."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int, c_void_p


def declare(L: ctypes.CDLL) -> None:
    from .rjmcmc_abi import RoiBatch

    L.bamf_net_param_count.restype = c_int
    L.bamf_net_create.argtypes = [c_int, c_int, c_int, c_void_p, POINTER(c_void_p)]
    L.bamf_net_destroy.argtypes = [c_void_p]
    L.bamf_net_get_params.argtypes = [c_void_p, c_void_p]
    L.bamf_net_forward.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_void_p]
    L.bamf_net_gradients.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]
    L.bamf_net_train_step.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_float, c_void_p, c_void_p]
    L.bamf_net_proposal_maps.argtypes = [c_void_p, POINTER(RoiBatch), c_float, c_void_p, c_void_p, c_void_p]
