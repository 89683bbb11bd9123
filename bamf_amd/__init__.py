"""bamf_amd — MI355X (gfx950) implementation of BAMF's data-parallel hot path.

The compute lives in libbamf_gpu.so (hand-written HIP kernels behind the C
ABI of include/bamf_gpu.h).  This package is the host-side mirror of the
reference's MATLAB-level interfaces:

    gauss_mle.gauss_mle   <- @RJ/gaussMLE.m   (fit type 'Basic')
    find_roi.find_roi     <- @RJ/findROI.m
    rjmcmc.rjmcmc         <- @RJ/rjmcmc.m     (-> cRJMCMC mex layout)
    rjmcmc.RJMCMCBatch    batched chains over many ROIs (persistent kernel)

There is no CPU fallback: if the library is missing every call raises.
"""
from ._lib import BamfError, lib  # noqa: F401

__all__ = ["BamfError", "lib"]
