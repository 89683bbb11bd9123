"""bamf_amd — MI355X (gfx950) implementation of BAMF's data-parallel hot path.

The compute lives in libbamf_gpu.so (hand-written HIP kernels behind the C
ABI of include/bamf_gpu.h).  This package is the host-side mirror of the
reference's MATLAB-level interfaces:

    gauss_mle.gauss_mle   <- @RJ/gaussMLE.m   (Basic, Sigma, SigmaYX, Z; CCD, SCMOS)
    find_roi.find_roi     <- @RJ/findROI.m
    priors.find_priors    <- @RJ/findPriors.m
    psf.find_psf          <- @RJ/findPSF.m
    tiles.Tiling          <- @RJ/makeSubRegions.m, removeOverlapping.m, stitchROIs.m
    rjmcmc.crjmcmc        <- cRJMCMC (source/c/cRJMCMC.cpp, the mex layout)
    rjmcmc.rjmcmc         <- @RJ/rjmcmc.m
    rjmcmc.run_batch      batched chains over many ROIs (persistent kernel)
    rjmcmc.find_map       <- @RJ/findMap.m, batched
    refine.mcmc           <- @RJ/mcmc.m
    refine.analyze_tiles  <- @RJ/analyzeROIs.m's tile loop and ClustInfo
    mexsim.call           the mex wrappers of csrc/mex, run without MATLAB

There is no CPU fallback: if the library is missing every call raises.
"""
from ._lib import BamfError, lib  # noqa: F401

__all__ = ["BamfError", "lib"]
