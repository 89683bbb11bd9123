"""The mex wrappers' host side without a GPU: each wrapper, built against the
project's mx runtime, loads and validates MATLAB-style arguments (struct
fields, classes, dimensions, strings) before any device call, raising its
mexErrMsgIdAndTxt message as BamfError."""
import numpy as np
import pytest

from bamf_amd import mexsim
from bamf_amd._lib import BamfError


def err(name, nlhs, *args):
    with pytest.raises(BamfError) as e:
        mexsim.call(name, nlhs, *args)
    return str(e.value)


def test_crjmcmc_argument_checks():
    assert "11 arguments" in err("cRJMCMC_gpu", 5, np.zeros((8, 8), np.float32))
    e_active = {"N": 0.0, "I": np.zeros(0, np.float32), "X": np.zeros(0, np.float32), "Y": np.zeros(0, np.float32),
                "BG": 10.0}
    args = [np.zeros((8, 8), np.float32), e_active, {"N_Trials": 10.0}, 8, 8] + [np.zeros(10, np.float32)] * 3 + \
        [np.zeros((0, 0), np.float32), np.zeros(1, np.float32), 0]
    assert "lacks field PSFsigma" in err("cRJMCMC_gpu", 5, *args)
    args[2] = 3.0
    assert "MCMC must be a struct" in err("cRJMCMC_gpu", 5, *args)
    args[2] = {"PSFsigma": 1.3, "N_Trials": 10.0, "N_Burnin": 10.0, "I_stdFg": 1.0, "I_stdBg": 1.0, "X_stdFg": 1.0,
               "X_stdBg": 1.0, "P_Burnin": np.zeros(4, np.float32)}
    assert "P_Burnin must be single with at least 8" in err("cRJMCMC_gpu", 5, *args)


def test_gauss_mle_argument_checks():
    d = np.zeros((7, 7, 4), np.float32)
    assert "FitType" in err("gaussMLE_gpu", 2, d, "Bogus", "CCD", 1.3)
    assert "FitType" in err("gaussMLE_gpu", 2, d, "Basic", "EMCCD", 1.3)
    assert "single" in err("gaussMLE_gpu", 2, d.astype(np.float64), "Basic", "CCD", 1.3)
    assert "SZ x SZ x N" in err("gaussMLE_gpu", 2, np.zeros((7, 6, 4), np.float32), "Basic", "CCD", 1.3)
    # 'Z' reads its calibration from the ZStruct fields by name
    z = {"Ax": 0.1, "Ay": 0.1, "Bx": 0.0, "By": 0.0, "Gamma": 0.4}
    assert "ZStruct.D is missing" in err("gaussMLE_gpu", 2, d, "Z", "CCD", np.array([1.3, 1.3]), z)
    assert "SCMOS needs" in err("gaussMLE_gpu", 2, d, "Basic", "SCMOS", 1.3)


def test_find_roi_priors_psf_argument_checks():
    assert "Image single" in err("findROI_gpu", 3, np.zeros((16, 16, 2)))
    assert "findPriors_gpu(" in err("findPriors_gpu", 3, np.zeros((16, 16), np.float32), 1.3, 5.0, 1.0, 1.0, "Yes")
    assert "BoxSize^2 x N" in err("findPSF_gpu", 1, np.zeros((8, 7, 3), np.float32), 1.0, 0.0, 1.3, 8.0)


def test_value_conversion_round_trip():
    """Struct arrays, char arrays and numeric classes survive the trip into the
    runtime (the wrapper sees them) and outputs come back column-major."""
    # a struct array (list of dicts) is accepted where a struct is read by name;
    # its first element is what the wrapper reads
    d = np.zeros((7, 7, 2), np.float32)
    z = [{"Ax": 0.1, "Ay": 0.1, "Bx": 0.0, "By": 0.0, "Gamma": 0.4}] * 2
    assert "ZStruct.D is missing" in err("gaussMLE_gpu", 2, d, "Z", "CCD", np.array([1.3, 1.3]), z)
    with pytest.raises(TypeError):
        mexsim.call("gaussMLE_gpu", 2, np.zeros(3, np.complex64), "Basic", "CCD", 1.3)
