// MATLAB mex wrapper: SampledPSFm = findPSF_gpu(ROIStack, Gain, Offset, PSFSigma, BoxSize)
// with the arguments and output of @RJ/findPSF.m:1 (a drop-in for RJ.findPSF at
// @RJ/findPSF_SMA.m:70 and @RJ/RJ.m:328), over bamf_find_psf_host.
// Inside MATLAB: mex -R2018a findPSF_gpu.cpp -L<dir> -lbamf_gpu; bamf_amd/build.py
// also builds it with g++ against the project's mx runtime (mx/mex.h, mx/mx.cpp)
// into bamf_amd/mex/, where bamf_amd/mexsim.py calls it.
#include <cstdint>

#include "mex.h"
#include "../../../include/bamf_gpu.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs < 5 || !mxIsSingle(prhs[0]))
        mexErrMsgIdAndTxt("bamf:args", "findPSF_gpu(ROIStack single, Gain, Offset, PSFSigma, BoxSize)");
    const mwSize nd = mxGetNumberOfDimensions(prhs[0]);
    const mwSize* d = mxGetDimensions(prhs[0]);
    const int64_t n = nd >= 3 ? (int64_t)d[2] : 1;
    const int box = (int)mxGetScalar(prhs[4]);
    if ((int64_t)d[0] != box || (int64_t)d[1] != box) mexErrMsgIdAndTxt("bamf:args", "ROIStack must be BoxSize^2 x N");
    mwSize od[4] = {64, 64, 4, 4};
    plhs[0] = mxCreateNumericArray(4, od, mxSINGLE_CLASS, mxREAL);
    int32_t size[4];
    const int rc = bamf_find_psf_host((const float*)mxGetData(prhs[0]), box, n, (float)mxGetScalar(prhs[1]),
                                      (float)mxGetScalar(prhs[2]), (float)mxGetScalar(prhs[3]),
                                      (float*)mxGetData(plhs[0]), size, nullptr, nullptr);
    if (rc != BAMF_OK) mexErrMsgIdAndTxt("bamf:run", "%s", bamf_last_error());
    (void)nlhs;
}
