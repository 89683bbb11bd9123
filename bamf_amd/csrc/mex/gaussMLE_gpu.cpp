// MATLAB mex entry with the argument list and outputs of gaussMLE for the
// fit the hot path uses (@RJ/gaussMLE.m:64-211, FitType 'Basic', CameraType
// 'CCD'), fitting on the GPU through bamf_gauss_mle_host.  Build inside
// MATLAB (mex.h is needed; not compiled by bamf_amd/build.py):
//   mex -R2018a gaussMLE_gpu.cpp -I<repo>/include -L<repo>/bamf_amd -lbamf_gpu
//
// [Results, Statistics] = gaussMLE_gpu(Data, FitType, CameraType, Sigma)
#include <cstdint>
#include <string>
#include <vector>

#include "bamf_gpu.h"
#include "mex.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs < 4) mexErrMsgIdAndTxt("bamf:args", "gaussMLE_gpu(Data, FitType, CameraType, Sigma)");
    char fit[32] = {0}, cam[32] = {0};
    mxGetString(prhs[1], fit, sizeof(fit));
    mxGetString(prhs[2], cam, sizeof(cam));
    if (std::string(fit) != "Basic" || std::string(cam) != "CCD")
        mexErrMsgIdAndTxt("bamf:args", "only FitType 'Basic' with CameraType 'CCD' is implemented");
    const mxArray* D = prhs[0];
    if (!mxIsSingle(D)) mexErrMsgIdAndTxt("bamf:args", "Data must be single (SZ, SZ, N)");
    const mwSize nd = mxGetNumberOfDimensions(D);
    const mwSize* dims = mxGetDimensions(D);
    const int sz = (int)dims[0];
    if (nd < 2 || dims[1] != dims[0]) mexErrMsgIdAndTxt("bamf:args", "Data must be SZ x SZ x N");
    const int64_t n = nd >= 3 ? (int64_t)dims[2] : 1;
    // gaussMLE.m:154 passes mean(Sigma)
    const double* sg = mxGetPr(prhs[3]);
    double sum = 0;
    const size_t ns = mxGetNumberOfElements(prhs[3]);
    for (size_t i = 0; i < ns; i++) sum += sg[i];
    const float sigma = (float)(sum / (double)(ns ? ns : 1));
    std::vector<float> out(9 * (size_t)n);
    double secs = 0;
    float* o = out.data();
    const int rc = bamf_gauss_mle_host((const float*)mxGetData(D), sz, n, sigma, 20, o, o + n, o + 2 * n, o + 3 * n,
                                       o + 4 * n, o + 5 * n, o + 6 * n, o + 7 * n, o + 8 * n, &secs);
    if (rc != BAMF_OK) mexErrMsgIdAndTxt("bamf:gpu", "%s", bamf_last_error());
    // Results fields in gaussMLE.m:183-193 order
    static const char* names[9] = {"Y", "X", "Photons", "BG", "Y_SE", "X_SE", "Photons_SE", "BG_SE", "LogLikelihood"};
    plhs[0] = mxCreateStructMatrix(1, 1, 9, names);
    for (int f = 0; f < 9; f++) {
        mxArray* v = mxCreateNumericMatrix((mwSize)n, 1, mxSINGLE_CLASS, mxREAL);
        float* d = (float*)mxGetData(v);
        for (int64_t i = 0; i < n; i++) d[i] = o[(size_t)f * n + i];
        mxSetFieldByNumber(plhs[0], 0, f, v);
    }
    if (nlhs > 1) {
        static const char* sn[4] = {"FitTime", "FitPerSecond", "NKernelsCalls", "NFitsPerCall"};
        plhs[1] = mxCreateStructMatrix(1, 1, 4, sn);
        mxSetFieldByNumber(plhs[1], 0, 0, mxCreateDoubleScalar(secs));
        mxSetFieldByNumber(plhs[1], 0, 1, mxCreateDoubleScalar(secs > 0 ? (double)n / secs : 0.0));
        mxSetFieldByNumber(plhs[1], 0, 2, mxCreateDoubleScalar(1.0));
        mxSetFieldByNumber(plhs[1], 0, 3, mxCreateDoubleScalar((double)n));
    }
}
