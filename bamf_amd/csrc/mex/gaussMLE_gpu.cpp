// MATLAB mex entry with the argument list and outputs of gaussMLE for the
// fits of FitType 'Basic', 'Sigma', 'SigmaYX' and 'Z' with CameraType 'CCD' or 'SCMOS'
// (@RJ/gaussMLE.m:64-211), fitting on the GPU through bamf_gauss_mle_fit_host.
// Inside MATLAB:
//   mex -R2018a gaussMLE_gpu.cpp -I<repo>/include -L<repo>/bamf_amd -lbamf_gpu
// bamf_amd/build.py also builds it with g++ against the project's mx runtime
// (mx/mex.h, mx/mx.cpp) into bamf_amd/mex/, where bamf_amd/mexsim.py calls it.
//
// [Results, Statistics] = gaussMLE_gpu(Data, FitType, 'CCD', Sigma[, ZStruct])
// [Results, Statistics] = gaussMLE_gpu(Data, FitType, 'SCMOS', Sigma, VarianceImage, BoxCorners[, ZStruct])
// (BoxCorners as the kernels read d_Coords: the corner pair of box b at
// entries 2b, 2b+1, i.e. a 2 x N array.)
#include <cmath>
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
    const std::string ft(fit);
    const bool scmos = std::string(cam) == "SCMOS";
    if ((ft != "Basic" && ft != "Sigma" && ft != "SigmaYX" && ft != "Z") || (!scmos && std::string(cam) != "CCD"))
        mexErrMsgIdAndTxt("bamf:args", "FitType 'Basic', 'Sigma', 'SigmaYX' or 'Z' with CameraType 'CCD' or 'SCMOS'");
    const int fit_type = ft == "SigmaYX" ? BAMF_FIT_SIGMAYX
                         : ft == "Sigma" ? BAMF_FIT_SIGMA
                         : ft == "Z"     ? BAMF_FIT_Z
                                         : BAMF_FIT_BASIC;
    const int np = fit_type == BAMF_FIT_SIGMAYX ? 6 : fit_type == BAMF_FIT_BASIC ? 4 : 5;  // gaussMLE.m:93-109
    const mxArray* D = prhs[0];
    if (!mxIsSingle(D)) mexErrMsgIdAndTxt("bamf:args", "Data must be single (SZ, SZ, N)");
    const mwSize nd = mxGetNumberOfDimensions(D);
    const mwSize* dims = mxGetDimensions(D);
    const int sz = (int)dims[0];
    if (nd < 2 || dims[1] != dims[0]) mexErrMsgIdAndTxt("bamf:args", "Data must be SZ x SZ x N");
    const int64_t n = nd >= 3 ? (int64_t)dims[2] : 1;
    // gaussMLE.m:154 passes mean(Sigma); 'Z' passes Sigma(1), the ZStruct
    // fields and Sigma(2) (gaussMLE.m:155-158)
    const double* sg = mxGetPr(prhs[3]);
    double sum = 0;
    const size_t ns = mxGetNumberOfElements(prhs[3]);
    for (size_t i = 0; i < ns; i++) sum += sg[i];
    float sigma = (float)(sum / (double)(ns ? ns : 1));
    float zcal[7] = {0, 0, 0, 0, 0, 0, 0};
    // gaussMLE.m:69-84: VarianceImage and BoxCorners come before ZStruct for sCMOS
    const int zarg = scmos ? 6 : 4;
    if (fit_type == BAMF_FIT_Z) {
        if (nrhs <= zarg || !mxIsStruct(prhs[zarg]) || ns < 2)
            mexErrMsgIdAndTxt("bamf:args", "FitType 'Z' needs Sigma = [SigmaX SigmaY] and ZStruct");
        static const char* zf[6] = {"Ax", "Ay", "Bx", "By", "Gamma", "D"};
        for (int k = 0; k < 6; k++) {
            const mxArray* v = mxGetField(prhs[zarg], 0, zf[k]);
            if (!v) mexErrMsgIdAndTxt("bamf:args", "ZStruct.%s is missing", zf[k]);
            zcal[k] = (float)mxGetScalar(v);
        }
        sigma = (float)sg[0];
        zcal[6] = (float)sg[1];
    }
    std::vector<float> p(np * (size_t)n), c(np * (size_t)n), ll((size_t)n);
    double secs = 0;
    int rc;
    if (scmos) {
        if (nrhs < 6 || !mxIsSingle(prhs[4]) || !mxIsSingle(prhs[5]) ||
            mxGetNumberOfElements(prhs[5]) < 2 * (size_t)n || mxGetM(prhs[4]) != mxGetN(prhs[4]))
            mexErrMsgIdAndTxt("bamf:args", "SCMOS needs single VarianceImage (square) and BoxCorners (2 x N)");
        std::vector<float> z0((size_t)n, 0.0f);  // gaussMLE.m:167
        rc = bamf_gauss_mle_scmos_fit_host((const float*)mxGetData(D), sz, n, (const float*)mxGetData(prhs[5]),
                                           (const float*)mxGetData(prhs[4]), (int)mxGetM(prhs[4]), sigma, 20, fit_type,
                                           zcal, z0.data(), p.data(), c.data(), ll.data(), &secs);
    } else {
        rc = bamf_gauss_mle_fit_host((const float*)mxGetData(D), sz, n, sigma, 20, fit_type, zcal, p.data(), c.data(),
                                     ll.data(), &secs);
    }
    if (rc != BAMF_OK) mexErrMsgIdAndTxt("bamf:gpu", "%s", bamf_last_error());
    // Results fields in gaussMLE.m:183-196 order: parameters, their SEs, the
    // fit type's extra pair, LogLikelihood
    static const char* names_b[9] = {"Y", "X", "Photons", "BG", "Y_SE", "X_SE", "Photons_SE", "BG_SE",
                                     "LogLikelihood"};
    static const char* names_s[11] = {"Y", "X", "Photons", "BG", "Y_SE", "X_SE", "Photons_SE", "BG_SE",
                                      "Sigma", "Sigma_SE", "LogLikelihood"};
    static const char* names_z[11] = {"Y", "X", "Photons", "BG", "Y_SE", "X_SE", "Photons_SE", "BG_SE",
                                      "Z", "Z_SE", "LogLikelihood"};
    static const char* names_yx[13] = {"Y", "X", "Photons", "BG", "Y_SE", "X_SE", "Photons_SE", "BG_SE",
                                       "SigmaY", "SigmaX", "SigmaY_SE", "SigmaX_SE", "LogLikelihood"};
    const int nf = np == 6 ? 13 : np == 5 ? 11 : 9;
    plhs[0] = mxCreateStructMatrix(1, 1, nf,
                                   np == 6 ? names_yx : fit_type == BAMF_FIT_Z ? names_z : np == 5 ? names_s : names_b);
    auto column = [&](const float* src, bool root) {
        mxArray* v = mxCreateNumericMatrix((mwSize)n, 1, mxSINGLE_CLASS, mxREAL);
        float* d = (float*)mxGetData(v);
        for (int64_t i = 0; i < n; i++) d[i] = root ? std::sqrt(src[i]) : src[i];
        return v;
    };
    for (int k = 0; k < 4; k++) {
        mxSetFieldByNumber(plhs[0], 0, k, column(p.data() + (size_t)k * n, false));
        mxSetFieldByNumber(plhs[0], 0, 4 + k, column(c.data() + (size_t)k * n, true));
    }
    if (np == 5) {
        mxSetFieldByNumber(plhs[0], 0, 8, column(p.data() + 4 * (size_t)n, false));
        mxSetFieldByNumber(plhs[0], 0, 9, column(c.data() + 4 * (size_t)n, true));
    }
    if (np == 6) {  // gaussMLE.m:199-202
        mxSetFieldByNumber(plhs[0], 0, 8, column(p.data() + 4 * (size_t)n, false));
        mxSetFieldByNumber(plhs[0], 0, 9, column(p.data() + 5 * (size_t)n, false));
        mxSetFieldByNumber(plhs[0], 0, 10, column(c.data() + 4 * (size_t)n, true));
        mxSetFieldByNumber(plhs[0], 0, 11, column(c.data() + 5 * (size_t)n, true));
    }
    mxSetFieldByNumber(plhs[0], 0, nf - 1, column(ll.data(), false));
    if (nlhs > 1) {
        static const char* sn[4] = {"FitTime", "FitPerSecond", "NKernelsCalls", "NFitsPerCall"};
        plhs[1] = mxCreateStructMatrix(1, 1, 4, sn);
        mxSetFieldByNumber(plhs[1], 0, 0, mxCreateDoubleScalar(secs));
        mxSetFieldByNumber(plhs[1], 0, 1, mxCreateDoubleScalar(secs > 0 ? (double)n / secs : 0.0));
        mxSetFieldByNumber(plhs[1], 0, 2, mxCreateDoubleScalar(1.0));
        mxSetFieldByNumber(plhs[1], 0, 3, mxCreateDoubleScalar((double)n));
    }
}
