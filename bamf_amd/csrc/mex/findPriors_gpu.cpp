// MATLAB mex entry for the prior construction of @RJ/findPriors.m:35-80 on
// the GPU through bamf_find_priors_host.  Inside MATLAB:
//   mex -R2018a findPriors_gpu.cpp -I<repo>/include -L<repo>/bamf_amd -lbamf_gpu
// bamf_amd/build.py also builds it with g++ against the project's mx runtime
// (mx/mex.h, mx/mx.cpp) into bamf_amd/mex/, where bamf_amd/mexsim.py calls it.
//
// [P_EmPhotons, P_BgEmPhotons, P_Offset] = findPriors_gpu(Data, PSF_Sigma, SMAStruct, P_dP, BgPriorRange, Y_N_BgPrior)
// (the RJStruct fields findPriors.m:111-113 sets; SMAStruct: MeanPhotons,
// MinPValue, MinPhotons, optional MaxXY_SE)
#include <cstdint>
#include <string>
#include <vector>

#include "bamf_gpu.h"
#include "mex.h"

static double field(const mxArray* s, const char* name, double dflt, bool required) {
    const mxArray* v = mxGetField(s, 0, name);
    if (!v) {
        if (required) mexErrMsgIdAndTxt("bamf:args", "SMAStruct.%s is missing", name);
        return dflt;
    }
    return mxGetScalar(v);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs < 6 || !mxIsSingle(prhs[0]) || !mxIsStruct(prhs[2]))
        mexErrMsgIdAndTxt("bamf:args",
                          "findPriors_gpu(single Data, PSF_Sigma, SMAStruct, P_dP, BgPriorRange, Y_N_BgPrior)");
    const mwSize nd = mxGetNumberOfDimensions(prhs[0]);
    const mwSize* dims = mxGetDimensions(prhs[0]);
    const int ys = (int)dims[0], xs = (int)dims[1], nf = nd >= 3 ? (int)dims[2] : 1;
    char yn[8] = {0};
    mxGetString(prhs[5], yn, sizeof(yn));
    const int64_t cap = 1 << 22;
    std::vector<double> sig((size_t)cap), bg((size_t)cap), off((size_t)cap);
    int64_t len = 0;
    double info[8];
    const int rc = bamf_find_priors_host(
        (const float*)mxGetData(prhs[0]), ys, xs, nf, (float)mxGetScalar(prhs[1]),
        (float)field(prhs[2], "MeanPhotons", 0, true), (float)field(prhs[2], "MaxXY_SE", 0.2, false),
        (float)field(prhs[2], "MinPValue", 0, true), (float)field(prhs[2], "MinPhotons", 0, true),
        (float)mxGetScalar(prhs[3]), (float)mxGetScalar(prhs[4]), std::string(yn) == "Yes" ? 1 : 0, sig.data(),
        bg.data(), off.data(), cap, &len, info);
    if (rc != BAMF_OK) mexErrMsgIdAndTxt("bamf:gpu", "%s", bamf_last_error());
    if (len > cap) mexErrMsgIdAndTxt("bamf:gpu", "prior grid of %lld points exceeds %lld", (long long)len,
                                     (long long)cap);
    const std::vector<double>* outs[3] = {&sig, &bg, &off};
    for (int k = 0; k < 3 && k < (nlhs > 0 ? nlhs : 1); k++) {
        plhs[k] = mxCreateDoubleMatrix(1, (mwSize)len, mxREAL);
        double* d = mxGetPr(plhs[k]);
        for (int64_t i = 0; i < len; i++) d[i] = (*outs[k])[(size_t)i];
    }
}
