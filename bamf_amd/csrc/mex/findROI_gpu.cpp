// MATLAB mex entry with the argument list and outputs of findROI
// (@RJ/findROI.m:1-163), detecting on the GPU through bamf_find_roi_host.
// Inside MATLAB:
//   mex -R2018a findROI_gpu.cpp -I<repo>/include -L<repo>/bamf_amd -lbamf_gpu
// bamf_amd/build.py also builds it with g++ against the project's mx runtime
// (mx/mex.h, mx/mx.cpp) into bamf_amd/mex/, where bamf_amd/mexsim.py calls it.
//
// [SMD, ROIStack, Res] = findROI_gpu(Image, Sigma1, Sigma2, Boxsize, Minval)
// with findROI.m:60-75's defaults for missing arguments.
#include <cmath>
#include <cstdint>
#include <vector>

#include "bamf_gpu.h"
#include "mex.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs < 1 || !mxIsSingle(prhs[0])) mexErrMsgIdAndTxt("bamf:args", "findROI_gpu(Image single, ...)");
    const mxArray* I = prhs[0];
    const mwSize nd = mxGetNumberOfDimensions(I);
    const mwSize* dims = mxGetDimensions(I);
    const int ysize = (int)dims[0], xsize = (int)dims[1], nframes = nd >= 3 ? (int)dims[2] : 1;
    const double pi = 3.14159265358979323846;
    const float sigma1 = nrhs > 1 ? (float)mxGetScalar(prhs[1]) : 1.3f;
    const float sigma2 = nrhs > 2 ? (float)mxGetScalar(prhs[2]) : 2.0f * 1.3f;
    const int boxsize = nrhs > 3 ? (int)mxGetScalar(prhs[3]) : 7;
    const float minval = nrhs > 4 ? (float)mxGetScalar(prhs[4]) : (float)(0.25 * 2000.0 / (2 * pi * 2 * 1.3 * 1.3));
    const int64_t npix = (int64_t)ysize * xsize * nframes;
    int64_t cap = npix / 16 + 1024, count = 0;
    std::vector<int32_t> xc, yc, fr;
    std::vector<float> stack;
    std::vector<uint8_t> res((size_t)npix);
    for (int pass = 0; pass < 2; pass++) {
        xc.resize((size_t)cap);
        yc.resize((size_t)cap);
        fr.resize((size_t)cap);
        stack.resize((size_t)cap * boxsize * boxsize);
        const int rc = bamf_find_roi_host((const float*)mxGetData(I), ysize, xsize, nframes, sigma1, sigma2, boxsize,
                                          minval, cap, xc.data(), yc.data(), fr.data(), stack.data(), res.data(),
                                          &count);
        if (rc != BAMF_OK) mexErrMsgIdAndTxt("bamf:gpu", "%s", bamf_last_error());
        if (count <= cap) break;
        cap = count;
    }
    static const char* names[3] = {"XBoxCorner", "YBoxCorner", "FrameNum"};
    plhs[0] = mxCreateStructMatrix(1, 1, 3, names);
    const int32_t* src[3] = {xc.data(), yc.data(), fr.data()};
    for (int f = 0; f < 3; f++) {
        mxArray* v = mxCreateDoubleMatrix((mwSize)count, 1, mxREAL);
        double* d = mxGetPr(v);
        for (int64_t i = 0; i < count; i++) d[i] = (double)src[f][i];
        mxSetFieldByNumber(plhs[0], 0, f, v);
    }
    if (nlhs > 1) {
        const mwSize sd[3] = {(mwSize)boxsize, (mwSize)boxsize, (mwSize)count};
        // ROIStack = zeros(...) is double in findROI.m:155
        plhs[1] = mxCreateNumericArray(3, sd, mxDOUBLE_CLASS, mxREAL);
        double* d = mxGetPr(plhs[1]);
        for (size_t i = 0; i < (size_t)count * boxsize * boxsize; i++) d[i] = stack[i];
    }
    if (nlhs > 2) {
        plhs[2] = mxCreateNumericArray(nd, dims, mxDOUBLE_CLASS, mxREAL);
        double* d = mxGetPr(plhs[2]);
        for (int64_t i = 0; i < npix; i++) d[i] = res[(size_t)i];
    }
}
