// MATLAB mex entry with the argument list and outputs of cRJMCMC
// (source/c/cRJMCMC.cpp:52-305), running the chain on the GPU through
// bamf_crjmcmc.  Inside MATLAB:
//   mex -R2018a cRJMCMC_gpu.cpp -I<repo>/include -L<repo>/bamf_amd -lbamf_gpu
// bamf_amd/build.py also builds it with g++ against the project's mx runtime
// (mx/mex.h, mx/mx.cpp) into bamf_amd/mex/, where bamf_amd/mexsim.py calls it.
//
// [PGrid, PBack, Chain, ProposedChain, Stat] =
//     cRJMCMC_gpu(Data, E_active, MCMC, Xsize, Ysize, OffSetPDF, SignalPDF, BackPDF,
//                 sCMOScorr, SampledPSF, OptModel [, Seed])
#include <cstdint>
#include <vector>

#include "bamf_gpu.h"
#include "mex.h"

namespace {

float fieldf(const mxArray* s, const char* f) {
    const mxArray* v = mxGetField(s, 0, f);
    if (!v) mexErrMsgIdAndTxt("bamf:field", "MCMC struct lacks field %s", f);
    return (float)mxGetScalar(v);
}

// a single-precision field of at least n elements (n = 0: may be empty)
const float* single_field(const mxArray* s, const char* f, size_t n) {
    const mxArray* v = mxGetField(s, 0, f);
    if (!v || !mxIsSingle(v) || mxGetNumberOfElements(v) < n)
        mexErrMsgIdAndTxt("bamf:field", "field %s must be single with at least %zu elements", f, n);
    return (const float*)mxGetData(v);
}

int longest_side(const mxArray* a) { return (int)(mxGetM(a) > mxGetN(a) ? mxGetM(a) : mxGetN(a)); }

// Chain struct array of NLoop entries, the first NLoop - 1 filled
// (cRJMCMC.cpp:159-229).  LLR, PR, JumpType and Accepted come from `meta`:
// the accepted chain's records for both outputs, as cRJMCMC.cpp:265-272 does.
mxArray* chain_struct(const std::vector<bamf_chain_rec>& rec, const std::vector<bamf_chain_rec>& meta,
                      const std::vector<float>& em, int64_t nrec) {
    static const char* names[12] = {"N", "X", "Y", "Photons", "Signal", "LLR", "PR", "JumpType", "Accepted",
                                    "BG", "ABG", "BBG"};
    const mwSize n = (mwSize)(nrec > 0 ? nrec : 0);
    mxArray* s = mxCreateStructMatrix(n, 1, 12, names);
    for (int64_t k = 0; k + 1 < nrec; k++) {
        const bamf_chain_rec& h = rec[k];
        const int nn = h.n < 1 ? 1 : h.n;  // an empty state keeps its one placeholder entry
        const float* e = &em[4 * (size_t)h.em_index];
        mxSetFieldByNumber(s, k, 0, mxCreateDoubleScalar(e[2] == 0.0f ? 0.0 : (double)h.n));
        for (int f = 0; f < 4; f++) {
            mxArray* v = mxCreateNumericMatrix(1, nn, mxDOUBLE_CLASS, mxREAL);
            double* d = (double*)mxGetData(v);
            static const int src[4] = {0, 1, 2, 3};  // X, Y, Photons, Signal
            for (int i = 0; i < nn; i++) d[i] = (double)e[4 * i + src[f]];
            mxSetFieldByNumber(s, k, 1 + f, v);
        }
        const bamf_chain_rec& m = meta[k];
        mxSetFieldByNumber(s, k, 5, mxCreateDoubleScalar((double)m.llr));
        mxSetFieldByNumber(s, k, 6, mxCreateDoubleScalar((double)m.pr));
        mxSetFieldByNumber(s, k, 7, mxCreateDoubleScalar((double)m.jump_type));
        mxSetFieldByNumber(s, k, 8, mxCreateDoubleScalar((double)m.accepted));
        mxSetFieldByNumber(s, k, 9, mxCreateDoubleScalar((double)h.bg));
        mxSetFieldByNumber(s, k, 10, mxCreateDoubleScalar((double)h.abg));
        mxSetFieldByNumber(s, k, 11, mxCreateDoubleScalar((double)h.bbg));
    }
    return s;
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs < 11) mexErrMsgIdAndTxt("bamf:args", "cRJMCMC_gpu needs the 11 arguments of cRJMCMC");
    if (!mxIsStruct(prhs[2])) mexErrMsgIdAndTxt("bamf:args", "MCMC must be a struct");
    for (int k : {0, 5, 6, 7, 8, 9})
        if (!mxIsSingle(prhs[k])) mexErrMsgIdAndTxt("bamf:args", "argument %d must be single", k + 1);
    const mxArray* E = prhs[1];
    const mxArray* M = prhs[2];
    bamf_mcmc_params p = {};
    p.psf_sigma = fieldf(M, "PSFsigma");
    p.n_trials = (int32_t)fieldf(M, "N_Trials");
    p.n_burnin = (int32_t)fieldf(M, "N_Burnin");
    p.i_std_fg = fieldf(M, "I_stdFg");
    p.i_std_bg = fieldf(M, "I_stdBg");
    p.x_std_fg = fieldf(M, "X_stdFg");
    p.x_std_bg = fieldf(M, "X_stdBg");
    const float* pb = single_field(M, "P_Burnin", 8);  // single[8], as make_Params reads
    const float* pt = single_field(M, "P_Trials", 8);
    for (int i = 0; i < 8; i++) {
        p.p_burnin[i] = pb[i];
        p.p_trials[i] = pt[i];
    }
    p.rho = fieldf(M, "Rho");
    p.bnd_out = fieldf(M, "Bnd_out");
    p.grid_zoom = (int32_t)fieldf(M, "Grid_Zoom");
    p.split_std = fieldf(M, "Split_std");
    p.icutoff = fieldf(M, "Icutoff");
    p.bg_std = fieldf(M, "BG_std");
    p.abbg_std = fieldf(M, "ABBG_std");
    p.dx = fieldf(M, "DX");
    p.drift_x = fieldf(M, "DriftX");
    p.drift_y = fieldf(M, "DriftY");
    p.is_backg = (int32_t)fieldf(M, "IsBackg");

    const int xsize = (int)mxGetScalar(prhs[3]), ysize = (int)mxGetScalar(prhs[4]);
    if (!mxIsStruct(E) || !mxGetField(E, 0, "N") || !mxGetField(E, 0, "BG"))
        mexErrMsgIdAndTxt("bamf:field", "E_active needs fields N, I, X, Y and BG");
    const int n0 = (int)mxGetScalar(mxGetField(E, 0, "N"));
    const float* ei = single_field(E, "I", (size_t)(n0 > 0 ? n0 : 0));
    const float* ex = single_field(E, "X", (size_t)(n0 > 0 ? n0 : 0));
    const float* ey = single_field(E, "Y", (size_t)(n0 > 0 ? n0 : 0));
    const uint64_t seed = nrhs > 11 ? (uint64_t)mxGetScalar(prhs[11]) : 0;
    // SampledPSF (prhs[9]) with its size from MCMC.PSFsize (cRJMCMC.cpp:35), as
    // rjmcmc.m:91 sets it; the array's own dimensions when the field is absent
    const int opt_model = (int)mxGetScalar(prhs[10]);
    int32_t psf_size[4] = {1, 1, 1, 1};
    const mxArray* ps = mxGetField(M, 0, "PSFsize");
    if (ps && mxIsInt32(ps) && mxGetNumberOfElements(ps) >= 4) {
        for (int k = 0; k < 4; k++) psf_size[k] = (int32_t)((const int32_t*)mxGetData(ps))[k];
    } else {
        const mwSize nd = mxGetNumberOfDimensions(prhs[9]);
        const mwSize* d = mxGetDimensions(prhs[9]);
        for (mwSize k = 0; k < nd && k < 4; k++) psf_size[k] = (int32_t)d[k];
    }
    if (opt_model == 1 && (!mxIsSingle(prhs[9]) ||
                           (int64_t)mxGetNumberOfElements(prhs[9]) !=
                               (int64_t)psf_size[0] * psf_size[1] * psf_size[2] * psf_size[3]))
        mexErrMsgIdAndTxt("bamf:args", "SampledPSF must be single with PSFsize elements");
    const size_t npx = (size_t)(xsize > 0 ? xsize : 0) * (size_t)(ysize > 0 ? ysize : 0);
    if (mxGetNumberOfElements(prhs[0]) != npx)
        mexErrMsgIdAndTxt("bamf:args", "Data must hold Xsize * Ysize pixels");
    if (mxGetNumberOfElements(prhs[8]) != 0 && mxGetNumberOfElements(prhs[8]) != npx)
        mexErrMsgIdAndTxt("bamf:args", "sCMOScorr must be empty or hold Xsize * Ysize pixels");
    const int64_t total = (int64_t)p.n_burnin + p.n_trials;
    const int32_t rec_cap = (int32_t)(total + 1);
    const int32_t em_cap = (int32_t)(total * 16 + 1024);
    std::vector<bamf_chain_rec> acc(rec_cap), prop(rec_cap);
    std::vector<float> acc_em(4 * (size_t)em_cap), prop_em(4 * (size_t)em_cap);
    int64_t counts[3] = {0, 0, 0};
    double stat[16];
    mwSize dims[2] = {(mwSize)ysize * p.grid_zoom, (mwSize)xsize * p.grid_zoom};
    plhs[0] = mxCreateNumericArray(2, dims, mxSINGLE_CLASS, mxREAL);
    plhs[1] = mxCreateNumericArray(2, dims, mxSINGLE_CLASS, mxREAL);
    const int rc = bamf_crjmcmc(
        &p, (const float*)mxGetData(prhs[0]), xsize, ysize, n0, ei, ex, ey, (float)mxGetScalar(mxGetField(E, 0, "BG")),
        (const float*)mxGetData(prhs[5]), longest_side(prhs[5]),
        (const float*)mxGetData(prhs[6]), longest_side(prhs[6]), (const float*)mxGetData(prhs[7]),
        longest_side(prhs[7]), (const float*)mxGetData(prhs[8]), (const float*)mxGetData(prhs[9]), psf_size,
        opt_model, seed,
        (float*)mxGetData(plhs[0]), (float*)mxGetData(plhs[1]), acc.data(), prop.data(), acc_em.data(),
        prop_em.data(), rec_cap, em_cap, counts, stat);
    if (rc != BAMF_OK) mexErrMsgIdAndTxt("bamf:run", "%s", bamf_last_error());
    if (counts[1] > em_cap || counts[2] > em_cap) mexErrMsgIdAndTxt("bamf:cap", "chain record capacity exceeded");
    plhs[2] = chain_struct(acc, acc, acc_em, counts[0]);
    plhs[3] = chain_struct(prop, acc, prop_em, counts[0]);
    static const char* sn[16] = {"NTrials", "Accept_Jump", "Accept_Split", "Accept_Merge", "Accept_Birth",
                                 "Accept_Death", "Accept_JumpFg", "Accept_SplitFg", "Accept_MergeFg",
                                 "Accept_BirthFg", "Accept_DeathFg", "Accept_JumpBg", "Accept_SplitBg",
                                 "Accept_MergeBg", "Accept_BirthBg", "Accept_DeathBg"};
    plhs[4] = mxCreateStructMatrix(1, 1, 16, sn);
    for (int i = 0; i < 16; i++) mxSetFieldByNumber(plhs[4], 0, i, mxCreateDoubleScalar(stat[i]));
    (void)nlhs;
}
