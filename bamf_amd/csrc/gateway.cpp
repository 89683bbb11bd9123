// Host-array gateway with cRJMCMC's argument and output layout
// (source/c/cRJMCMC.cpp:52-305) over the device chain: one ROI per call,
// host memory in and out, the device work done by bamf_rjmcmc_run.
// A MATLAB mex function (csrc/mex/cRJMCMC_gpu.cpp) only maps mxArrays onto
// this entry point.
#include <cstring>
#include <vector>

#include "common.hpp"

using bamf::set_error;

namespace {

struct DevBufs {
    std::vector<void*> ptrs;
    hipError_t err = hipSuccess;
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 1);
        if (err == hipSuccess) ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    template <class T>
    T* upload(const T* h, size_t count) {
        T* d = alloc<T>(count);
        if (err == hipSuccess && count) err = hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice);
        return d;
    }
    ~DevBufs() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

extern "C" {

int bamf_crjmcmc(const bamf_mcmc_params* mcmc, const float* data, int xsize, int ysize, int n0, const float* i0,
                 const float* x0, const float* y0, float bg0, const float* off_pdf, int off_len,
                 const float* sig_pdf, int sig_len, const float* back_pdf, int back_len, const float* scmos,
                 int opt_model, uint64_t seed, float* pgrid, float* pback, bamf_chain_rec* acc_rec,
                 bamf_chain_rec* prop_rec, float* acc_em, float* prop_em, int32_t rec_cap, int32_t em_cap,
                 int64_t counts[3], double stat[16]) {
    BAMF_CHECK_ARG(mcmc && data && off_pdf && sig_pdf && back_pdf, "null argument");
    BAMF_CHECK_ARG(xsize >= 1 && ysize >= 1 && xsize <= 32 && ysize <= 32, "ROI sides must be in 1..32");
    BAMF_CHECK_ARG(n0 >= 0 && (n0 == 0 || (i0 && x0 && y0)), "E_active needs I, X and Y for N > 0");
    BAMF_CHECK_ARG(opt_model == 0, "OptModel = 1 (the sampled PSF) is not supported");
    BAMF_CHECK_ARG(!acc_rec == !prop_rec && !acc_rec == !acc_em && !acc_rec == !prop_em,
                   "chain outputs need all four arrays");
    BAMF_CHECK_ARG(!acc_rec || (counts && rec_cap >= 0 && em_cap >= 0), "chain outputs need counts and capacities");
    bamf_rjmcmc_plan* plan = nullptr;
    int rc = bamf_rjmcmc_plan_create(mcmc, off_pdf, off_len, sig_pdf, sig_len, back_pdf, back_len, &plan);
    if (rc != BAMF_OK) return rc;
    const int P = xsize * ysize, Z = mcmc->grid_zoom;
    const int64_t grid = (int64_t)P * Z * Z;
    DevBufs b;
    const int64_t zero = 0;
    bamf_roi_batch batch;
    std::memset(&batch, 0, sizeof(batch));
    batch.n_roi = 1;
    batch.d_pix_offset = b.upload(&zero, 1);
    batch.d_xsize = b.upload(&xsize, 1);
    batch.d_ysize = b.upload(&ysize, 1);
    batch.d_data = b.upload(data, P);
    if (scmos) batch.d_scmos = b.upload(scmos, P);
    // BG = 0 becomes 1 inside the run, as cRJMCMC.cpp:132-133 does
    batch.d_bg0 = b.upload(&bg0, 1);
    if (n0 > 0) {
        batch.d_n0 = b.upload(&n0, 1);
        batch.d_em0_offset = b.upload(&zero, 1);
        batch.d_i0 = b.upload(i0, n0);
        batch.d_x0 = b.upload(x0, n0);
        batch.d_y0 = b.upload(y0, n0);
    }
    bamf_rjmcmc_out out;
    std::memset(&out, 0, sizeof(out));
    int32_t* d_status = b.alloc<int32_t>(1);
    int32_t* d_counters = b.alloc<int32_t>(30);
    out.d_status = d_status;
    out.d_counters = d_counters;
    float* d_pgrid = nullptr;
    float* d_pback = nullptr;
    if (pgrid || pback) {
        d_pgrid = b.alloc<float>(grid);
        d_pback = b.alloc<float>(grid);
        if (b.err == hipSuccess) b.err = hipMemset(d_pgrid, 0, grid * sizeof(float));
        if (b.err == hipSuccess) b.err = hipMemset(d_pback, 0, grid * sizeof(float));
        out.d_pgrid = d_pgrid;
        out.d_pback = d_pback;
        out.d_grid_offset = b.upload(&zero, 1);
    }
    if (acc_rec) {
        out.d_acc_rec = b.alloc<bamf_chain_rec>(rec_cap);
        out.d_prop_rec = b.alloc<bamf_chain_rec>(rec_cap);
        out.d_acc_em = b.alloc<float>(4 * (size_t)em_cap);
        out.d_prop_em = b.alloc<float>(4 * (size_t)em_cap);
        out.d_rec_count = b.alloc<int64_t>(3);
        out.rec_cap = rec_cap;
        out.em_cap = em_cap;
    }
    if (b.err != hipSuccess) {
        bamf_rjmcmc_plan_destroy(plan);
        set_error("gateway device setup failed: %s", hipGetErrorString(b.err));
        return BAMF_E_HIP;
    }
    rc = bamf_rjmcmc_run(plan, &batch, seed, BAMF_RJ_PLAIN, nullptr, &out, nullptr);
    if (rc == BAMF_OK && hipDeviceSynchronize() != hipSuccess) {
        set_error("chain run failed on the device");
        rc = BAMF_E_HIP;
    }
    int32_t status = 0, cnt[30];
    if (rc == BAMF_OK) {
        hipError_t e = hipMemcpy(&status, d_status, 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(cnt, d_counters, sizeof(cnt), hipMemcpyDeviceToHost);
        if (e == hipSuccess && pgrid) e = hipMemcpy(pgrid, d_pgrid, grid * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && pback) e = hipMemcpy(pback, d_pback, grid * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && acc_rec) {
            e = hipMemcpy(counts, out.d_rec_count, 3 * sizeof(int64_t), hipMemcpyDeviceToHost);
            const int64_t nr = counts[0] < rec_cap ? counts[0] : rec_cap;
            const int64_t na = counts[1] < em_cap ? counts[1] : em_cap;
            const int64_t np = counts[2] < em_cap ? counts[2] : em_cap;
            if (e == hipSuccess && nr)
                e = hipMemcpy(acc_rec, out.d_acc_rec, nr * sizeof(bamf_chain_rec), hipMemcpyDeviceToHost);
            if (e == hipSuccess && nr)
                e = hipMemcpy(prop_rec, out.d_prop_rec, nr * sizeof(bamf_chain_rec), hipMemcpyDeviceToHost);
            if (e == hipSuccess && na) e = hipMemcpy(acc_em, out.d_acc_em, na * 16, hipMemcpyDeviceToHost);
            if (e == hipSuccess && np) e = hipMemcpy(prop_em, out.d_prop_em, np * 16, hipMemcpyDeviceToHost);
        }
        if (e != hipSuccess) {
            set_error("gateway copy-back failed: %s", hipGetErrorString(e));
            rc = BAMF_E_HIP;
        }
    }
    bamf_rjmcmc_plan_destroy(plan);
    if (rc != BAMF_OK) return rc;
    if (status != 0) {
        set_error("chain ended with status %d", status);
        return status == 2 ? BAMF_E_UNSUPPORTED : BAMF_E_CAPACITY;
    }
    if (stat) {
        // Stat struct (cRJMCMC.cpp:283-305): NTrials, then A/N for jump, split,
        // merge, birth, death; the Fg set; the Bg set (counters start at N = 1)
        stat[0] = (double)mcmc->n_trials + mcmc->n_burnin;
        for (int g = 0; g < 3; g++)
            for (int k = 0; k < 5; k++) {
                const int n = cnt[10 * g + k], acc = cnt[10 * g + 5 + k];
                stat[1 + 5 * g + k] = (double)acc / (double)n;
            }
    }
    return BAMF_OK;
}

}  // extern "C"
