// Host-array gateways with the argument and output layouts of the three
// MATLAB-level entry points, each over the device entry points:
//   bamf_crjmcmc        cRJMCMC (source/c/cRJMCMC.cpp:52-305), one ROI per call;
//   bamf_gauss_mle_host gaussMLE(Data, 'Basic', 'CCD', Sigma) (@RJ/gaussMLE.m:64-211);
//   bamf_gauss_mle_fit_host the same for FitType 'Basic', 'Sigma', 'SigmaYX' or 'Z', raw outputs;
//   bamf_find_roi_host  findROI(Image, Sigma1, Sigma2, Boxsize, Minval) (@RJ/findROI.m:1-163).
// The MATLAB mex functions under csrc/mex only map mxArrays onto these.
#include <algorithm>
#include <chrono>
#include <cmath>
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
        if (err != hipSuccess) p = nullptr;
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

namespace {

// gaussMLE.m:113-180 on host arrays: fits in calls of at most 2^24 (the
// reference splits by free device memory the same way; fits are independent).
// coords/gain non-null: the sCMOS kernels.
int mle_fit_host(const float* data, int sz, int64_t nfits, const float* coords, const float* gain, int mapsz,
                 float sigma, int iterations, int fit_type, const float* zcal, const float* z0, float* params,
                 float* crlb, float* ll, double* fit_seconds) {
    BAMF_CHECK_ARG(data && params && crlb && ll, "null argument");
    BAMF_CHECK_ARG(sz >= 1 && nfits >= 0 && iterations >= 0, "bad size");
    BAMF_CHECK_ARG(fit_type >= BAMF_FIT_BASIC && fit_type <= BAMF_FIT_Z, "fit_type must be BASIC, SIGMA, SIGMAYX or Z");
    BAMF_CHECK_ARG(fit_type != BAMF_FIT_Z || zcal, "fit type Z needs its calibration");
    const bool sc = coords != nullptr;
    BAMF_CHECK_ARG(!sc || (gain && mapsz >= sz && (fit_type != BAMF_FIT_Z || z0)), "sCMOS fits need the gain map");
    const int np = fit_type == BAMF_FIT_SIGMAYX ? 6 : fit_type == BAMF_FIT_BASIC ? 4 : 5;
    if (nfits == 0) {
        if (fit_seconds) *fit_seconds = 0;
        return BAMF_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t chunk = std::min<int64_t>(nfits, int64_t(1) << 24);
    DevBufs b;
    float* d_data = b.alloc<float>((size_t)chunk * sz * sz);
    float* d_p = b.alloc<float>(np * (size_t)chunk);
    float* d_c = b.alloc<float>(np * (size_t)chunk);
    float* d_l = b.alloc<float>((size_t)chunk);
    float* d_xy = sc ? b.alloc<float>(2 * (size_t)chunk) : nullptr;
    float* d_z0 = sc && fit_type == BAMF_FIT_Z ? b.alloc<float>((size_t)chunk) : nullptr;
    const float* d_gain = sc ? b.upload(gain, (size_t)mapsz * mapsz) : nullptr;
    std::vector<float> hp(np * (size_t)chunk), hc(np * (size_t)chunk);
    for (int64_t f0 = 0; f0 < nfits && b.err == hipSuccess; f0 += chunk) {
        const int n = (int)std::min<int64_t>(chunk, nfits - f0);
        b.err = hipMemcpy(d_data, data + (size_t)f0 * sz * sz, (size_t)n * sz * sz * 4, hipMemcpyHostToDevice);
        if (b.err == hipSuccess && d_xy) b.err = hipMemcpy(d_xy, coords + 2 * f0, 2 * (size_t)n * 4, hipMemcpyHostToDevice);
        if (b.err == hipSuccess && d_z0) b.err = hipMemcpy(d_z0, z0 + f0, (size_t)n * 4, hipMemcpyHostToDevice);
        if (b.err != hipSuccess) break;
        int rc;
        if (sc) {
            switch (fit_type) {
            case BAMF_FIT_SIGMA:
                rc = bamf_gauss_mle_scmos_xynbs(d_data, d_xy, d_gain, sigma, sz, mapsz, iterations, d_p, d_c, d_l, n,
                                                nullptr);
                break;
            case BAMF_FIT_SIGMAYX:
                rc = bamf_gauss_mle_scmos_xynbsxsy(d_data, d_xy, d_gain, sigma, sz, mapsz, iterations, d_p, d_c, d_l,
                                                   n, nullptr);
                break;
            case BAMF_FIT_Z:
                rc = bamf_gauss_mle_scmos_xynbz(d_data, d_xy, d_gain, d_z0, sigma, zcal[0], zcal[1], zcal[2], zcal[3],
                                                zcal[4], zcal[5], zcal[6], sz, mapsz, iterations, d_p, d_c, d_l, n,
                                                nullptr);
                break;
            default:
                rc = bamf_gauss_mle_scmos_xynb(d_data, d_xy, d_gain, sigma, sz, mapsz, iterations, d_p, d_c, d_l, n,
                                               nullptr);
            }
        } else {
            switch (fit_type) {
            case BAMF_FIT_SIGMA: rc = bamf_gauss_mle_xynbs(d_data, sigma, sz, iterations, d_p, d_c, d_l, n, nullptr); break;
            case BAMF_FIT_SIGMAYX:
                rc = bamf_gauss_mle_xynbsxsy(d_data, sigma, sz, iterations, d_p, d_c, d_l, n, nullptr);
                break;
            case BAMF_FIT_Z:
                rc = bamf_gauss_mle_xynbz(d_data, sigma, zcal[0], zcal[1], zcal[2], zcal[3], zcal[4], zcal[5],
                                          zcal[6], sz, iterations, d_p, d_c, d_l, n, nullptr);
                break;
            default: rc = bamf_gauss_mle_xynb(d_data, sigma, sz, iterations, d_p, d_c, d_l, n, nullptr);
            }
        }
        if (rc != BAMF_OK) return rc;
        b.err = hipMemcpy(hp.data(), d_p, np * (size_t)n * 4, hipMemcpyDeviceToHost);
        if (b.err == hipSuccess) b.err = hipMemcpy(hc.data(), d_c, np * (size_t)n * 4, hipMemcpyDeviceToHost);
        if (b.err == hipSuccess) b.err = hipMemcpy(ll + f0, d_l, (size_t)n * 4, hipMemcpyDeviceToHost);
        if (b.err != hipSuccess) break;
        for (int k = 0; k < np; k++) {  // parameter-major over all nfits
            std::memcpy(params + (size_t)k * nfits + f0, hp.data() + (size_t)k * n, (size_t)n * 4);
            std::memcpy(crlb + (size_t)k * nfits + f0, hc.data() + (size_t)k * n, (size_t)n * 4);
        }
    }
    if (b.err != hipSuccess) {
        set_error("gaussMLE gateway: %s", hipGetErrorString(b.err));
        return BAMF_E_HIP;
    }
    if (fit_seconds) fit_seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return BAMF_OK;
}

}  // namespace

extern "C" {

int bamf_gauss_mle_fit_host(const float* data, int sz, int64_t nfits, float sigma, int iterations, int fit_type,
                            const float* zcal, float* params, float* crlb, float* ll, double* fit_seconds) {
    return mle_fit_host(data, sz, nfits, nullptr, nullptr, 0, sigma, iterations, fit_type, zcal, nullptr, params,
                        crlb, ll, fit_seconds);
}

int bamf_gauss_mle_scmos_fit_host(const float* data, int sz, int64_t nfits, const float* coords,
                                  const float* gain_ratio, int mapsz, float sigma, int iterations, int fit_type,
                                  const float* zcal, const float* z0, float* params, float* crlb, float* ll,
                                  double* fit_seconds) {
    BAMF_CHECK_ARG(coords && gain_ratio, "sCMOS fits need box corners and the gain map");
    return mle_fit_host(data, sz, nfits, coords, gain_ratio, mapsz, sigma, iterations, fit_type, zcal, z0, params,
                        crlb, ll, fit_seconds);
}

int bamf_gauss_mle_host(const float* data, int sz, int64_t nfits, float sigma, int iterations, float* y, float* x,
                        float* photons, float* bg, float* y_se, float* x_se, float* photons_se, float* bg_se,
                        float* ll, double* fit_seconds) {
    BAMF_CHECK_ARG(data && y && x && photons && bg && ll, "null argument");
    BAMF_CHECK_ARG(sz >= 1 && nfits >= 0 && iterations >= 0, "bad size");
    std::vector<float> p(4 * (size_t)nfits), c(4 * (size_t)nfits);
    const int rc = bamf_gauss_mle_fit_host(data, sz, nfits, sigma, iterations, BAMF_FIT_BASIC, nullptr, p.data(),
                                           c.data(), ll, fit_seconds);
    if (rc != BAMF_OK) return rc;
    // gaussMLE.m:183-190: parameter k = 0..3 is Y, X, Photons, BG; SE = sqrt(CRLB)
    float* out[4] = {y, x, photons, bg};
    float* se[4] = {y_se, x_se, photons_se, bg_se};
    for (int k = 0; k < 4; k++) {
        std::memcpy(out[k], p.data() + (size_t)k * nfits, (size_t)nfits * 4);
        if (se[k])
            for (int64_t i = 0; i < nfits; i++) se[k][i] = std::sqrt(c[(size_t)k * nfits + i]);
    }
    return BAMF_OK;
}

int bamf_find_roi_host(const float* image, int ysize, int xsize, int nframes, float sigma1, float sigma2,
                       int boxsize, float minval, int64_t capacity, int32_t* x_corner, int32_t* y_corner,
                       int32_t* frame_num, float* roi_stack, uint8_t* res, int64_t* count) {
    BAMF_CHECK_ARG(image && count, "null argument");
    BAMF_CHECK_ARG(xsize >= 1 && ysize >= 1 && nframes >= 0 && boxsize >= 1 && capacity >= 0, "bad size");
    *count = 0;
    if (nframes == 0) return BAMF_OK;
    const int64_t frame = (int64_t)xsize * ysize;
    // frames per device pass: findROI.m:98-104 sizes its chunks as 16 bytes
    // per pixel of device memory; here the chunk is bounded by the free
    // memory and by 2^31 pixels.  The filters act frame by frame, so the
    // chunking does not change any result; detections keep find() order
    // because chunks are consecutive frame ranges.
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        set_error("findROI gateway: hipMemGetInfo failed");
        return BAMF_E_HIP;
    }
    const size_t per_frame = (size_t)frame * 4 * 4 + (size_t)frame / 8 + 64;
    int64_t fpc = (int64_t)((free_b / 2) / per_frame);
    fpc = std::min<int64_t>(fpc, ((int64_t)1 << 31) / frame);
    fpc = std::max<int64_t>(1, std::min<int64_t>(fpc, nframes));
    const size_t ws_bytes = bamf_find_roi_workspace_bytes(xsize, ysize, (int)fpc);
    const int64_t npix_max = frame * fpc;
    DevBufs b;
    float* d_stack = b.alloc<float>((size_t)npix_max);
    void* d_ws = b.alloc<uint8_t>(ws_bytes);
    uint32_t* d_bits = res ? b.alloc<uint32_t>((size_t)(npix_max + 31) / 32) : nullptr;
    const int64_t cap = std::max<int64_t>(capacity, 1);
    int32_t* d_row = b.alloc<int32_t>((size_t)cap);
    int32_t* d_col = b.alloc<int32_t>((size_t)cap);
    int32_t* d_frm = b.alloc<int32_t>((size_t)cap);
    int64_t* d_count = b.alloc<int64_t>(1);
    float* d_crop = roi_stack ? b.alloc<float>((size_t)cap * boxsize * boxsize) : nullptr;
    std::vector<uint32_t> bits;
    int64_t total = 0;
    for (int f0 = 0; f0 < nframes && b.err == hipSuccess; f0 += (int)fpc) {
        const int nf = (int)std::min<int64_t>(fpc, nframes - f0);
        const int64_t npix = frame * nf;
        b.err = hipMemcpy(d_stack, image + (size_t)f0 * frame, (size_t)npix * 4, hipMemcpyHostToDevice);
        if (b.err != hipSuccess) break;
        const int64_t room = std::max<int64_t>(capacity - total, 0);
        int rc = bamf_find_roi(d_stack, xsize, ysize, nf, sigma1, sigma2, boxsize, minval, d_ws, ws_bytes, d_bits,
                               nullptr, d_row, d_col, d_frm, room, d_count, nullptr);
        if (rc != BAMF_OK) return rc;
        int64_t n = 0;
        b.err = hipMemcpy(&n, d_count, 8, hipMemcpyDeviceToHost);
        if (b.err != hipSuccess) break;
        const int64_t m = std::min(n, room);
        if (m > 0) {
            if (y_corner) b.err = hipMemcpy(y_corner + total, d_row, m * 4, hipMemcpyDeviceToHost);
            if (b.err == hipSuccess && x_corner) b.err = hipMemcpy(x_corner + total, d_col, m * 4, hipMemcpyDeviceToHost);
            if (b.err == hipSuccess && frame_num) {
                b.err = hipMemcpy(frame_num + total, d_frm, m * 4, hipMemcpyDeviceToHost);
                for (int64_t i = 0; i < m; i++) frame_num[total + i] += f0;
            }
            if (b.err == hipSuccess && roi_stack) {
                rc = bamf_find_roi_crop(d_stack, xsize, ysize, d_row, d_col, d_frm, m, boxsize, d_crop, nullptr);
                if (rc != BAMF_OK) return rc;
                b.err = hipMemcpy(roi_stack + (size_t)total * boxsize * boxsize, d_crop,
                                  (size_t)m * boxsize * boxsize * 4, hipMemcpyDeviceToHost);
            }
        }
        if (b.err == hipSuccess && res) {
            bits.resize((size_t)(npix + 31) / 32);
            b.err = hipMemcpy(bits.data(), d_bits, bits.size() * 4, hipMemcpyDeviceToHost);
            for (int64_t i = 0; i < npix; i++) res[(size_t)f0 * frame + i] = (bits[i >> 5] >> (i & 31)) & 1u;
        }
        total += n;
    }
    if (b.err != hipSuccess) {
        set_error("findROI gateway: %s", hipGetErrorString(b.err));
        return BAMF_E_HIP;
    }
    *count = total;
    return BAMF_OK;
}

}  // extern "C"
