// Host-array gateways with the argument and output layouts of the three
// MATLAB-level entry points, each over the device entry points:
//   bamf_crjmcmc        cRJMCMC (source/c/cRJMCMC.cpp:52-305), one ROI per call;
//   bamf_gauss_mle_host gaussMLE(Data, 'Basic', 'CCD', Sigma) (@RJ/gaussMLE.m:64-211);
//   bamf_gauss_mle_fit_host the same for FitType 'Basic', 'Sigma', 'SigmaYX' or 'Z', raw outputs;
//   bamf_find_roi_host  findROI(Image, Sigma1, Sigma2, Boxsize, Minval) (@RJ/findROI.m:1-163);
//   bamf_find_priors_host findPriors (@RJ/findPriors.m:35-80) on the three above and the
//                       device threshold and kernel density of priors.hip;
//   bamf_find_psf_host  findPSF (@RJ/findPSF.m:30-119): the 'Sigma' fits and the
//                       shift-average of psf.hip, then the table split on the host.
// The MATLAB mex functions under csrc/mex only map mxArrays onto these.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace {
constexpr int64_t kFindRoiMaxFramesPerCall = 65535;  // bamf_find_roi's limit (one grid dimension per frame)
}

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
                 const float* sampled_psf, const int32_t* psf_size, int opt_model, uint64_t seed, float* pgrid, float* pback, bamf_chain_rec* acc_rec,
                 bamf_chain_rec* prop_rec, float* acc_em, float* prop_em, int32_t rec_cap, int32_t em_cap,
                 int64_t counts[3], double stat[16]) {
    BAMF_CHECK_ARG(mcmc && data && off_pdf && sig_pdf && back_pdf, "null argument");
    BAMF_CHECK_ARG(xsize >= 1 && ysize >= 1 && xsize <= 32 && ysize <= 32, "ROI sides must be in 1..32");
    BAMF_CHECK_ARG(n0 >= 0 && (n0 == 0 || (i0 && x0 && y0)), "E_active needs I, X and Y for N > 0");
    BAMF_CHECK_ARG(opt_model == 0 || opt_model == 1, "OptModel must be 0 (erf PSF) or 1 (sampled PSF)");
    BAMF_CHECK_ARG(opt_model == 0 || (sampled_psf && psf_size), "OptModel = 1 needs SampledPSF and PSFsize");
    BAMF_CHECK_ARG(!acc_rec == !prop_rec && !acc_rec == !acc_em && !acc_rec == !prop_em,
                   "chain outputs need all four arrays");
    BAMF_CHECK_ARG(!acc_rec || (counts && rec_cap >= 0 && em_cap >= 0), "chain outputs need counts and capacities");
    bamf_rjmcmc_plan* plan = nullptr;
    int rc = bamf_rjmcmc_plan_create(mcmc, off_pdf, off_len, sig_pdf, sig_len, back_pdf, back_len, &plan);
    if (rc != BAMF_OK) return rc;
    if (opt_model == 1 && (rc = bamf_rjmcmc_set_sampled_psf(plan, sampled_psf, psf_size)) != BAMF_OK) {
        bamf_rjmcmc_plan_destroy(plan);
        return rc;
    }
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
    if (sc) {
        // the kernels read gain[(cy + j) * mapsz + cx + i] for 0 <= i, j < sz: every box must lie on the map
        for (int64_t i = 0; i < nfits; i++) {
            const float cx = coords[2 * i], cy = coords[2 * i + 1];
            BAMF_CHECK_ARG(cx >= 0 && cy >= 0 && cx + sz <= mapsz && cy + sz <= mapsz,
                           "sCMOS fit %lld: box corner (%g, %g) of a %d x %d box is not on the %d x %d gain map",
                           (long long)i, cx, cy, sz, sz, mapsz, mapsz);
        }
    }
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
    fpc = std::min<int64_t>(fpc, kFindRoiMaxFramesPerCall);  // bamf_find_roi's per-call frame limit
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

namespace {

// psi(x) and psi'(x) for x > 0: the recurrences psi(x) = psi(x+1) - 1/x and
// psi'(x) = psi'(x+1) + 1/x^2 up to x >= 8, then the asymptotic (Bernoulli)
// series.
double digamma(double x) {
    double r = 0.0;
    while (x < 8.0) {
        r -= 1.0 / x;
        x += 1.0;
    }
    const double f = 1.0 / (x * x);
    return r + std::log(x) - 0.5 / x -
           f * (1.0 / 12 - f * (1.0 / 120 - f * (1.0 / 252 - f * (1.0 / 240 - f * (1.0 / 132)))));
}
double trigamma(double x) {
    double r = 0.0;
    while (x < 8.0) {
        r += 1.0 / (x * x);
        x += 1.0;
    }
    const double f = 1.0 / (x * x);
    return r + 1.0 / x + f / 2.0 +
           f / x * (1.0 / 6 - f * (1.0 / 30 - f * (1.0 / 42 - f * (1.0 / 30 - f * 5.0 / 66))));
}

// Maximum-likelihood Gamma(a, b) (shape, scale) of positive samples, as
// fitdist(x, 'Gamma') estimates it: log a - psi(a) = log(mean) - mean(log),
// b = mean / a; Newton from the closed-form starting point of Minka (2002).
void gamma_mle(const std::vector<double>& x, double* a_out, double* b_out) {
    double m = 0.0, ml = 0.0;
    for (double v : x) {
        m += v;
        ml += std::log(v);
    }
    m /= (double)x.size();
    ml /= (double)x.size();
    const double s = std::log(m) - ml;
    double a = (3.0 - s + std::sqrt((s - 3.0) * (s - 3.0) + 24.0 * s)) / (12.0 * s);
    for (int it = 0; it < 100; it++) {
        const double f = std::log(a) - digamma(a) - s, fp = 1.0 / a - trigamma(a);
        double an = a - f / fp;
        if (!(an > 0.0)) an = a / 2.0;
        const bool done = std::fabs(an - a) <= 1e-14 * a;
        a = an;
        if (done) break;
    }
    *a_out = a;
    *b_out = m / a;
}

double gampdf(double x, double a, double b) {
    if (x < 0.0) return 0.0;
    if (x == 0.0) return a < 1.0 ? INFINITY : a == 1.0 ? 1.0 / b : 0.0;
    return std::exp((a - 1.0) * std::log(x) - x / b - std::lgamma(a) - a * std::log(b));
}

// quantile(x, p) of MATLAB for a vector: sorted samples at the positions
// (i - 0.5)/n, linear interpolation between them, clamped to the extremes.
double matlab_quantile(std::vector<float> v, double p) {
    std::sort(v.begin(), v.end());
    const double n = (double)v.size();
    const double pos = n * p + 0.5;  // 1-based
    if (pos <= 1.0) return v.front();
    if (pos >= n) return v.back();
    const int64_t f = (int64_t)std::floor(pos);
    return (double)v[f - 1] + (pos - (double)f) * ((double)v[f] - (double)v[f - 1]);
}

// The prior grid 0:dp:top (MATLAB colon: floor(top/dp) + 1 points, with the
// colon's tolerance of a few ulps) and the kernel density of x on it.
int kde_on_grid(const std::vector<float>& x, double dp, std::vector<double>* pdf) {
    float top = 0.f;
    for (float v : x) top = std::max(top, v);
    const double end = (double)top + 1000.0 * dp;  // findPriors.m:47
    const int64_t ng = (int64_t)std::floor(end / dp * (1.0 + 4e-16)) + 1;
    BAMF_CHECK_ARG(ng > 0 && ng < (1 << 26), "prior grid of %lld points", (long long)ng);
    DevBufs b;
    float* d_x = b.upload(x.data(), x.size());
    double* d_pdf = b.alloc<double>((size_t)ng);
    if (b.err != hipSuccess) {
        set_error("findPriors gateway: %s", hipGetErrorString(b.err));
        return BAMF_E_HIP;
    }
    int rc = bamf_kde_pdf(d_x, (int64_t)x.size(), 75.0, 0.0, dp, (int)ng, d_pdf, nullptr);
    if (rc) return rc;
    pdf->resize((size_t)ng);
    b.err = hipMemcpy(pdf->data(), d_pdf, (size_t)ng * 8, hipMemcpyDeviceToHost);
    if (b.err != hipSuccess) {
        set_error("findPriors gateway: %s", hipGetErrorString(b.err));
        return BAMF_E_HIP;
    }
    return BAMF_OK;
}

}  // namespace

extern "C" {

int bamf_find_priors_host(const float* frames, int ysize, int xsize, int nframes, float psf_sigma,
                          float mean_photons, float max_xy_se, float min_pvalue, float min_photons, float p_dp,
                          float bg_prior_range, int bg_prior, double* signal_pdf, double* bg_pdf, double* offset_pdf,
                          int64_t capacity, int64_t* len, double info[8]) {
    BAMF_CHECK_ARG(frames && len && signal_pdf && bg_pdf && offset_pdf, "null argument");
    BAMF_CHECK_ARG(psf_sigma > 0.f && p_dp > 0.f && capacity >= 0, "bad argument");
    // findPriors.m:35-40: Boxsize 8, Sigma2 = 2 Sigma1, Minval from MeanPhotons
    const int box = 8;
    const double minval = (double)mean_photons / (4.0 * 3.141592653589793 * (double)psf_sigma * psf_sigma) / 4.0;
    int64_t cap = std::max<int64_t>(1024, (int64_t)ysize * xsize * nframes / 64), count = 0;
    std::vector<int32_t> xc, yc, fr;
    std::vector<float> stack;
    for (int pass = 0; pass < 2; pass++) {
        xc.assign((size_t)cap, 0);
        yc.assign((size_t)cap, 0);
        fr.assign((size_t)cap, 0);
        stack.assign((size_t)cap * box * box, 0.f);
        const int rc = bamf_find_roi_host(frames, ysize, xsize, nframes, psf_sigma, 2.0f * psf_sigma, box,
                                          (float)minval, cap, xc.data(), yc.data(), fr.data(), stack.data(),
                                          nullptr, &count);
        if (rc) return rc;
        if (count <= cap) break;
        cap = count;
    }
    const int64_t nf = count;
    BAMF_CHECK_ARG(nf > 0, "findPriors: no ROI found");
    // findPriors.m:40: gaussMLE(ROIStack, 'Basic', 'CCD', Sigma1)
    std::vector<float> prm(4 * (size_t)nf), crl(4 * (size_t)nf), ll((size_t)nf);
    int rc = bamf_gauss_mle_fit_host(stack.data(), box, nf, psf_sigma, 20, BAMF_FIT_BASIC, nullptr, prm.data(),
                                     crl.data(), ll.data(), nullptr);
    if (rc) return rc;
    // findPriors.m:41: threhsSMA on the device
    std::vector<uint8_t> keep((size_t)nf);
    {
        DevBufs b;
        const float* d_p = b.upload(prm.data(), prm.size());
        const float* d_c = b.upload(crl.data(), crl.size());
        const float* d_l = b.upload(ll.data(), ll.size());
        uint8_t* d_k = b.alloc<uint8_t>((size_t)nf);
        if (b.err == hipSuccess)
            rc = bamf_fit_keep(d_p, d_c, d_l, nf, box, max_xy_se, min_pvalue, min_photons, d_k, nullptr);
        if (rc) return rc;
        if (b.err == hipSuccess) b.err = hipMemcpy(keep.data(), d_k, (size_t)nf, hipMemcpyDeviceToHost);
        if (b.err != hipSuccess) {
            set_error("findPriors gateway: %s", hipGetErrorString(b.err));
            return BAMF_E_HIP;
        }
    }
    // isreal(SMD.X_SE) and isreal(SMD.Y_SE) test whole arrays: one negative
    // CRLB (a complex square root) rejects every fit (threhsSMA.m:9-12)
    for (int64_t i = 0; i < nf; i++)
        if (crl[i] < 0.f || crl[(size_t)nf + i] < 0.f) {
            std::fill(keep.begin(), keep.end(), 0);
            break;
        }
    std::vector<float> photon, bgv;
    for (int64_t i = 0; i < nf; i++)
        if (keep[i]) {
            photon.push_back(prm[2 * (size_t)nf + i]);
            bgv.push_back(prm[3 * (size_t)nf + i]);
        }
    BAMF_CHECK_ARG(photon.size() >= 2, "findPriors: %zu fits pass the threshold", photon.size());
    const size_t nkept = photon.size();
    // findPriors.m:43-47: trim at the 0.98 quantile, kernel density on the grid
    if (photon.size() > 200) {
        const float id = (float)matlab_quantile(photon, 0.98);
        photon.erase(std::remove_if(photon.begin(), photon.end(), [&](float v) { return v > id; }), photon.end());
    }
    std::vector<double> pdf;
    if ((rc = kde_on_grid(photon, p_dp, &pdf))) return rc;
    // findPriors.m:50-59: a second mode above the highest one cuts the tail
    int refit = 0;
    {
        const size_t n = pdf.size();
        std::vector<int> sg(n);
        for (size_t i = 0; i < n; i++) {
            const double gr = n == 1 ? 0.0
                              : i == 0 ? pdf[1] - pdf[0]
                              : i == n - 1 ? pdf[n - 1] - pdf[n - 2]
                                           : (pdf[i + 1] - pdf[i - 1]) / 2.0;
            sg[i] = (gr > 0) - (gr < 0);
        }
        std::vector<int64_t> ind;  // 1-based, where sign(gradient) changes
        for (size_t i = 1; i < n; i++)
            if (sg[i] != sg[i - 1]) ind.push_back((int64_t)i + 1);
        if (!ind.empty()) {
            double mx = -INFINITY;
            for (int64_t k : ind) mx = std::max(mx, pdf[k - 1]);
            int64_t imax = -1;  // first index of the whole density holding that value
            for (size_t i = 0; i < n && imax < 0; i++)
                if (pdf[i] == mx) imax = (int64_t)i + 1;
            std::vector<int64_t> rem;
            for (int64_t k : ind)
                if (k > imax) rem.push_back(k);
            if (rem.size() >= 2) {
                const double cut = (double)rem[0] * p_dp;
                photon.erase(std::remove_if(photon.begin(), photon.end(), [&](float v) { return (double)v > cut; }),
                             photon.end());
                if ((rc = kde_on_grid(photon, p_dp, &pdf))) return rc;
                refit = 1;
            }
        }
    }
    // findPriors.m:66-74: Gamma fit of the kept backgrounds; exponential background prior
    std::vector<double> bgd(bgv.begin(), bgv.end());
    double ga, gb;
    gamma_mle(bgd, &ga, &gb);
    double mean_ph = 0.0;
    for (float v : photon) mean_ph += (double)v;
    mean_ph /= (double)photon.size();
    const double mean_bg = mean_ph / (double)bg_prior_range;
    const int64_t ng = (int64_t)pdf.size();
    *len = ng;
    for (int64_t i = 0; i < std::min(ng, capacity); i++) {
        const double x = (double)i * p_dp;
        signal_pdf[i] = pdf[(size_t)i];
        offset_pdf[i] = gampdf(x, ga, gb);
        bg_pdf[i] = bg_prior ? std::exp(-x / mean_bg) / mean_bg : 0.0;
    }
    if (info) {
        info[0] = (double)nf;
        info[1] = (double)nkept;
        info[2] = (double)photon.size();
        info[3] = ga;
        info[4] = gb;
        info[5] = refit;
        info[6] = mean_ph;
        info[7] = minval;
    }
    return BAMF_OK;
}

int bamf_find_psf_host(const float* roi_stack, int box, int64_t n, float gain, float offset, float psf_sigma,
                       float* sampled_psf, int32_t psf_size[4], float* fit_x, float* fit_y) {
    constexpr int B = 8, M = B + 24, SUB = 64, OS = 4, PADR = 112;
    BAMF_CHECK_ARG(box == B, "findPSF's fixed 64 x 64 x 4 x 4 split needs BoxSize 8 (got %d)", box);
    BAMF_CHECK_ARG(roi_stack && sampled_psf && psf_size && n >= 1 && n <= 0x7fffffff, "need ROIs and outputs");
    BAMF_CHECK_ARG(gain != 0.0f && psf_sigma > 0.0f, "gain must be nonzero and PSFSigma > 0");
    const size_t np = (size_t)n * B * B;
    std::vector<float> r(np);
    for (size_t i = 0; i < np; i++) r[i] = (roi_stack[i] - offset) / gain;  // findPSF.m:30
    DevBufs b;
    float* d_r = b.upload(r.data(), np);
    float* d_par = b.alloc<float>(5 * (size_t)n);
    float* d_crlb = b.alloc<float>(5 * (size_t)n);
    float* d_ll = b.alloc<float>((size_t)n);
    float* d_xy = b.alloc<float>(2 * (size_t)n);
    void* d_ws = b.alloc<char>((size_t)bamf_psf_workspace_bytes(n));
    float* d_img = b.alloc<float>(M * M);
    if (b.err != hipSuccess) {
        set_error("findPSF device setup failed: %s", hipGetErrorString(b.err));
        return BAMF_E_HIP;
    }
    // findPSF.m:31: the 'Sigma' fits give the positions to align
    int rc = bamf_gauss_mle_xynbs(d_r, psf_sigma, B, 20, d_par, d_crlb, d_ll, (int)n, nullptr);
    if (rc != BAMF_OK) return rc;
    std::vector<float> par(2 * (size_t)n), xy(2 * (size_t)n);
    BAMF_HIP_TRY(hipMemcpy(par.data(), d_par, 2 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    // positions in 1-based pixel-centre coordinates (the fit's, in which pixel k
    // spans [k, k + 1), plus one half): the convention under which this table
    // reproduces the PSF through NumModel
    for (int64_t i = 0; i < n; i++) {
        xy[i] = par[n + i] + 0.5f;  // X (second parameter, gaussMLE.m:184)
        xy[n + i] = par[i] + 0.5f;  // Y
        if (fit_x) fit_x[i] = xy[i];
        if (fit_y) fit_y[i] = xy[n + i];
    }
    BAMF_HIP_TRY(hipMemcpy(d_xy, xy.data(), 2 * (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    rc = bamf_psf_shift_average(d_r, d_xy, d_xy + n, n, B, d_ws, d_img, nullptr);
    if (rc != BAMF_OK) return rc;
    std::vector<float> img(M * M);
    BAMF_HIP_TRY(hipMemcpy(img.data(), d_img, sizeof(float) * M * M, hipMemcpyDeviceToHost));
    // findPSF.m:58-68: minus the mean of the four edge means, negatives zeroed, peak 1
    double e[4] = {0, 0, 0, 0};
    for (int k = 0; k < M; k++) {
        e[0] += img[0 + M * k];        // first row
        e[1] += img[(M - 1) + M * k];  // last row
        e[2] += img[k + M * (M - 1)];  // last column
        e[3] += img[k];                // first column
    }
    const float off = (float)((e[0] + e[1] + e[2] + e[3]) / (4.0 * M));
    float peak = 0.0f;
    for (float& v : img) {
        v = v - off;
        if (v < 0.0f) v = 0.0f;
        peak = std::max(peak, v);
    }
    for (float& v : img) v = v / peak;
    // findPSF.m:73-117: padded by 112 per side (256 x 256); phase (jj, ii) takes
    // rows l = (4 - jj) % 4 + 4a and columns k = (4 - ii) % 4 + 4b, the
    // middle third (21..43) moved by one for nonzero phases, then sum 1
    auto padded = [&](int l, int k) {
        const int y = l - PADR, x = k - PADR;
        return (y >= 0 && y < M && x >= 0 && x < M) ? img[y + M * x] : 0.0f;
    };
    std::vector<float> t(SUB * SUB);
    const int lo = 21, hi = 43;  // round(64/3), round(128/3) as 0-based targets of the move
    for (int ii = 0; ii < OS; ii++)
        for (int jj = 0; jj < OS; jj++) {
            const int l0 = (OS - jj) % OS, k0 = (OS - ii) % OS;
            for (int bcol = 0; bcol < SUB; bcol++)
                for (int a = 0; a < SUB; a++) t[a + SUB * bcol] = padded(l0 + OS * a, k0 + OS * bcol);
            if (jj > 0)
                for (int a = hi; a >= lo; a--)
                    for (int bcol = 0; bcol < SUB; bcol++) t[a + SUB * bcol] = t[a - 1 + SUB * bcol];
            if (ii > 0)
                for (int bcol = hi; bcol >= lo; bcol--)
                    for (int a = 0; a < SUB; a++) t[a + SUB * bcol] = t[a + SUB * (bcol - 1)];
            double sum = 0.0;
            for (float v : t) sum += v;
            float* dst = sampled_psf + (size_t)SUB * SUB * (jj + OS * ii);
            for (int i = 0; i < SUB * SUB; i++) dst[i] = (float)(t[i] / sum);
        }
    psf_size[0] = SUB;
    psf_size[1] = SUB;
    psf_size[2] = OS;
    psf_size[3] = OS;
    return BAMF_OK;
}

}  // extern "C"
