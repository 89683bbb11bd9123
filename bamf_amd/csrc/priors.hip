// Prior construction of @RJ/findPriors.m:35-80 on the device: the fit
// threshold of @RJ/threhsSMA.m (standard errors, chi-square p-value, photons,
// position and background), and the normal-kernel density of fitdist(...,
// 'Kernel', 'BandWidth', h) evaluated on the prior grid (pdf(PdEm, Xarray)).
//
// gfx950 design: one work-item per fit for the threshold (the p-value is the
// regularised incomplete gamma function in double precision: its power series
// below a + 1, else the Lentz continued fraction of the upper function, the
// textbook pair of expansions); the density is an n x ngrid sum of Gaussians,
// one work-item per grid point, the samples staged through LDS in tiles so a
// work-group reads each sample once per tile, accumulated in double (the
// reference evaluates it in double).
#include <cmath>

#include "common.hpp"

namespace bamf {
namespace {

// Regularised lower incomplete gamma P(a, x) for a > 0, x >= 0.
__device__ double gamma_p(double a, double x) {
    if (!(x > 0.0)) return 0.0;
    const double lg = lgamma(a);
    if (x < a + 1.0) {  // sum_n x^n / (a (a+1) ... (a+n)), times x^a e^-x / Gamma(a)
        double ap = a, term = 1.0 / a, sum = term;
        for (int n = 0; n < 1000; n++) {
            ap += 1.0;
            term *= x / ap;
            sum += term;
            if (fabs(term) < fabs(sum) * 1e-16) break;
        }
        return sum * exp(-x + a * log(x) - lg);
    }
    // Q(a, x) = e^-x x^a / Gamma(a) * 1 / (x + 1 - a - 1(1-a)/(x + 3 - a - ...)), modified Lentz
    const double tiny = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i < 1000; i++) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < 1e-16) break;
    }
    return 1.0 - exp(-x + a * log(x) - lg) * h;
}

// threhsSMA.m:1-17 for fit type 'Basic' (k = BoxSize^2 - 4 degrees of freedom).
// params/crlb are parameter-major [4][n] (Y, X, Photons, BG as gaussMLE.m maps
// them); SE = sqrt(CRLB).
__global__ void k_fit_keep(const float* __restrict__ params, const float* __restrict__ crlb,
                           const float* __restrict__ ll, int64_t n, int box, float max_xy_se, float min_pvalue,
                           float min_photons, uint8_t* __restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float y = params[i], x = params[n + i], ph = params[2 * n + i], bg = params[3 * n + i];
    const float y_se = sqrtf(crlb[i]), x_se = sqrtf(crlb[n + i]);
    const double k = (double)box * box - 4.0;
    const double x2 = -2.0 * (double)ll[i];
    const double pv = 1.0 - gamma_p(k / 2.0, x2 / 2.0);  // 1 - gammainc(X2/2, k/2)
    const bool ok = x_se < max_xy_se && y_se < max_xy_se && pv > (double)min_pvalue && ph > min_photons &&
                    x > 0.f && y > 0.f && x < (float)box && y < (float)box && bg > 0.f;
    keep[i] = ok ? 1 : 0;
}

constexpr int kKdeThreads = 256, kKdeTile = 1024;

// pdf[g] = 1/(n h) sum_i phi((x0 + g dx - s_i) / h), phi the standard normal density.
__global__ __launch_bounds__(kKdeThreads) void k_kde(const float* __restrict__ samples, int64_t n, double h,
                                                     double x0, double dx, int ngrid, double* __restrict__ pdf) {
    __shared__ double tile[kKdeTile];
    const int g = blockIdx.x * kKdeThreads + threadIdx.x;
    const double xg = x0 + (double)g * dx;
    const double rh = 1.0 / h;
    double acc = 0.0;
    for (int64_t t0 = 0; t0 < n; t0 += kKdeTile) {
        const int nt = (int)min<int64_t>(kKdeTile, n - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < nt; e += kKdeThreads) tile[e] = (double)samples[t0 + e];
        __syncthreads();
        if (g < ngrid)
            for (int e = 0; e < nt; e++) {
                const double u = (xg - tile[e]) * rh;
                acc += exp(-0.5 * u * u);
            }
    }
    if (g < ngrid) pdf[g] = acc / ((double)n * h * 2.5066282746310002);  // sqrt(2 pi)
}

}  // namespace
}  // namespace bamf

using namespace bamf;

extern "C" {

int bamf_fit_keep(const float* d_params, const float* d_crlb, const float* d_ll, int64_t nfits, int box,
                  float max_xy_se, float min_pvalue, float min_photons, uint8_t* d_keep, void* stream) {
    BAMF_CHECK_ARG(nfits >= 0 && box >= 3, "bad sizes");
    if (nfits == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_params && d_crlb && d_ll && d_keep, "null device pointer");
    hipLaunchKernelGGL(k_fit_keep, dim3((unsigned)ceil_div(nfits, 256)), dim3(256), 0, as_stream(stream), d_params,
                       d_crlb, d_ll, nfits, box, max_xy_se, min_pvalue, min_photons, d_keep);
    BAMF_LAUNCH_CHECK("k_fit_keep");
    return BAMF_OK;
}

int bamf_kde_pdf(const float* d_samples, int64_t n, double bandwidth, double x0, double dx, int ngrid,
                 double* d_pdf, void* stream) {
    BAMF_CHECK_ARG(n >= 1 && ngrid >= 0 && bandwidth > 0.0, "need samples, ngrid >= 0 and bandwidth > 0");
    if (ngrid == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_samples && d_pdf, "null device pointer");
    hipLaunchKernelGGL(k_kde, dim3((unsigned)ceil_div(ngrid, kKdeThreads)), dim3(kKdeThreads), 0, as_stream(stream),
                       d_samples, n, bandwidth, x0, dx, ngrid, d_pdf);
    BAMF_LAUNCH_CHECK("k_kde");
    return BAMF_OK;
}

}  // extern "C"
