// Single-emitter Newton MLE of theta = {x, y, N, bg} (fit type 'Basic', NV =
// 4), {x, y, N, bg, sigma} ('Sigma'), {x, y, N, bg, sigma_ii, sigma_jj}
// ('SigmaYX') or {x, y, N, bg, z} ('Z', astigmatic widths) with CRLB and
// log-likelihood.  Same inputs, outputs and iteration as kernel_MLEFit_XYNB_
// (source/cuda/cuda_gaussMLEv2/cuda_gaussMLEv2.cu:16-141), kernel_MLEFit_XYNBS_
// (:144-276), kernel_MLEFit_XYNBZ_ (:280-411) and kernel_MLEFit_XYNBSXSY_
// (:414-556).
//
// gfx950 design: one work-item per fit, the box staged through LDS so that
// global loads are contiguous, the pixel-integrated Gaussian evaluated in
// separable form (2*sz erf pairs and 2*sz exp pairs per Newton iteration
// instead of 2*sz^2 of each), and the initial-value smoothing filter of
// kernel_GaussFMaxMin2D (GPUgaussLib.cuh:230-258) applied as two 1-D passes.
// Per pixel the model, derivative and Newton sums are the reference's
// expressions evaluated in the reference's order.
#include "common.hpp"
#include "erf_fast.hpp"

namespace bamf {
namespace {

constexpr float kPi = 3.141592f;  // definitions.h:13
constexpr int kBlock = 64;

// Fit types (gaussMLE.m:93-109) and their parameter counts.
constexpr int kFtBasic = 0, kFtSigma = 1, kFtSigmaYX = 2, kFtZ = 3;
constexpr int nv_of(int ft) { return ft == kFtSigmaYX ? 6 : ft == kFtBasic ? 4 : 5; }

// sCMOS inputs (kernel_MLEFit_SCMOS*, cuda_gaussMLEv2.cu:559-1124): per fit the
// box corner (x, y) in the square map of variance/gain^2 ratios, the map's
// side, and for 'Z' the starting z of each fit.
struct ScIn {
    const float* coords;
    const float* gain;
    int mapsz;
    const float* z0;
};
// The map index of pixel (ii, jj), formed in float as the reference forms
// GRind = (Coords[1] + jj) * Mapsz + Coords[0] + ii (cuda_gaussMLEv2.cu:620).
__device__ inline float gain_at(const ScIn& sc, float cx, float cy, int ii, int jj) {
    return sc.gain[(int)((cy + jj) * sc.mapsz + cx + ii)];
}

// Astigmatism calibration of fit type 'Z' (kernel_MLEFit_XYNBZ_'s PSFSigma_y,
// Ax, Ay, Bx, By, gamma, d; PSFSigma_x is the kernels' sigma argument).
struct ZCal {
    float sy0, ax, ay, bx, by, gamma, d;
};
// GPUgaussLib.cuh:26-64: alpha(z) = 1 + (z/d)^2 + A (z/d)^3 + B (z/d)^4 and its derivatives.
__device__ inline float alpha_z(float z, float a, float b, float d) {
    const float t = z / d;
    return 1.0f + (t * t) + a * (t * t * t) + b * (t * t * t * t);
}
__device__ inline float dalphadz(float z, float a, float b, float d) {
    const float d2 = d * d;
    return (2.0f * z / d2 + 3.0f * a * (z * z) / (d2 * d) + 4.0f * b * (z * z * z) / (d2 * d2));
}
__device__ inline float d2alphadz2(float z, float a, float b, float d) {
    const float d2 = d * d;
    return (2.0f / d2 + 6.0f * a * z / (d2 * d) + 12.0f * b * (z * z) / (d2 * d2));
}

// NV x NV partial inverse: in-place LU without pivoting, diagonal of the
// inverse (same algorithm as kernel_MatInvN, MatInvLib.h:15-75).
template <int NV>
__device__ inline void mat_inv_diag(float* m, float* diag) {
    constexpr int sz = NV;
    float minv[sz * sz];
    for (int jj = 0; jj < sz; jj++) {
        for (int ii = 1; ii <= jj; ii++) {
            float t = 0.f;
            for (int kk = 0; kk <= ii - 1; kk++) t += m[ii + kk * sz] * m[kk + jj * sz];
            m[ii + jj * sz] -= t;
        }
        for (int ii = jj + 1; ii < sz; ii++) {
            float t = 0.f;
            for (int kk = 0; kk <= jj - 1; kk++) t += m[ii + kk * sz] * m[kk + jj * sz];
            m[ii + jj * sz] = (1 / m[jj + jj * sz]) * (m[ii + jj * sz] - t);
        }
    }
    for (int num = 0; num < sz; num++) {
        float yy[sz];
        yy[0] = (num == 0) ? 1.f : 0.f;
        for (int ii = 1; ii < sz; ii++) {
            float t = 0.f;
            for (int jj = 0; jj <= ii - 1; jj++) t += m[ii + jj * sz] * yy[jj];
            yy[ii] = ((ii == num) ? 1.f : 0.f) - t;
        }
        minv[sz - 1 + num * sz] = yy[sz - 1] / m[(sz - 1) + (sz - 1) * sz];
        for (int ii = sz - 2; ii >= 0; ii--) {
            float t = 0.f;
            for (int jj = ii + 1; jj < sz; jj++) t += m[ii + jj * sz] * minv[jj + num * sz];
            minv[ii + num * sz] = (1 / m[ii + ii * sz]) * (yy[ii] - t);
        }
    }
    for (int ii = 0; ii < sz; ii++) diag[ii] = minv[ii * sz + ii];
}

// Per-axis terms of the pixel-integrated Gaussian and its first and second
// derivatives (GPUgaussLib.cuh:14-24, 67-87) for pixel index i.  The
// reference's derivative is ((-N/sqrt(2pi))/sigma*(a-b))*PSF_other, so the
// axis factor d1 times the other axis's PSF reproduces it exactly.
// With SG, also the sigma-derivative factors of GPUgaussLib.cuh:90-111: the
// reference's dS = ((-N/sqrt(2pi))/S/S*(a(i-x+1) - b(i-x)))*PSF_other = s1 *
// PSF_other, and its second derivative -2/S*dS - s2 * PSF_other with
// s2 = (N/sqrt(2pi))/S^5*(a(i-x+1)^3 - b(i-x)^3).
struct AxisTerms {
    float psf, d1, d2, s1, s2;
};
template <bool SG>
__device__ inline AxisTerms axis_terms(int i, float x, float sigma, float sqrt_norm, float n,
                                       float c1, float c2) {
    AxisTerms t;
    // branch-free erf: the fits of a wavefront evaluate different ranges
    t.psf = 0.5f * (erf_nr_fast((i - x + 1.0f) * sqrt_norm) - erf_nr_fast((i - x) * sqrt_norm));
    float ta = (i + 1.0f - x) / sigma;
    float tb = (i - x) / sigma;
    float a = __expf(-0.5f * (ta * ta));
    float b = __expf(-0.5f * (tb * tb));
    t.d1 = -n / c1 / sigma * (a - b);
    t.d2 = -n / c1 / c2 * ((i + 1.0f - x) * a - (i - x) * b);
    if constexpr (SG) {
        const float p1 = i - x + 1.0f, p0 = i - x;
        t.s1 = -n / c1 / sigma / sigma * (a * p1 - b * p0);
        const float s2 = sigma * sigma;
        t.s2 = n / c1 / (s2 * s2 * sigma) * (a * (p1 * p1 * p1) - b * (p0 * p0 * p0));
    } else {
        t.s1 = t.s2 = 0.f;
    }
    return t;
}

template <int SZ, int FT, bool SC>
__global__ __launch_bounds__(kBlock) void k_mle_xynb(const float* __restrict__ data, float sigma,
                                                     int iterations, float* __restrict__ params,
                                                     float* __restrict__ crlb, float* __restrict__ ll,
                                                     int nfits, ZCal zc, ScIn sc) {
    constexpr int NV = nv_of(FT);
    constexpr int NP = SZ * SZ;
    // +1 word per box keeps the per-work-item stride odd (no LDS bank conflicts).
    __shared__ float s_box[kBlock * (NP + 1)];
    const int tx = threadIdx.x;
    const int64_t first = (int64_t)blockIdx.x * kBlock;
    const int nb = (int)min<int64_t>(kBlock, nfits - first);
    const float* src = data + first * NP;
    for (int k = tx; k < nb * NP; k += kBlock) {
        int b = k / NP, p = k - b * NP;
        s_box[b * (NP + 1) + p] = src[k];
    }
    __syncthreads();
    if (tx >= nb) return;
    const int64_t i0 = first + tx;
    const float* box = s_box + tx * (NP + 1);
    float d[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) d[p] = box[p];
    // sCMOS: the box's variance/gain^2 ratios, pixel (ii, jj) at g[SZ*jj + ii]
    float g[SC ? NP : 1];
    if constexpr (SC) {
        const float cx = sc.coords[2 * i0], cy = sc.coords[2 * i0 + 1];
#pragma unroll
        for (int jj = 0; jj < SZ; jj++)
#pragma unroll
            for (int ii = 0; ii < SZ; ii++) g[SZ * jj + ii] = gain_at(sc, cx, cy, ii, jj);
    }

    // --- initial values: centre of mass (GPUgaussLib.cuh:207-227) ---
    // SG: the axis widths vary (fitted, or functions of z); XY: one fitted width
    // per axis; ZF: widths from the fitted z through the astigmatism calibration
    constexpr bool SG = FT != kFtBasic, XY = FT == kFtSigmaYX, ZF = FT == kFtZ;
    float theta[NV];
    {
        float sx = 0.f, sy = 0.f, ss = 0.f;
#pragma unroll
        for (int ii = 0; ii < SZ; ii++)
#pragma unroll
            for (int jj = 0; jj < SZ; jj++) {
                float v = d[SZ * jj + ii];
                sx += v * ii;
                sy += v * jj;
                ss += v;
            }
        theta[0] = sx / ss;
        theta[1] = sy / ss;
    }
    // --- filtered max/min (GPUgaussLib.cuh:230-258), separable ---
    {
        const float norm = 1.0f / 2.0f / sigma / sigma;
        float g[SZ];
#pragma unroll
        for (int k = 0; k < SZ; k++) g[k] = expf(-(float)(k * k) * norm);
        float rs[SZ];  // sum_i g(|i-k|)
#pragma unroll
        for (int k = 0; k < SZ; k++) {
            float s = 0.f;
#pragma unroll
            for (int i = 0; i < SZ; i++) s += g[i > k ? i - k : k - i];
            rs[k] = s;
        }
        float t[NP];  // t[ii][ll] = sum_jj g(|ll-jj|) d[ii*SZ+jj]
#pragma unroll
        for (int ii = 0; ii < SZ; ii++)
#pragma unroll
            for (int ll = 0; ll < SZ; ll++) {
                float s = 0.f;
#pragma unroll
                for (int jj = 0; jj < SZ; jj++) s += g[ll > jj ? ll - jj : jj - ll] * d[ii * SZ + jj];
                t[ii * SZ + ll] = s;
            }
        float fmax_ = 0.0f, fmin_ = 10e10f;
#pragma unroll
        for (int kk = 0; kk < SZ; kk++)
#pragma unroll
            for (int ll = 0; ll < SZ; ll++) {
                float s = 0.f;
#pragma unroll
                for (int ii = 0; ii < SZ; ii++) s += g[ii > kk ? ii - kk : kk - ii] * t[ii * SZ + ll];
                float f = s / (rs[kk] * rs[ll]);
                fmax_ = fmaxf(fmax_, f);
                fmin_ = fminf(fmin_, f);
            }
        theta[3] = fmin_;
        if constexpr (ZF) {  // cuda_gaussMLEv2.cu:331-332
            theta[2] = fmaxf(0.0f, (fmax_ - theta[3]) * 2 * kPi * sigma * zc.sy0 * sqrtf(2.0f));
            if constexpr (SC) theta[4] = sc.z0[i0];  // z_initial (cuda_gaussMLEv2.cu:896)
            else theta[4] = 0;
        } else {
            theta[2] = fmaxf(0.0f, (fmax_ - theta[3]) * 2 * 2 * kPi * sigma * sigma);
            if constexpr (SG) theta[4] = sigma;
            if constexpr (XY) theta[5] = sigma;
        }
    }

    const float c1 = sqrtf(2.0f * kPi);
    // widths along ii (x) and jj (y): sigma, theta[4] for both, or theta[4] and theta[5]
    float sgx = sigma, snx = sqrtf(1.0f / 2.0f / sigma / sigma), c2x = sigma * sigma * sigma;
    float sgy = sgx, sny = snx, c2y = c2x;
    float dszx = 0.f, dszy = 0.f, ddszx = 0.f, ddszy = 0.f;  // 'Z': dS/dz and d2S/dz2 per axis
    auto set_widths = [&]() {
        if constexpr (ZF) {  // GPUgaussLib.cuh:155-202
            const float z = theta[4];
            const float alx = alpha_z(z - zc.gamma, zc.ax, zc.bx, zc.d);
            const float aly = alpha_z(z + zc.gamma, zc.ay, zc.by, zc.d);
            const float rax = sqrtf(alx), ray = sqrtf(aly);
            sgx = sigma * rax;
            sgy = zc.sy0 * ray;
            snx = sqrtf(1.0f / 2.0f / sgx / sgx);
            sny = sqrtf(1.0f / 2.0f / sgy / sgy);
            c2x = sgx * sgx * sgx;
            c2y = sgy * sgy * sgy;
            const float dsax = sigma / 2.0f / rax, dsay = zc.sy0 / 2.0f / ray;
            const float dax = dalphadz(z - zc.gamma, zc.ax, zc.bx, zc.d);
            const float day = dalphadz(z + zc.gamma, zc.ay, zc.by, zc.d);
            dszx = dsax * dax;
            dszy = dsay * day;
            const float d2x = -sigma / 4.0f / (alx * rax), d2y = -zc.sy0 / 4.0f / (aly * ray);
            ddszx = d2x * (dax * dax) + dsax * d2alphadz2(z - zc.gamma, zc.ax, zc.bx, zc.d);
            ddszy = d2y * (day * day) + dsay * d2alphadz2(z + zc.gamma, zc.ay, zc.by, zc.d);
            return;
        }
        if constexpr (SG) {
            sgx = theta[4];
            snx = sqrtf(1.0f / 2.0f / sgx / sgx);
            c2x = sgx * sgx * sgx;
            sgy = sgx, sny = snx, c2y = c2x;
        }
        if constexpr (XY) {
            sgy = theta[5];
            sny = sqrtf(1.0f / 2.0f / sgy / sgy);
            c2y = sgy * sgy * sgy;
        }
    };
    for (int it = 0; it < iterations; it++) {
        set_widths();
        AxisTerms ax[SZ], ay[SZ];
#pragma unroll
        for (int i = 0; i < SZ; i++) {
            ax[i] = axis_terms<SG>(i, theta[0], sgx, snx, theta[2], c1, c2x);
            ay[i] = axis_terms<SG>(i, theta[1], sgy, sny, theta[2], c1, c2y);
        }
        const float m2sx = -2.0f / sgx, m2sy = -2.0f / sgy;
        float num[NV], den[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) num[k] = den[k] = 0.f;
#pragma unroll
        for (int ii = 0; ii < SZ; ii++)
#pragma unroll
            for (int jj = 0; jj < SZ; jj++) {
                const float px = ax[ii].psf, py = ay[jj].psf;
                const float model = theta[3] + theta[2] * px * py;
                const float dat = d[SZ * jj + ii];
                const float u0 = ax[ii].d1 * py, v0 = ax[ii].d2 * py;
                const float u1 = ay[jj].d1 * px, v1 = ay[jj].d2 * px;
                const float u2 = px * py;
                float cf = 0.f, df = 0.f;
                if (model > 10e-3f) {
                    if constexpr (SC) {  // cuda_gaussMLEv2.cu:631-632
                        const float gv = g[SZ * jj + ii];
                        const float r = __builtin_amdgcn_rcpf(model + gv);
                        cf = (dat - model) * r;
                        df = (dat + gv) * r * r;
                    } else {
                        const float r = __builtin_amdgcn_rcpf(model);  // 1 ulp, for two divisions
                        cf = dat * r - 1;
                        df = dat * r * r;
                    }
                }
                cf = fminf(cf, 10e4f);
                df = fminf(df, 10e4f);
                num[0] += u0 * cf;
                den[0] += v0 * cf - (u0 * u0) * df;
                num[1] += u1 * cf;
                den[1] += v1 * cf - (u1 * u1) * df;
                num[2] += u2 * cf;
                den[2] += -(u2 * u2) * df;
                num[3] += cf;
                den[3] += -df;
                if constexpr (SG) {
                    const float dsx = ax[ii].s1 * py, dsy = ay[jj].s1 * px;
                    const float vx = m2sx * dsx - ax[ii].s2 * py, vy = m2sy * dsy - ay[jj].s2 * px;
                    if constexpr (ZF) {  // GPUgaussLib.cuh:178,198-199
                        const float u4 = dsx * dszx + dsy * dszy;
                        const float v4 = vx * (dszx * dszx) + dsx * ddszx + vy * (dszy * dszy) + dsy * ddszy;
                        num[4] += u4 * cf;
                        den[4] += v4 * cf - (u4 * u4) * df;
                    } else if constexpr (XY) {
                        num[4] += dsx * cf;
                        den[4] += vx * cf - (dsx * dsx) * df;
                        num[5] += dsy * cf;
                        den[5] += vy * cf - (dsy * dsy) * df;
                    } else {
                        const float u4 = dsx + dsy, v4 = vx + vy;
                        num[4] += u4 * cf;
                        den[4] += v4 * cf - (u4 * u4) * df;
                    }
                }
            }
        theta[0] -= fminf(fmaxf(num[0] / den[0], -1e0f), 1e0f);
        theta[1] -= fminf(fmaxf(num[1] / den[1], -1e0f), 1e0f);
        theta[2] -= fminf(fmaxf(num[2] / den[2], -theta[2]), theta[2]);
        theta[3] -= fminf(fmaxf(num[3] / den[3], -theta[3]), theta[3]);
        if constexpr (ZF) theta[4] -= fminf(fmaxf(num[4] / den[4], -1e-1f), 1e-1f);  // cuda_gaussMLEv2.cu:369
        else if constexpr (SG) theta[4] -= fminf(fmaxf(num[4] / den[4], -theta[4]), theta[4]);
        if constexpr (XY) theta[5] -= fminf(fmaxf(num[5] / den[5], -theta[5]), theta[5]);
        theta[2] = fmaxf(theta[2], 1.0f);
        theta[3] = fmaxf(theta[3], 0.01f);
        if constexpr (SG && !ZF) theta[4] = fminf(fmaxf(theta[4], 0.5f), SZ / 2.0f);  // cuda_gaussMLEv2.cu:237-238
        if constexpr (XY) theta[5] = fminf(fmaxf(theta[5], 0.5f), SZ / 2.0f);  // cuda_gaussMLEv2.cu:512-513
    }
    set_widths();

    // --- Fisher information, CRLB and log-likelihood (cuda_gaussMLEv2.cu:105-138, 243-266, 518-546) ---
    float M[NV * NV];
#pragma unroll
    for (int k = 0; k < NV * NV; k++) M[k] = 0.f;
    float div = 0.f;
    // sCMOS: the reference forms the pixel's ratio only in the Newton loop, so its
    // CRLB and log-likelihood loop uses the last pixel's (ii = jj = sz-1) for all
    [[maybe_unused]] const float gl = SC ? g[SC ? NP - 1 : 0] : 0.f;
    {
        AxisTerms ax[SZ], ay[SZ];
#pragma unroll
        for (int i = 0; i < SZ; i++) {
            ax[i] = axis_terms<SG>(i, theta[0], sgx, snx, theta[2], c1, c2x);
            ay[i] = axis_terms<SG>(i, theta[1], sgy, sny, theta[2], c1, c2y);
        }
#pragma unroll
        for (int ii = 0; ii < SZ; ii++)
#pragma unroll
            for (int jj = 0; jj < SZ; jj++) {
                const float px = ax[ii].psf, py = ay[jj].psf;
                const float model = theta[3] + theta[2] * px * py;
                const float dat = d[SZ * jj + ii];
                float du[NV];
                du[0] = ax[ii].d1 * py;
                du[1] = ay[jj].d1 * px;
                du[2] = px * py;
                du[3] = 1.0f;
                if constexpr (ZF) {
                    du[4] = ax[ii].s1 * py * dszx + ay[jj].s1 * px * dszy;
                } else if constexpr (XY) {
                    du[4] = ax[ii].s1 * py;
                    du[5] = ay[jj].s1 * px;
                } else if constexpr (SG) {
                    du[4] = ax[ii].s1 * py + ay[jj].s1 * px;
                }
                const float rm = __builtin_amdgcn_rcpf(SC ? model + gl : model);
#pragma unroll
                for (int k = 0; k < NV; k++)
#pragma unroll
                    for (int l = k; l < NV; l++) M[k * NV + l] += du[l] * du[k] * rm;
                if constexpr (SC) {  // cuda_gaussMLEv2.cu:676-679
                    if ((model + gl) > 0) {
                        if ((dat + gl) > 0) div += (dat + gl) * logf(model + gl) - model - (dat + gl) * logf(dat + gl) + dat;
                        else div += -model - gl;
                    }
                } else if (model > 0) {
                    if (dat > 0) div += dat * logf(model) - model - dat * logf(dat) + dat;
                    else div += -model;
                }
            }
#pragma unroll
        for (int k = 0; k < NV; k++)
#pragma unroll
            for (int l = k + 1; l < NV; l++) M[l * NV + k] = M[k * NV + l];
    }
    float diag[NV];
    mat_inv_diag<NV>(M, diag);
    const int64_t i = first + tx;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        params[(int64_t)nfits * k + i] = theta[k];
        crlb[(int64_t)nfits * k + i] = diag[k];
    }
    ll[i] = div;
}

// Any box size: same algorithm reading the box from global memory.
template <int FT, bool SC>
__global__ __launch_bounds__(kBlock) void k_mle_xynb_any(const float* __restrict__ data, float sigma,
                                                         int sz, int iterations, float* __restrict__ params,
                                                         float* __restrict__ crlb, float* __restrict__ ll,
                                                         int nfits, ZCal zc, ScIn sc) {
    constexpr int NV = nv_of(FT);
    // SG: the axis widths vary (fitted, or functions of z); XY: one fitted width
    // per axis; ZF: widths from the fitted z through the astigmatism calibration
    constexpr bool SG = FT != kFtBasic, XY = FT == kFtSigmaYX, ZF = FT == kFtZ;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nfits) return;
    const int64_t i0 = i;
    const float cx = SC ? sc.coords[2 * i] : 0.f, cy = SC ? sc.coords[2 * i + 1] : 0.f;
    const float* d = data + i * sz * sz;
    float theta[NV];
    {
        float sx = 0.f, sy = 0.f, ss = 0.f;
        for (int ii = 0; ii < sz; ii++)
            for (int jj = 0; jj < sz; jj++) {
                float v = d[sz * jj + ii];
                sx += v * ii;
                sy += v * jj;
                ss += v;
            }
        theta[0] = sx / ss;
        theta[1] = sy / ss;
    }
    {
        const float norm = 1.0f / 2.0f / sigma / sigma;
        float fmax_ = 0.0f, fmin_ = 10e10f;
        for (int kk = 0; kk < sz; kk++)
            for (int ll2 = 0; ll2 < sz; ll2++) {
                float s = 0.f, w = 0.f;
                for (int ii = 0; ii < sz; ii++) {
                    float gi = expf(-(float)((ii - kk) * (ii - kk)) * norm);
                    for (int jj = 0; jj < sz; jj++) {
                        float gg = gi * expf(-(float)((ll2 - jj) * (ll2 - jj)) * norm);
                        s += gg * d[ii * sz + jj];
                        w += gg;
                    }
                }
                float f = s / w;
                fmax_ = fmaxf(fmax_, f);
                fmin_ = fminf(fmin_, f);
            }
        theta[3] = fmin_;
        if constexpr (ZF) {  // cuda_gaussMLEv2.cu:331-332
            theta[2] = fmaxf(0.0f, (fmax_ - theta[3]) * 2 * kPi * sigma * zc.sy0 * sqrtf(2.0f));
            if constexpr (SC) theta[4] = sc.z0[i0];  // z_initial (cuda_gaussMLEv2.cu:896)
            else theta[4] = 0;
        } else {
            theta[2] = fmaxf(0.0f, (fmax_ - theta[3]) * 2 * 2 * kPi * sigma * sigma);
            if constexpr (SG) theta[4] = sigma;
            if constexpr (XY) theta[5] = sigma;
        }
    }
    const float c1 = sqrtf(2.0f * kPi);
    // widths along ii (x) and jj (y): sigma, theta[4] for both, or theta[4] and theta[5]
    float sgx = sigma, snx = sqrtf(1.0f / 2.0f / sigma / sigma), c2x = sigma * sigma * sigma;
    float sgy = sgx, sny = snx, c2y = c2x;
    float dszx = 0.f, dszy = 0.f, ddszx = 0.f, ddszy = 0.f;  // 'Z': dS/dz and d2S/dz2 per axis
    auto set_widths = [&]() {
        if constexpr (ZF) {  // GPUgaussLib.cuh:155-202
            const float z = theta[4];
            const float alx = alpha_z(z - zc.gamma, zc.ax, zc.bx, zc.d);
            const float aly = alpha_z(z + zc.gamma, zc.ay, zc.by, zc.d);
            const float rax = sqrtf(alx), ray = sqrtf(aly);
            sgx = sigma * rax;
            sgy = zc.sy0 * ray;
            snx = sqrtf(1.0f / 2.0f / sgx / sgx);
            sny = sqrtf(1.0f / 2.0f / sgy / sgy);
            c2x = sgx * sgx * sgx;
            c2y = sgy * sgy * sgy;
            const float dsax = sigma / 2.0f / rax, dsay = zc.sy0 / 2.0f / ray;
            const float dax = dalphadz(z - zc.gamma, zc.ax, zc.bx, zc.d);
            const float day = dalphadz(z + zc.gamma, zc.ay, zc.by, zc.d);
            dszx = dsax * dax;
            dszy = dsay * day;
            const float d2x = -sigma / 4.0f / (alx * rax), d2y = -zc.sy0 / 4.0f / (aly * ray);
            ddszx = d2x * (dax * dax) + dsax * d2alphadz2(z - zc.gamma, zc.ax, zc.bx, zc.d);
            ddszy = d2y * (day * day) + dsay * d2alphadz2(z + zc.gamma, zc.ay, zc.by, zc.d);
            return;
        }
        if constexpr (SG) {
            sgx = theta[4];
            snx = sqrtf(1.0f / 2.0f / sgx / sgx);
            c2x = sgx * sgx * sgx;
            sgy = sgx, sny = snx, c2y = c2x;
        }
        if constexpr (XY) {
            sgy = theta[5];
            sny = sqrtf(1.0f / 2.0f / sgy / sgy);
            c2y = sgy * sgy * sgy;
        }
    };
    for (int it = 0; it < iterations; it++) {
        set_widths();
        const float m2sx = -2.0f / sgx, m2sy = -2.0f / sgy;
        float num[NV], den[NV];
        for (int k = 0; k < NV; k++) num[k] = den[k] = 0.f;
        for (int ii = 0; ii < sz; ii++) {
            AxisTerms ax = axis_terms<SG>(ii, theta[0], sgx, snx, theta[2], c1, c2x);
            for (int jj = 0; jj < sz; jj++) {
                AxisTerms ay = axis_terms<SG>(jj, theta[1], sgy, sny, theta[2], c1, c2y);
                const float px = ax.psf, py = ay.psf;
                const float model = theta[3] + theta[2] * px * py;
                const float dat = d[sz * jj + ii];
                const float u0 = ax.d1 * py, v0 = ax.d2 * py, u1 = ay.d1 * px, v1 = ay.d2 * px, u2 = px * py;
                float cf = 0.f, df = 0.f;
                if (model > 10e-3f) {
                    if constexpr (SC) {  // cuda_gaussMLEv2.cu:631-632
                        const float gv = gain_at(sc, cx, cy, ii, jj);
                        const float r = __builtin_amdgcn_rcpf(model + gv);
                        cf = (dat - model) * r;
                        df = (dat + gv) * r * r;
                    } else {
                        const float r = __builtin_amdgcn_rcpf(model);  // 1 ulp, for two divisions
                        cf = dat * r - 1;
                        df = dat * r * r;
                    }
                }
                cf = fminf(cf, 10e4f);
                df = fminf(df, 10e4f);
                num[0] += u0 * cf; den[0] += v0 * cf - (u0 * u0) * df;
                num[1] += u1 * cf; den[1] += v1 * cf - (u1 * u1) * df;
                num[2] += u2 * cf; den[2] += -(u2 * u2) * df;
                num[3] += cf;      den[3] += -df;
                if constexpr (SG) {
                    const float dsx = ax.s1 * py, dsy = ay.s1 * px;
                    const float vx = m2sx * dsx - ax.s2 * py, vy = m2sy * dsy - ay.s2 * px;
                    if constexpr (ZF) {  // GPUgaussLib.cuh:178,198-199
                        const float u4 = dsx * dszx + dsy * dszy;
                        const float v4 = vx * (dszx * dszx) + dsx * ddszx + vy * (dszy * dszy) + dsy * ddszy;
                        num[4] += u4 * cf;
                        den[4] += v4 * cf - (u4 * u4) * df;
                    } else if constexpr (XY) {
                        num[4] += dsx * cf;
                        den[4] += vx * cf - (dsx * dsx) * df;
                        num[5] += dsy * cf;
                        den[5] += vy * cf - (dsy * dsy) * df;
                    } else {
                        const float u4 = dsx + dsy, v4 = vx + vy;
                        num[4] += u4 * cf;
                        den[4] += v4 * cf - (u4 * u4) * df;
                    }
                }
            }
        }
        theta[0] -= fminf(fmaxf(num[0] / den[0], -1e0f), 1e0f);
        theta[1] -= fminf(fmaxf(num[1] / den[1], -1e0f), 1e0f);
        theta[2] -= fminf(fmaxf(num[2] / den[2], -theta[2]), theta[2]);
        theta[3] -= fminf(fmaxf(num[3] / den[3], -theta[3]), theta[3]);
        if constexpr (ZF) theta[4] -= fminf(fmaxf(num[4] / den[4], -1e-1f), 1e-1f);  // cuda_gaussMLEv2.cu:369
        else if constexpr (SG) theta[4] -= fminf(fmaxf(num[4] / den[4], -theta[4]), theta[4]);
        if constexpr (XY) theta[5] -= fminf(fmaxf(num[5] / den[5], -theta[5]), theta[5]);
        theta[2] = fmaxf(theta[2], 1.0f);
        theta[3] = fmaxf(theta[3], 0.01f);
        if constexpr (SG && !ZF) theta[4] = fminf(fmaxf(theta[4], 0.5f), sz / 2.0f);
        if constexpr (XY) theta[5] = fminf(fmaxf(theta[5], 0.5f), sz / 2.0f);
    }
    set_widths();
    float M[NV * NV];
    for (int k = 0; k < NV * NV; k++) M[k] = 0.f;
    float div = 0.f;
    // sCMOS: the reference forms the pixel's ratio only in the Newton loop, so its
    // CRLB and log-likelihood loop uses the last pixel's (ii = jj = sz-1) for all
    [[maybe_unused]] const float gl = SC ? gain_at(sc, cx, cy, sz - 1, sz - 1) : 0.f;
    for (int ii = 0; ii < sz; ii++) {
        AxisTerms ax = axis_terms<SG>(ii, theta[0], sgx, snx, theta[2], c1, c2x);
        for (int jj = 0; jj < sz; jj++) {
            AxisTerms ay = axis_terms<SG>(jj, theta[1], sgy, sny, theta[2], c1, c2y);
            const float px = ax.psf, py = ay.psf;
            const float model = theta[3] + theta[2] * px * py;
            const float dat = d[sz * jj + ii];
            float du[NV];
            du[0] = ax.d1 * py;
            du[1] = ay.d1 * px;
            du[2] = px * py;
            du[3] = 1.0f;
            if constexpr (ZF) {
                du[4] = ax.s1 * py * dszx + ay.s1 * px * dszy;
            } else if constexpr (XY) {
                du[4] = ax.s1 * py;
                du[5] = ay.s1 * px;
            } else if constexpr (SG) {
                du[4] = ax.s1 * py + ay.s1 * px;
            }
            for (int k = 0; k < NV; k++)
                for (int l = k; l < NV; l++) M[k * NV + l] += du[l] * du[k] / (SC ? model + gl : model);
            if constexpr (SC) {  // cuda_gaussMLEv2.cu:676-679
                if ((model + gl) > 0) {
                    if ((dat + gl) > 0) div += (dat + gl) * logf(model + gl) - model - (dat + gl) * logf(dat + gl) + dat;
                    else div += -model - gl;
                }
            } else if (model > 0) {
                if (dat > 0) div += dat * logf(model) - model - dat * logf(dat) + dat;
                else div += -model;
            }
        }
    }
    for (int k = 0; k < NV; k++)
        for (int l = k + 1; l < NV; l++) M[l * NV + k] = M[k * NV + l];
    float diag[NV];
    mat_inv_diag<NV>(M, diag);
    for (int k = 0; k < NV; k++) {
        params[(int64_t)nfits * k + i] = theta[k];
        crlb[(int64_t)nfits * k + i] = diag[k];
    }
    ll[i] = div;
}

template <int FT, bool SC = false>
int launch_mle(const float* d_data, float psf_sigma, int sz, int iterations, float* d_params, float* d_crlb,
               float* d_ll, int nfits, void* stream, ZCal zc = {}, ScIn sc = {}) {
    BAMF_CHECK_ARG(nfits >= 0, "nfits must be >= 0 (got %d)", nfits);
    BAMF_CHECK_ARG(sz >= 1 && sz <= 64, "sz must be in [1, 64] (got %d)", sz);
    BAMF_CHECK_ARG(iterations >= 0, "iterations must be >= 0");
    BAMF_CHECK_ARG(psf_sigma > 0.f, "psf_sigma must be > 0");
    if (nfits == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_data && d_params && d_crlb && d_ll, "null device pointer");
    if (SC) {
        BAMF_CHECK_ARG(sc.coords && sc.gain && sc.mapsz >= sz, "sCMOS fits need box corners and a gain map of side >= sz");
        BAMF_CHECK_ARG(FT != kFtZ || sc.z0, "the sCMOS 'Z' fit needs its starting z");
    }
    dim3 grid((unsigned)ceil_div(nfits, kBlock)), block(kBlock);
    hipStream_t s = as_stream(stream);
    switch (sz) {
    case 5: hipLaunchKernelGGL((k_mle_xynb<5, FT, SC>), grid, block, 0, s, d_data, psf_sigma, iterations, d_params, d_crlb, d_ll, nfits, zc, sc); break;
    case 6: hipLaunchKernelGGL((k_mle_xynb<6, FT, SC>), grid, block, 0, s, d_data, psf_sigma, iterations, d_params, d_crlb, d_ll, nfits, zc, sc); break;
    case 7: hipLaunchKernelGGL((k_mle_xynb<7, FT, SC>), grid, block, 0, s, d_data, psf_sigma, iterations, d_params, d_crlb, d_ll, nfits, zc, sc); break;
    case 8: hipLaunchKernelGGL((k_mle_xynb<8, FT, SC>), grid, block, 0, s, d_data, psf_sigma, iterations, d_params, d_crlb, d_ll, nfits, zc, sc); break;
    case 9: hipLaunchKernelGGL((k_mle_xynb<9, FT, SC>), grid, block, 0, s, d_data, psf_sigma, iterations, d_params, d_crlb, d_ll, nfits, zc, sc); break;
    default:
        hipLaunchKernelGGL((k_mle_xynb_any<FT, SC>), grid, block, 0, s, d_data, psf_sigma, sz, iterations, d_params, d_crlb, d_ll, nfits, zc, sc);
    }
    BAMF_LAUNCH_CHECK("k_mle_xynb");
    return BAMF_OK;
}

}  // namespace
}  // namespace bamf

extern "C" int bamf_gauss_mle_xynb(const float* d_data, float psf_sigma, int sz, int iterations,
                                   float* d_params, float* d_crlb, float* d_ll, int nfits, void* stream) {
    return bamf::launch_mle<bamf::kFtBasic>(d_data, psf_sigma, sz, iterations, d_params, d_crlb, d_ll, nfits, stream);
}

extern "C" int bamf_gauss_mle_xynbs(const float* d_data, float psf_sigma, int sz, int iterations,
                                    float* d_params, float* d_crlb, float* d_ll, int nfits, void* stream) {
    return bamf::launch_mle<bamf::kFtSigma>(d_data, psf_sigma, sz, iterations, d_params, d_crlb, d_ll, nfits, stream);
}

extern "C" int bamf_gauss_mle_xynbsxsy(const float* d_data, float psf_sigma, int sz, int iterations,
                                       float* d_params, float* d_crlb, float* d_ll, int nfits, void* stream) {
    return bamf::launch_mle<bamf::kFtSigmaYX>(d_data, psf_sigma, sz, iterations, d_params, d_crlb, d_ll, nfits,
                                               stream);
}

extern "C" int bamf_gauss_mle_xynbz(const float* d_data, float psf_sigma_x, float ax, float ay, float bx, float by,
                                    float gamma, float d, float psf_sigma_y, int sz, int iterations, float* d_params,
                                    float* d_crlb, float* d_ll, int nfits, void* stream) {
    BAMF_CHECK_ARG(psf_sigma_y > 0.f && d != 0.f, "psf_sigma_y must be > 0 and d nonzero");
    return bamf::launch_mle<bamf::kFtZ>(d_data, psf_sigma_x, sz, iterations, d_params, d_crlb, d_ll, nfits, stream,
                                        bamf::ZCal{psf_sigma_y, ax, ay, bx, by, gamma, d});
}

// The sCMOS kernels (cuda_gaussMLEv2.cu:559-1124), same argument lists.
extern "C" int bamf_gauss_mle_scmos_xynb(const float* d_data, const float* d_coords, const float* d_gain_ratio,
                                         float psf_sigma, int sz, int mapsz, int iterations, float* d_params,
                                         float* d_crlb, float* d_ll, int nfits, void* stream) {
    return bamf::launch_mle<bamf::kFtBasic, true>(d_data, psf_sigma, sz, iterations, d_params, d_crlb, d_ll, nfits,
                                                  stream, {}, bamf::ScIn{d_coords, d_gain_ratio, mapsz, nullptr});
}

extern "C" int bamf_gauss_mle_scmos_xynbs(const float* d_data, const float* d_coords, const float* d_gain_ratio,
                                          float psf_sigma, int sz, int mapsz, int iterations, float* d_params,
                                          float* d_crlb, float* d_ll, int nfits, void* stream) {
    return bamf::launch_mle<bamf::kFtSigma, true>(d_data, psf_sigma, sz, iterations, d_params, d_crlb, d_ll, nfits,
                                                  stream, {}, bamf::ScIn{d_coords, d_gain_ratio, mapsz, nullptr});
}

extern "C" int bamf_gauss_mle_scmos_xynbsxsy(const float* d_data, const float* d_coords, const float* d_gain_ratio,
                                             float psf_sigma, int sz, int mapsz, int iterations, float* d_params,
                                             float* d_crlb, float* d_ll, int nfits, void* stream) {
    return bamf::launch_mle<bamf::kFtSigmaYX, true>(d_data, psf_sigma, sz, iterations, d_params, d_crlb, d_ll,
                                                    nfits, stream, {},
                                                    bamf::ScIn{d_coords, d_gain_ratio, mapsz, nullptr});
}

extern "C" int bamf_gauss_mle_scmos_xynbz(const float* d_data, const float* d_coords, const float* d_gain_ratio,
                                          const float* d_x0, float psf_sigma_x, float ax, float ay, float bx, float by,
                                          float gamma, float d, float psf_sigma_y, int sz, int mapsz, int iterations,
                                          float* d_params, float* d_crlb, float* d_ll, int nfits, void* stream) {
    BAMF_CHECK_ARG(psf_sigma_y > 0.f && d != 0.f, "psf_sigma_y must be > 0 and d nonzero");
    return bamf::launch_mle<bamf::kFtZ, true>(d_data, psf_sigma_x, sz, iterations, d_params, d_crlb, d_ll, nfits,
                                              stream, bamf::ZCal{psf_sigma_y, ax, ay, bx, by, gamma, d},
                                              bamf::ScIn{d_coords, d_gain_ratio, mapsz, d_x0});
}
