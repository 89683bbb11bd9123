/*
 * Oracle: single-emitter Newton MLE of (x, y, I, bg) and of (x, y, I, bg,
 * sigma) with CRLB and log-likelihood.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates kernel_MLEFit_XYNB_ (source/cuda/cuda_gaussMLEv2/cuda_gaussMLEv2.cu:16-141),
 * kernel_MLEFit_XYNBS_ (cuda_gaussMLEv2.cu:144-276), kernel_MLEFit_XYNBSXSY_
 * (cuda_gaussMLEv2.cu:414-556: one width per axis), kernel_MLEFit_XYNBZ_
 * (cuda_gaussMLEv2.cu:280-411: astigmatic z, with kernel_alpha / dalphadz /
 * d2alphadz2 and kernel_DerivativeIntGauss2Dz, GPUgaussLib.cuh:26-64, 140-205)
 * with their sigma derivatives
 * kernel_DerivativeIntGauss1DSigma / 2DSigma (GPUgaussLib.cuh:90-137)
 * and its helpers kernel_IntGauss1D / kernel_DerivativeIntGauss1D
 * (GPUgaussLib.cuh:14-24, 67-87), kernel_CenterofMass2D (GPUgaussLib.cuh:207-227),
 * kernel_GaussFMaxMin2D (GPUgaussLib.cuh:230-258) and kernel_MatInvN
 * (MatInvLib.h:15-75), one fit at a time, in the reference's loop order.
 * pi is 3.141592f (definitions.h:13).
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

#define NV 4
static const float kPi = 3.141592f;

/* GPUgaussLib.cuh:14-24 */
float orc_int_gauss_1d(int ii, float x, float sigma) {
    const float norm = 1.0f / 2.0f / sigma / sigma;
    return 1.0f / 2.0f * (erff((ii - x + 1.0f) * sqrtf(norm)) - erff((ii - x) * sqrtf(norm)));
}

/* GPUgaussLib.cuh:67-87.  pow(v, 2.0f) is written v*v (both correctly rounded). */
static void deriv_int_gauss_1d(int ii, float x, float sigma, float n, float psfy,
                               float* dudt, float* d2udt2) {
    float ta = (ii + 1.0f - x) / sigma;
    float tb = (ii - x) / sigma;
    float a = expf(-1.0f / 2.0f * (ta * ta));
    float b = expf(-1.0f / 2.0f * (tb * tb));
    *dudt = -n / sqrtf(2.0f * kPi) / sigma * (a - b) * psfy;
    if (d2udt2)
        *d2udt2 = -n / sqrtf(2.0f * kPi) / powf(sigma, 3.0f) * ((ii + 1.0f - x) * a - (ii - x) * b) * psfy;
}

/* GPUgaussLib.cuh:207-227 */
void orc_center_of_mass_2d(int sz, const float* data, float* x, float* y) {
    float tmpx = 0.0f, tmpy = 0.0f, tmpsum = 0.0f;
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++) {
            tmpx += data[sz * jj + ii] * ii;
            tmpy += data[sz * jj + ii] * jj;
            tmpsum += data[sz * jj + ii];
        }
    *x = tmpx / tmpsum;
    *y = tmpy / tmpsum;
}

/* GPUgaussLib.cuh:230-258: brute-force normalised Gaussian smoothing; max and min. */
void orc_gauss_fmaxmin_2d(int sz, float sigma, const float* data, float* maxn, float* minbg) {
    float filteredpixel, sum;
    *maxn = 0.0f;
    *minbg = 10e10f;
    const float norm = 1.0f / 2.0f / sigma / sigma;
    for (int kk = 0; kk < sz; kk++)
        for (int ll = 0; ll < sz; ll++) {
            filteredpixel = 0.0f;
            sum = 0.0f;
            for (int ii = 0; ii < sz; ii++)
                for (int jj = 0; jj < sz; jj++) {
                    float di = (float)(ii - kk), dj = (float)(ll - jj);
                    filteredpixel += expf(-(di * di) * norm) * expf(-(dj * dj) * norm) * data[ii * sz + jj];
                    sum += expf(-(di * di) * norm) * expf(-(dj * dj) * norm);
                }
            filteredpixel /= sum;
            *maxn = fmaxf(*maxn, filteredpixel);
            *minbg = fminf(*minbg, filteredpixel);
        }
}

/* MatInvLib.h:15-75: in-place LU without pivoting, column solves, diag of inverse. */
void orc_mat_inv_n(float* m, float* minv, float* diag, int sz) {
    int ii, jj, kk, num, b;
    float tmp1 = 0;
    float yy[25];
    for (jj = 0; jj < sz; jj++) {
        for (ii = 0; ii <= jj; ii++)
            if (ii > 0) {
                for (kk = 0; kk <= ii - 1; kk++) tmp1 += m[ii + kk * sz] * m[kk + jj * sz];
                m[ii + jj * sz] -= tmp1;
                tmp1 = 0;
            }
        for (ii = jj + 1; ii < sz; ii++)
            if (jj > 0) {
                for (kk = 0; kk <= jj - 1; kk++) tmp1 += m[ii + kk * sz] * m[kk + jj * sz];
                m[ii + jj * sz] = (1 / m[jj + jj * sz]) * (m[ii + jj * sz] - tmp1);
                tmp1 = 0;
            } else {
                m[ii + jj * sz] = (1 / m[jj + jj * sz]) * m[ii + jj * sz];
            }
    }
    tmp1 = 0;
    for (num = 0; num < sz; num++) {
        yy[0] = (num == 0) ? 1.0f : 0.0f;
        for (ii = 1; ii < sz; ii++) {
            b = (ii == num) ? 1 : 0;
            for (jj = 0; jj <= ii - 1; jj++) tmp1 += m[ii + jj * sz] * yy[jj];
            yy[ii] = b - tmp1;
            tmp1 = 0;
        }
        minv[sz - 1 + num * sz] = yy[sz - 1] / m[(sz - 1) + (sz - 1) * sz];
        for (ii = sz - 2; ii >= 0; ii--) {
            for (jj = ii + 1; jj < sz; jj++) tmp1 += m[ii + jj * sz] * minv[jj + num * sz];
            minv[ii + num * sz] = (1 / m[ii + ii * sz]) * (yy[ii] - tmp1);
            tmp1 = 0;
        }
    }
    if (diag)
        for (ii = 0; ii < sz; ii++) diag[ii] = minv[ii * sz + ii];
}

/* cuda_gaussMLEv2.cu:16-141, one fit per loop iteration instead of per thread. */
static void mle_one(const float* s_data, const float* gr, float sigma, int sz, int iterations,
                    float* theta, float* diagout, float* llout) {
    float m[NV * NV], diag[NV], minv[NV * NV];
    float dudt[NV], d2udt2[NV], num[NV], den[NV];
    const float maxjump[NV] = {1e0f, 1e0f, 1e2f, 2e0f};
    float nmax, model, cf, df, data, psfx, psfy, div;
    /* sCMOS: the pixel's variance/gain^2 (gr != NULL).  Like the reference, the
     * Fisher and log-likelihood loop below reuses the value of the last pixel of
     * the last Newton iteration (cuda_gaussMLEv2.cu:620-621 set it only there). */
    float gainr = 0.0f;
    (void)maxjump;
    memset(m, 0, sizeof(m));
    memset(minv, 0, sizeof(minv));
    orc_center_of_mass_2d(sz, s_data, &theta[0], &theta[1]);
    orc_gauss_fmaxmin_2d(sz, sigma, s_data, &nmax, &theta[3]);
    theta[2] = fmaxf(0.0f, (nmax - theta[3]) * 2 * 2 * kPi * sigma * sigma);

    for (int kk = 0; kk < iterations; kk++) {
        memset(num, 0, sizeof(num));
        memset(den, 0, sizeof(den));
        for (int ii = 0; ii < sz; ii++)
            for (int jj = 0; jj < sz; jj++) {
                psfx = orc_int_gauss_1d(ii, theta[0], sigma);
                psfy = orc_int_gauss_1d(jj, theta[1], sigma);
                model = theta[3] + theta[2] * psfx * psfy;
                data = s_data[sz * jj + ii];
                deriv_int_gauss_1d(ii, theta[0], sigma, theta[2], psfy, &dudt[0], &d2udt2[0]);
                deriv_int_gauss_1d(jj, theta[1], sigma, theta[2], psfx, &dudt[1], &d2udt2[1]);
                dudt[2] = psfx * psfy;
                d2udt2[2] = 0.0f;
                dudt[3] = 1.0f;
                d2udt2[3] = 0.0f;
                cf = 0.0f;
                df = 0.0f;
                if (gr) {
                    /* cuda_gaussMLEv2.cu:620-632 */
                    gainr = gr[sz * jj + ii];
                    if (model > 10e-3f) cf = (data - model) / (model + gainr);
                    if (model > 10e-3f) df = (data + gainr) / ((model + gainr) * (model + gainr));
                } else {
                    if (model > 10e-3f) cf = data / model - 1;
                    if (model > 10e-3f) df = data / (model * model);
                }
                cf = fminf(cf, 10e4f);
                df = fminf(df, 10e4f);
                for (int ll = 0; ll < NV; ll++) {
                    num[ll] += dudt[ll] * cf;
                    den[ll] += d2udt2[ll] * cf - (dudt[ll] * dudt[ll]) * df;
                }
            }
        /* cuda_gaussMLEv2.cu:93-100 */
        theta[0] -= fminf(fmaxf(num[0] / den[0], -1e0f), 1e0f);
        theta[1] -= fminf(fmaxf(num[1] / den[1], -1e0f), 1e0f);
        theta[2] -= fminf(fmaxf(num[2] / den[2], -theta[2]), theta[2]);
        theta[3] -= fminf(fmaxf(num[3] / den[3], -theta[3]), theta[3]);
        theta[2] = fmaxf(theta[2], 1.0f);
        theta[3] = fmaxf(theta[3], 0.01f);
    }

    /* cuda_gaussMLEv2.cu:105-133: Fisher matrix and log-likelihood. */
    div = 0.0f;
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++) {
            psfx = orc_int_gauss_1d(ii, theta[0], sigma);
            psfy = orc_int_gauss_1d(jj, theta[1], sigma);
            model = theta[3] + theta[2] * psfx * psfy;
            data = s_data[sz * jj + ii];
            deriv_int_gauss_1d(ii, theta[0], sigma, theta[2], psfy, &dudt[0], NULL);
            deriv_int_gauss_1d(jj, theta[1], sigma, theta[2], psfx, &dudt[1], NULL);
            dudt[2] = psfx * psfy;
            dudt[3] = 1.0f;
            for (int k = 0; k < NV; k++)
                for (int l = k; l < NV; l++) {
                    m[k * NV + l] += dudt[l] * dudt[k] / (gr ? model + gainr : model);
                    m[l * NV + k] = m[k * NV + l];
                }
            if (gr) { /* cuda_gaussMLEv2.cu:676-679 */
                if ((model + gainr) > 0) {
                    if ((data + gainr) > 0)
                        div += (data + gainr) * logf(model + gainr) - model - (data + gainr) * logf(data + gainr) + data;
                    else
                        div += -model - gainr;
                }
            } else if (model > 0) {
                if (data > 0)
                    div += data * logf(model) - model - data * logf(data) + data;
                else
                    div += -model;
            }
        }
    orc_mat_inv_n(m, minv, diag, NV);
    for (int k = 0; k < NV; k++) diagout[k] = diag[k];
    *llout = div;
}

void orc_mle_xynb(const float* data, float psf_sigma, int sz, int iterations,
                  float* params, float* crlb, float* ll, int nfits) {
    for (int i = 0; i < nfits; i++) {
        float theta[NV], diag[NV], l;
        mle_one(data + (size_t)sz * sz * i, NULL, psf_sigma, sz, iterations, theta, diag, &l);
        for (int k = 0; k < NV; k++) {
            params[(size_t)nfits * k + i] = theta[k];
            crlb[(size_t)nfits * k + i] = diag[k];
        }
        ll[i] = l;
    }
}

/* GPUgaussLib.cuh:90-111.  pow(v, 2.0f) is written v*v, pow(v, 3) v*v*v. */
static void deriv_int_gauss_1d_sigma(int ii, float x, float sx, float n, float psfy, float* dudt, float* d2udt2) {
    float ta = (ii + 1.0f - x) / sx;
    float tb = (ii - x) / sx;
    float ax = expf(-1.0f / 2.0f * (ta * ta));
    float bx = expf(-1.0f / 2.0f * (tb * tb));
    *dudt = -n / sqrtf(2.0f * kPi) / sx / sx * (ax * (ii - x + 1.0f) - bx * (ii - x)) * psfy;
    if (d2udt2) {
        const float p1 = ii - x + 1.0f, p0 = ii - x;
        *d2udt2 = -2.0f / sx * dudt[0] -
                  n / sqrtf(2.0f * kPi) / powf(sx, 5.0f) * (ax * (p1 * p1 * p1) - bx * (p0 * p0 * p0)) * psfy;
    }
}

/* GPUgaussLib.cuh:114-137 */
static void deriv_int_gauss_2d_sigma(int ii, int jj, float x, float y, float s, float n, float psfx, float psfy,
                                     float* dudt, float* d2udt2) {
    float dsx, dsy, ddsx, ddsy;
    deriv_int_gauss_1d_sigma(ii, x, s, n, psfy, &dsx, &ddsx);
    deriv_int_gauss_1d_sigma(jj, y, s, n, psfx, &dsy, &ddsy);
    *dudt = dsx + dsy;
    if (d2udt2) *d2udt2 = ddsx + ddsy;
}

#define NVS 5
/* cuda_gaussMLEv2.cu:144-276, one fit per loop iteration instead of per thread. */
static void mle_one_s(const float* s_data, const float* gr, float sigma, int sz, int iterations, float* theta,
                      float* diagout, float* llout) {
    float m[NVS * NVS], diag[NVS], minv[NVS * NVS];
    float dudt[NVS], d2udt2[NVS], num[NVS], den[NVS];
    float nmax, model, cf, df, data, psfx, psfy, div;
    /* sCMOS: the pixel's variance/gain^2 (gr != NULL).  Like the reference, the
     * Fisher and log-likelihood loop below reuses the value of the last pixel of
     * the last Newton iteration (cuda_gaussMLEv2.cu:620-621 set it only there). */
    float gainr = 0.0f;
    memset(m, 0, sizeof(m));
    memset(minv, 0, sizeof(minv));
    orc_center_of_mass_2d(sz, s_data, &theta[0], &theta[1]);
    orc_gauss_fmaxmin_2d(sz, sigma, s_data, &nmax, &theta[3]);
    theta[2] = fmaxf(0.0f, (nmax - theta[3]) * 2 * 2 * kPi * sigma * sigma);
    theta[4] = sigma;

    for (int kk = 0; kk < iterations; kk++) {
        memset(num, 0, sizeof(num));
        memset(den, 0, sizeof(den));
        for (int ii = 0; ii < sz; ii++)
            for (int jj = 0; jj < sz; jj++) {
                psfx = orc_int_gauss_1d(ii, theta[0], theta[4]);
                psfy = orc_int_gauss_1d(jj, theta[1], theta[4]);
                model = theta[3] + theta[2] * psfx * psfy;
                data = s_data[sz * jj + ii];
                deriv_int_gauss_1d(ii, theta[0], theta[4], theta[2], psfy, &dudt[0], &d2udt2[0]);
                deriv_int_gauss_1d(jj, theta[1], theta[4], theta[2], psfx, &dudt[1], &d2udt2[1]);
                deriv_int_gauss_2d_sigma(ii, jj, theta[0], theta[1], theta[4], theta[2], psfx, psfy, &dudt[4],
                                         &d2udt2[4]);
                dudt[2] = psfx * psfy;
                d2udt2[2] = 0.0f;
                dudt[3] = 1.0f;
                d2udt2[3] = 0.0f;
                cf = 0.0f;
                df = 0.0f;
                if (gr) {
                    /* cuda_gaussMLEv2.cu:620-632 */
                    gainr = gr[sz * jj + ii];
                    if (model > 10e-3f) cf = (data - model) / (model + gainr);
                    if (model > 10e-3f) df = (data + gainr) / ((model + gainr) * (model + gainr));
                } else {
                    if (model > 10e-3f) cf = data / model - 1;
                    if (model > 10e-3f) df = data / (model * model);
                }
                cf = fminf(cf, 10e4f);
                df = fminf(df, 10e4f);
                for (int ll = 0; ll < NVS; ll++) {
                    num[ll] += dudt[ll] * cf;
                    den[ll] += d2udt2[ll] * cf - (dudt[ll] * dudt[ll]) * df;
                }
            }
        /* cuda_gaussMLEv2.cu:229-240 */
        theta[0] -= fminf(fmaxf(num[0] / den[0], -1e0f), 1e0f);
        theta[1] -= fminf(fmaxf(num[1] / den[1], -1e0f), 1e0f);
        theta[2] -= fminf(fmaxf(num[2] / den[2], -theta[2]), theta[2]);
        theta[3] -= fminf(fmaxf(num[3] / den[3], -theta[3]), theta[3]);
        theta[4] -= fminf(fmaxf(num[4] / den[4], -theta[4]), theta[4]);
        theta[2] = fmaxf(theta[2], 1.0f);
        theta[3] = fmaxf(theta[3], 0.01f);
        theta[4] = fmaxf(theta[4], 0.5f);
        theta[4] = fminf(theta[4], sz / 2.0f);
    }

    /* cuda_gaussMLEv2.cu:243-266: Fisher matrix and log-likelihood. */
    div = 0.0f;
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++) {
            psfx = orc_int_gauss_1d(ii, theta[0], theta[4]);
            psfy = orc_int_gauss_1d(jj, theta[1], theta[4]);
            model = theta[3] + theta[2] * psfx * psfy;
            data = s_data[sz * jj + ii];
            deriv_int_gauss_1d(ii, theta[0], theta[4], theta[2], psfy, &dudt[0], NULL);
            deriv_int_gauss_1d(jj, theta[1], theta[4], theta[2], psfx, &dudt[1], NULL);
            deriv_int_gauss_2d_sigma(ii, jj, theta[0], theta[1], theta[4], theta[2], psfx, psfy, &dudt[4], NULL);
            dudt[2] = psfx * psfy;
            dudt[3] = 1.0f;
            for (int k = 0; k < NVS; k++)
                for (int l = k; l < NVS; l++) {
                    m[k * NVS + l] += dudt[l] * dudt[k] / (gr ? model + gainr : model);
                    m[l * NVS + k] = m[k * NVS + l];
                }
            if (gr) { /* cuda_gaussMLEv2.cu:676-679 */
                if ((model + gainr) > 0) {
                    if ((data + gainr) > 0)
                        div += (data + gainr) * logf(model + gainr) - model - (data + gainr) * logf(data + gainr) + data;
                    else
                        div += -model - gainr;
                }
            } else if (model > 0) {
                if (data > 0)
                    div += data * logf(model) - model - data * logf(data) + data;
                else
                    div += -model;
            }
        }
    orc_mat_inv_n(m, minv, diag, NVS);
    for (int k = 0; k < NVS; k++) diagout[k] = diag[k];
    *llout = div;
}

void orc_mle_xynbs(const float* data, float psf_sigma, int sz, int iterations,
                   float* params, float* crlb, float* ll, int nfits) {
    for (int i = 0; i < nfits; i++) {
        float theta[NVS], diag[NVS], l;
        mle_one_s(data + (size_t)sz * sz * i, NULL, psf_sigma, sz, iterations, theta, diag, &l);
        for (int k = 0; k < NVS; k++) {
            params[(size_t)nfits * k + i] = theta[k];
            crlb[(size_t)nfits * k + i] = diag[k];
        }
        ll[i] = l;
    }
}

#define NVXY 6
/* cuda_gaussMLEv2.cu:414-556: theta = {x, y, N, bg, sigma along ii, sigma along jj}. */
static void mle_one_sxsy(const float* s_data, const float* gr, float sigma, int sz, int iterations, float* theta,
                         float* diagout, float* llout) {
    float m[NVXY * NVXY], diag[NVXY], minv[NVXY * NVXY];
    float dudt[NVXY], d2udt2[NVXY], num[NVXY], den[NVXY];
    float nmax, model, cf, df, data, psfx, psfy, div;
    /* sCMOS: the pixel's variance/gain^2 (gr != NULL).  Like the reference, the
     * Fisher and log-likelihood loop below reuses the value of the last pixel of
     * the last Newton iteration (cuda_gaussMLEv2.cu:620-621 set it only there). */
    float gainr = 0.0f;
    memset(m, 0, sizeof(m));
    memset(minv, 0, sizeof(minv));
    orc_center_of_mass_2d(sz, s_data, &theta[0], &theta[1]);
    orc_gauss_fmaxmin_2d(sz, sigma, s_data, &nmax, &theta[3]);
    theta[2] = fmaxf(0.0f, (nmax - theta[3]) * 2 * 2 * kPi * sigma * sigma);
    theta[4] = sigma;
    theta[5] = sigma;
    for (int kk = 0; kk < iterations; kk++) {
        memset(num, 0, sizeof(num));
        memset(den, 0, sizeof(den));
        for (int ii = 0; ii < sz; ii++)
            for (int jj = 0; jj < sz; jj++) {
                psfx = orc_int_gauss_1d(ii, theta[0], theta[4]);
                psfy = orc_int_gauss_1d(jj, theta[1], theta[5]);
                model = theta[3] + theta[2] * psfx * psfy;
                data = s_data[sz * jj + ii];
                deriv_int_gauss_1d(ii, theta[0], theta[4], theta[2], psfy, &dudt[0], &d2udt2[0]);
                deriv_int_gauss_1d(jj, theta[1], theta[5], theta[2], psfx, &dudt[1], &d2udt2[1]);
                deriv_int_gauss_1d_sigma(ii, theta[0], theta[4], theta[2], psfy, &dudt[4], &d2udt2[4]);
                deriv_int_gauss_1d_sigma(jj, theta[1], theta[5], theta[2], psfx, &dudt[5], &d2udt2[5]);
                dudt[2] = psfx * psfy;
                d2udt2[2] = 0.0f;
                dudt[3] = 1.0f;
                d2udt2[3] = 0.0f;
                cf = 0.0f;
                df = 0.0f;
                if (gr) {
                    /* cuda_gaussMLEv2.cu:620-632 */
                    gainr = gr[sz * jj + ii];
                    if (model > 10e-3f) cf = (data - model) / (model + gainr);
                    if (model > 10e-3f) df = (data + gainr) / ((model + gainr) * (model + gainr));
                } else {
                    if (model > 10e-3f) cf = data / model - 1;
                    if (model > 10e-3f) df = data / (model * model);
                }
                cf = fminf(cf, 10e4f);
                df = fminf(df, 10e4f);
                for (int ll = 0; ll < NVXY; ll++) {
                    num[ll] += dudt[ll] * cf;
                    den[ll] += d2udt2[ll] * cf - (dudt[ll] * dudt[ll]) * df;
                }
            }
        /* cuda_gaussMLEv2.cu:500-513 */
        theta[0] -= fminf(fmaxf(num[0] / den[0], -1e0f), 1e0f);
        theta[1] -= fminf(fmaxf(num[1] / den[1], -1e0f), 1e0f);
        theta[2] -= fminf(fmaxf(num[2] / den[2], -theta[2]), theta[2]);
        theta[3] -= fminf(fmaxf(num[3] / den[3], -theta[3]), theta[3]);
        theta[4] -= fminf(fmaxf(num[4] / den[4], -theta[4]), theta[4]);
        theta[5] -= fminf(fmaxf(num[5] / den[5], -theta[5]), theta[5]);
        theta[2] = fmaxf(theta[2], 1.0f);
        theta[3] = fmaxf(theta[3], 0.01f);
        theta[4] = fmaxf(theta[4], 0.5f);
        theta[4] = fminf(theta[4], sz / 2.0f);
        theta[5] = fmaxf(theta[5], 0.5f);
        theta[5] = fminf(theta[5], sz / 2.0f);
    }
    /* cuda_gaussMLEv2.cu:518-546 */
    div = 0.0f;
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++) {
            psfx = orc_int_gauss_1d(ii, theta[0], theta[4]);
            psfy = orc_int_gauss_1d(jj, theta[1], theta[5]);
            model = theta[3] + theta[2] * psfx * psfy;
            data = s_data[sz * jj + ii];
            deriv_int_gauss_1d(ii, theta[0], theta[4], theta[2], psfy, &dudt[0], NULL);
            deriv_int_gauss_1d(jj, theta[1], theta[5], theta[2], psfx, &dudt[1], NULL);
            deriv_int_gauss_1d_sigma(ii, theta[0], theta[4], theta[2], psfy, &dudt[4], NULL);
            deriv_int_gauss_1d_sigma(jj, theta[1], theta[5], theta[2], psfx, &dudt[5], NULL);
            dudt[2] = psfx * psfy;
            dudt[3] = 1.0f;
            for (int k = 0; k < NVXY; k++)
                for (int l = k; l < NVXY; l++) {
                    m[k * NVXY + l] += dudt[l] * dudt[k] / (gr ? model + gainr : model);
                    m[l * NVXY + k] = m[k * NVXY + l];
                }
            if (gr) { /* cuda_gaussMLEv2.cu:676-679 */
                if ((model + gainr) > 0) {
                    if ((data + gainr) > 0)
                        div += (data + gainr) * logf(model + gainr) - model - (data + gainr) * logf(data + gainr) + data;
                    else
                        div += -model - gainr;
                }
            } else if (model > 0) {
                if (data > 0)
                    div += data * logf(model) - model - data * logf(data) + data;
                else
                    div += -model;
            }
        }
    orc_mat_inv_n(m, minv, diag, NVXY);
    for (int k = 0; k < NVXY; k++) diagout[k] = diag[k];
    *llout = div;
}

void orc_mle_xynbsxsy(const float* data, float psf_sigma, int sz, int iterations,
                      float* params, float* crlb, float* ll, int nfits) {
    for (int i = 0; i < nfits; i++) {
        float theta[NVXY], diag[NVXY], l;
        mle_one_sxsy(data + (size_t)sz * sz * i, NULL, psf_sigma, sz, iterations, theta, diag, &l);
        for (int k = 0; k < NVXY; k++) {
            params[(size_t)nfits * k + i] = theta[k];
            crlb[(size_t)nfits * k + i] = diag[k];
        }
        ll[i] = l;
    }
}

/* GPUgaussLib.cuh:26-64.  pow(v, k) for integer k is written as products. */
static float alpha_z(float z, float a, float b, float d) {
    const float t = z / d;
    return 1.0f + (t * t) + a * (t * t * t) + b * (t * t * t * t);
}
static float dalphadz(float z, float a, float b, float d) {
    return (2.0f * z / (d * d) + 3.0f * a * (z * z) / (d * d * d) + 4.0f * b * (z * z * z) / powf(d, 4.0f));
}
static float d2alphadz2(float z, float a, float b, float d) {
    return (2.0f / (d * d) + 6.0f * a * z / (d * d * d) + 12.0f * b * (z * z) / powf(d, 4.0f));
}

typedef struct { float sx0, ax, ay, bx, by, gamma, d, sy0; } ZCal;

/* GPUgaussLib.cuh:140-205 */
static void deriv_int_gauss_2d_z(int ii, int jj, const float* theta, const ZCal* c, float* ppsfx, float* ppsfy,
                                 float* dudt, float* d2udt2) {
    float dsx, dsy, ddsx, ddsy, ddx, ddy;
    const float z = theta[4];
    const float alphax = alpha_z(z - c->gamma, c->ax, c->bx, c->d);
    const float alphay = alpha_z(z + c->gamma, c->ay, c->by, c->d);
    const float sx = c->sx0 * sqrtf(alphax);
    const float sy = c->sy0 * sqrtf(alphay);
    const float psfx = orc_int_gauss_1d(ii, theta[0], sx);
    const float psfy = orc_int_gauss_1d(jj, theta[1], sy);
    *ppsfx = psfx;
    *ppsfy = psfy;
    deriv_int_gauss_1d(ii, theta[0], sx, theta[2], psfy, &dudt[0], &ddx);
    deriv_int_gauss_1d(jj, theta[1], sy, theta[2], psfx, &dudt[1], &ddy);
    deriv_int_gauss_1d_sigma(ii, theta[0], sx, theta[2], psfy, &dsx, &ddsx);
    deriv_int_gauss_1d_sigma(jj, theta[1], sy, theta[2], psfx, &dsy, &ddsy);
    const float dsdalpha_x = c->sx0 / 2.0f / sqrtf(alphax);
    const float dsdalpha_y = c->sy0 / 2.0f / sqrtf(alphay);
    const float dsdzx = dsdalpha_x * dalphadz(z - c->gamma, c->ax, c->bx, c->d);
    const float dsdzy = dsdalpha_y * dalphadz(z + c->gamma, c->ay, c->by, c->d);
    dudt[4] = dsx * dsdzx + dsy * dsdzy;
    if (d2udt2) {
        d2udt2[0] = ddx;
        d2udt2[1] = ddy;
        const float d2x = -c->sx0 / 4.0f / powf(alphax, 1.5f);
        const float d2y = -c->sy0 / 4.0f / powf(alphay, 1.5f);
        const float dax = dalphadz(z - c->gamma, c->ax, c->bx, c->d);
        const float day = dalphadz(z + c->gamma, c->ay, c->by, c->d);
        const float ddsddzx = d2x * (dax * dax) + dsdalpha_x * d2alphadz2(z - c->gamma, c->ax, c->bx, c->d);
        const float ddsddzy = d2y * (day * day) + dsdalpha_y * d2alphadz2(z + c->gamma, c->ay, c->by, c->d);
        d2udt2[4] = ddsx * (dsdzx * dsdzx) + dsx * ddsddzx + ddsy * (dsdzy * dsdzy) + dsy * ddsddzy;
    }
}

/* cuda_gaussMLEv2.cu:280-411: theta = {x, y, N, bg, z}. */
static void mle_one_z(const float* s_data, const float* gr, float z0, const ZCal* c, int sz, int iterations,
                      float* theta, float* diagout, float* llout) {
    float m[NVS * NVS], diag[NVS], minv[NVS * NVS];
    float dudt[NVS], d2udt2[NVS], num[NVS], den[NVS];
    float nmax, model, cf, df, data, psfx, psfy, div;
    /* sCMOS: the pixel's variance/gain^2 (gr != NULL).  Like the reference, the
     * Fisher and log-likelihood loop below reuses the value of the last pixel of
     * the last Newton iteration (cuda_gaussMLEv2.cu:620-621 set it only there). */
    float gainr = 0.0f;
    memset(m, 0, sizeof(m));
    memset(minv, 0, sizeof(minv));
    orc_center_of_mass_2d(sz, s_data, &theta[0], &theta[1]);
    orc_gauss_fmaxmin_2d(sz, c->sx0, s_data, &nmax, &theta[3]);
    theta[2] = fmaxf(0.0f, (nmax - theta[3]) * 2 * kPi * c->sx0 * c->sy0 * sqrtf(2.0f));
    theta[4] = z0; /* 0 (cuda_gaussMLEv2.cu:332), or z_initial of the sCMOS kernel (:896) */
    for (int kk = 0; kk < iterations; kk++) {
        memset(num, 0, sizeof(num));
        memset(den, 0, sizeof(den));
        for (int ii = 0; ii < sz; ii++)
            for (int jj = 0; jj < sz; jj++) {
                deriv_int_gauss_2d_z(ii, jj, theta, c, &psfx, &psfy, dudt, d2udt2);
                model = theta[3] + theta[2] * psfx * psfy;
                data = s_data[sz * jj + ii];
                dudt[2] = psfx * psfy;
                d2udt2[2] = 0.0f;
                dudt[3] = 1.0f;
                d2udt2[3] = 0.0f;
                cf = 0.0f;
                df = 0.0f;
                if (gr) {
                    /* cuda_gaussMLEv2.cu:620-632 */
                    gainr = gr[sz * jj + ii];
                    if (model > 10e-3f) cf = (data - model) / (model + gainr);
                    if (model > 10e-3f) df = (data + gainr) / ((model + gainr) * (model + gainr));
                } else {
                    if (model > 10e-3f) cf = data / model - 1;
                    if (model > 10e-3f) df = data / (model * model);
                }
                cf = fminf(cf, 10e4f);
                df = fminf(df, 10e4f);
                for (int ll = 0; ll < NVS; ll++) {
                    num[ll] += dudt[ll] * cf;
                    den[ll] += d2udt2[ll] * cf - (dudt[ll] * dudt[ll]) * df;
                }
            }
        /* cuda_gaussMLEv2.cu:365-373 */
        theta[0] -= fminf(fmaxf(num[0] / den[0], -1e0f), 1e0f);
        theta[1] -= fminf(fmaxf(num[1] / den[1], -1e0f), 1e0f);
        theta[2] -= fminf(fmaxf(num[2] / den[2], -theta[2]), theta[2]);
        theta[3] -= fminf(fmaxf(num[3] / den[3], -theta[3]), theta[3]);
        theta[4] -= fminf(fmaxf(num[4] / den[4], -1e-1f), 1e-1f);
        theta[2] = fmaxf(theta[2], 1.0f);
        theta[3] = fmaxf(theta[3], 0.01f);
    }
    div = 0.0f;
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++) {
            deriv_int_gauss_2d_z(ii, jj, theta, c, &psfx, &psfy, dudt, NULL);
            model = theta[3] + theta[2] * psfx * psfy;
            data = s_data[sz * jj + ii];
            dudt[2] = psfx * psfy;
            dudt[3] = 1.0f;
            for (int k = 0; k < NVS; k++)
                for (int l = k; l < NVS; l++) {
                    m[k * NVS + l] += dudt[l] * dudt[k] / (gr ? model + gainr : model);
                    m[l * NVS + k] = m[k * NVS + l];
                }
            if (gr) { /* cuda_gaussMLEv2.cu:676-679 */
                if ((model + gainr) > 0) {
                    if ((data + gainr) > 0)
                        div += (data + gainr) * logf(model + gainr) - model - (data + gainr) * logf(data + gainr) + data;
                    else
                        div += -model - gainr;
                }
            } else if (model > 0) {
                if (data > 0)
                    div += data * logf(model) - model - data * logf(data) + data;
                else
                    div += -model;
            }
        }
    orc_mat_inv_n(m, minv, diag, NVS);
    for (int k = 0; k < NVS; k++) diagout[k] = diag[k];
    *llout = div;
}

void orc_mle_xynbz(const float* data, float sx0, float ax, float ay, float bx, float by, float gamma, float d,
                   float sy0, int sz, int iterations, float* params, float* crlb, float* ll, int nfits) {
    const ZCal c = {sx0, ax, ay, bx, by, gamma, d, sy0};
    for (int i = 0; i < nfits; i++) {
        float theta[NVS], diag[NVS], l;
        mle_one_z(data + (size_t)sz * sz * i, NULL, 0.0f, &c, sz, iterations, theta, diag, &l);
        for (int k = 0; k < NVS; k++) {
            params[(size_t)nfits * k + i] = theta[k];
            crlb[(size_t)nfits * k + i] = diag[k];
        }
        ll[i] = l;
    }
}

/* The sCMOS kernels (cuda_gaussMLEv2.cu:559-1124): fit type 0 _SCMOSXYNB_, 1
 * _SCMOSXYNBS_, 2 _SCMOSXYNBSXSY_, 3 _SCMOSXYNBZ_ (zcal = {sx0, ax, ay, bx, by,
 * gamma, d, sy0}, z0 one starting z per fit).  Pixel (ii, jj) of fit i reads the
 * gain-ratio map at (coords[2i+1] + jj) * mapsz + coords[2i] + ii, formed in
 * float as the reference forms GRind. */
void orc_mle_scmos(int fit_type, const float* data, const float* coords, const float* gain, int mapsz,
                   const float* z0, float psf_sigma, const float* zcal, int sz, int iterations, float* params,
                   float* crlb, float* ll, int nfits) {
    const int np = fit_type == 2 ? 6 : fit_type == 0 ? 4 : 5;
    float gr[64 * 64];
    ZCal c = {psf_sigma, 0, 0, 0, 0, 0, 1, psf_sigma};
    if (fit_type == 3) {
        c.sx0 = zcal[0]; c.ax = zcal[1]; c.ay = zcal[2]; c.bx = zcal[3]; c.by = zcal[4];
        c.gamma = zcal[5]; c.d = zcal[6]; c.sy0 = zcal[7];
    }
    for (int i = 0; i < nfits; i++) {
        float theta[6], diag[6], l;
        for (int ii = 0; ii < sz; ii++)
            for (int jj = 0; jj < sz; jj++) {
                const int idx = (int)((coords[2 * i + 1] + jj) * mapsz + coords[2 * i] + ii);
                gr[sz * jj + ii] = gain[idx];
            }
        const float* d = data + (size_t)sz * sz * i;
        switch (fit_type) {
        case 0: mle_one(d, gr, psf_sigma, sz, iterations, theta, diag, &l); break;
        case 1: mle_one_s(d, gr, psf_sigma, sz, iterations, theta, diag, &l); break;
        case 2: mle_one_sxsy(d, gr, psf_sigma, sz, iterations, theta, diag, &l); break;
        default: mle_one_z(d, gr, z0[i], &c, sz, iterations, theta, diag, &l); break;
        }
        for (int k = 0; k < np; k++) {
            params[(size_t)nfits * k + i] = theta[k];
            crlb[(size_t)nfits * k + i] = diag[k];
        }
        ll[i] = l;
    }
}
