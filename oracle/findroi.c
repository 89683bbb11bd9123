/*
 * Oracle: difference-of-Gaussian ROI finder.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates the five kernels of source/cuda/cuda_FindROI/cuda_FindROI.cu
 * (gaussX :10-46, gaussY :48-83, subtract :85-92, maxX :94-117, maxY :119-147)
 * as host loops over (frame, row/column), and the MATLAB driver
 * @RJ/findROI.m:60-225 (coefficients :169-182, kernel chaining :114-117 and
 * :201-224, find/ind2sub :125-126, box corners :130-163).
 *
 * Layout follows the kernels: a frame is `ysize` runs of `xsize` contiguous
 * floats (kernel_gaussX's base = idz*X*Y + idy*X), frames back to back.
 * findROI.m passes Xsize = size(Image,2), Ysize = size(Image,1).
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include "oracle.h"

/* findROI.m:169-182: Young-van Vliet coefficients, double on the host, then
 * converted to single by CUDAKernel/feval for the float kernel arguments. */
void orc_gauss_coeffs(float sigma, float* b0, float* b1, float* b2, float* b3, float* B) {
    double s = (double)sigma, q;
    if (s > 2.5)
        q = 0.98711 * s - 0.96330;
    else
        q = 3.97156 - 4.14554 * sqrt(1 - 0.26891 * s);
    double qq = q * q, qqq = qq * q;
    double B0 = 1.57825 + 2.44413 * q + 1.4281 * qq + 0.422205 * qqq;
    double B1 = 2.44413 * q + 2.85619 * qq + 1.26661 * qqq;
    double B2 = -1.4281 * qq - 1.26661 * qqq;
    double B3 = 0.422205 * qqq;
    double BB = 1 - (B1 + B2 + B3) / B0;
    *b0 = (float)B0; *b1 = (float)B1; *b2 = (float)B2; *b3 = (float)B3; *B = (float)BB;
}

/* One causal + anti-causal pass over n samples at stride st (cuda_FindROI.cu:23-45). */
static void iir_run(float* d, int n, size_t st, float b0, float b1, float b2, float b3, float B) {
    float w0, w1, w2, w3, temp;
    w1 = w2 = w3 = d[0];
    for (int ii = 0; ii < n; ii++) {
        w0 = d[ii * st];
        temp = w0 * B + (b1 * w1 + b2 * w2 + b3 * w3) / b0;
        d[ii * st] = temp;
        w3 = w2; w2 = w1; w1 = temp;
    }
    w1 = w2 = w3 = d[(size_t)(n - 1) * st];
    for (int ii = n - 1; ii >= 0; ii--) {
        w0 = d[ii * st];
        temp = w0 * B + (b1 * w1 + b2 * w2 + b3 * w3) / b0;
        d[ii * st] = temp;
        w3 = w2; w2 = w1; w1 = temp;
    }
}

/* cuda_FindROI.cu:10-46 */
void orc_gauss_x(float* d, int xsize, int ysize, int nframes,
                 float b0, float b1, float b2, float b3, float B) {
    for (int z = 0; z < nframes; z++)
        for (int y = 0; y < ysize; y++)
            iir_run(d + (size_t)z * xsize * ysize + (size_t)y * xsize, xsize, 1, b0, b1, b2, b3, B);
}

/* cuda_FindROI.cu:48-83 */
void orc_gauss_y(float* d, int xsize, int ysize, int nframes,
                 float b0, float b1, float b2, float b3, float B) {
    for (int z = 0; z < nframes; z++)
        for (int x = 0; x < xsize; x++)
            iir_run(d + (size_t)z * xsize * ysize + x, ysize, (size_t)xsize, b0, b1, b2, b3, B);
}

/* cuda_FindROI.cu:85-92 */
void orc_subtract(float* a, const float* b, size_t n) {
    for (size_t i = 0; i < n; i++) a[i] = a[i] - b[i];
}

/* cuda_FindROI.cu:94-117: row-window max seeded at minval; +v if the pixel is the max, else -max. */
void orc_max_x(const float* a, float* b, int xsize, int ysize, int nframes, int k, float minval) {
    for (int z = 0; z < nframes; z++)
        for (int y = 0; y < ysize; y++) {
            const float* row = a + (size_t)z * xsize * ysize + (size_t)y * xsize;
            float* out = b + (size_t)z * xsize * ysize + (size_t)y * xsize;
            for (int x = 0; x < xsize; x++) {
                int start = x - k < 0 ? 0 : x - k;
                int end = x + k > xsize - 1 ? xsize - 1 : x + k;
                float maxval = minval, inpixel = row[x];
                for (int ii = start; ii < end + 1; ii++) maxval = fmaxf(maxval, row[ii]);
                float t1 = (float)(maxval > inpixel) * -maxval;
                float t2 = (float)(maxval == inpixel) * maxval;
                out[x] = t1 + t2;
            }
        }
}

/* cuda_FindROI.cu:119-147: column-window max of |b|; flag 1 where |max - b| < 1e-6. */
void orc_max_y(float* a, const float* b, int xsize, int ysize, int nframes, int k, float minval) {
    for (int z = 0; z < nframes; z++)
        for (int x = 0; x < xsize; x++)
            for (int y = 0; y < ysize; y++) {
                size_t base = (size_t)z * xsize * ysize;
                size_t idx = x + (size_t)xsize * y + base;
                int start = y - k < 0 ? 0 : y - k;
                int end = y + k > ysize - 1 ? ysize - 1 : y + k;
                float maxval = minval, inpixel = b[idx];
                for (int ii = start; ii <= end; ii++)
                    maxval = fmaxf(maxval, fabsf(b[(size_t)ii * xsize + x + base]));
                a[idx] = ((double)fabsf(maxval - inpixel) < 1e-6) ? 1.0f : 0.0f;
            }
}

/* findROI.m:60-163 for one chunk holding the whole stack. */
int64_t orc_find_roi(const float* stack, int xsize, int ysize, int nframes,
                     float sigma1, float sigma2, int boxsize, float minval,
                     float* res, int64_t cap, int64_t* lin_idx,
                     int32_t* start_row, int32_t* start_col, int32_t* frame) {
    size_t n = (size_t)xsize * ysize * nframes;
    float* da = (float*)malloc(n * sizeof(float));
    float* db = (float*)malloc(n * sizeof(float));
    float b0, b1, b2, b3, B;
    /* gauss_inplace(Image, Sigma1) and (Image, Sigma2): findROI.m:114-115, 166-200 */
    memcpy(da, stack, n * sizeof(float));
    orc_gauss_coeffs(sigma1, &b0, &b1, &b2, &b3, &B);
    orc_gauss_x(da, xsize, ysize, nframes, b0, b1, b2, b3, B);
    orc_gauss_y(da, xsize, ysize, nframes, b0, b1, b2, b3, B);
    memcpy(db, stack, n * sizeof(float));
    orc_gauss_coeffs(sigma2, &b0, &b1, &b2, &b3, &B);
    orc_gauss_x(db, xsize, ysize, nframes, b0, b1, b2, b3, B);
    orc_gauss_y(db, xsize, ysize, nframes, b0, b1, b2, b3, B);
    /* subtract_inplace: SubIm = D_A - D_B (findROI.m:116, 201-209) */
    orc_subtract(da, db, n);
    /* local_max: maxX(SubIm -> Im1), maxY(Im1 -> Res)  (findROI.m:117, 210-224) */
    orc_max_x(da, db, xsize, ysize, nframes, boxsize, minval);
    orc_max_y(res, db, xsize, ysize, nframes, boxsize, minval);
    free(da);
    free(db);

    /* find(Res) in linear (column-major) order, ind2sub with size(Res) = [Ysize Xsize Z]
     * (findROI.m:125-126), then box corners (findROI.m:130-148). */
    int64_t count = 0;
    int halfbox = (int)floor(boxsize / 2.0 + 0.5);
    for (size_t i = 0; i < n; i++) {
        if (res[i] == 0.0f) continue;
        if (count < cap) {
            int64_t lin = (int64_t)i;
            int ii = (int)(lin % ysize) + 1;
            int jj = (int)((lin / ysize) % xsize) + 1;
            int kk = (int)(lin / ((int64_t)ysize * xsize)) + 1;
            int sr = ii - halfbox + 1, sc = jj - halfbox + 1;
            if (sr < 1) sr = 1;
            if (sc < 1) sc = 1;
            int er = sr + boxsize - 1, ec = sc + boxsize - 1;
            if (er > ysize) sr = ysize - boxsize + 1;
            if (ec > xsize) sc = xsize - boxsize + 1;
            if (lin_idx) lin_idx[count] = lin;
            if (start_row) start_row[count] = sr;
            if (start_col) start_col[count] = sc;
            if (frame) frame[count] = kk;
        }
        count++;
    }
    return count;
}
