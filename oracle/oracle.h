/*
 * CPU oracle for the BAMF hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This library is a plain-C restatement of the reference algorithms
 * (MohamadFazel/BAMF) used as the checker for the HIP implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the product library (bamf_amd/libbamf_gpu.so) never links or calls it.
 *
 * Every function cites the reference file:line it follows.  The reference
 * C++ (source/c/LibRJMCMC.cpp) includes <mex.h>, a MATLAB header that this
 * image does not have, so it is not compiled here; the restatement is pinned
 * by closed forms, brute force and textbook routines in tests/.
 *
 * Arithmetic semantics: IEEE single precision, no FMA contraction
 * (-ffp-contract=off), the same as the reference's x86-64 / MSVC builds.
 */
#ifndef BAMF_ORACLE_H
#define BAMF_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 (Salmon et al., SC'11), published algorithm ---------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Raw 31-bit draw j of step `step` of chain `roi` (same layout as the GPU):
 * word step % 4 of Philox4x32-10 at counter (step / 4, j, seed_hi, 0), key
 * (roi, seed_lo), shifted right by one. */
uint32_t orc_chain_raw(uint64_t seed, uint32_t roi, uint32_t step, uint32_t j);

/* ---- gaussMLE XYNB (cuda_gaussMLEv2.cu:16-141) --------------------------- */
void orc_mle_xynb(const float* data, float psf_sigma, int sz, int iterations,
                  float* params, float* crlb, float* ll, int nfits);
/* kernel_MLEFit_XYNBS_ (cuda_gaussMLEv2.cu:144-276): theta = {x, y, N, bg, sigma}. */
void orc_mle_xynbs(const float* data, float psf_sigma, int sz, int iterations,
                   float* params, float* crlb, float* ll, int nfits);
/* kernel_MLEFit_XYNBSXSY_ (cuda_gaussMLEv2.cu:414-556): theta = {x, y, N, bg, sigma_ii, sigma_jj}. */
void orc_mle_xynbsxsy(const float* data, float psf_sigma, int sz, int iterations,
                      float* params, float* crlb, float* ll, int nfits);
/* kernel_MLEFit_XYNBZ_ (cuda_gaussMLEv2.cu:280-411): theta = {x, y, N, bg, z}, widths
 * sx0 sqrt(alpha(z - gamma; ax, bx, d)) along ii and sy0 sqrt(alpha(z + gamma; ay, by, d)) along jj. */
void orc_mle_xynbz(const float* data, float sx0, float ax, float ay, float bx, float by, float gamma, float d,
                   float sy0, int sz, int iterations, float* params, float* crlb, float* ll, int nfits);
/* The sCMOS kernels (cuda_gaussMLEv2.cu:559-1124): fit_type 0 XYNB, 1 XYNBS, 2 XYNBSXSY, 3 XYNBZ. */
void orc_mle_scmos(int fit_type, const float* data, const float* coords, const float* gain, int mapsz,
                   const float* z0, float psf_sigma, const float* zcal, int sz, int iterations, float* params,
                   float* crlb, float* ll, int nfits);
/* Helpers exposed for unit tests. */
float orc_int_gauss_1d(int ii, float x, float sigma);
void  orc_center_of_mass_2d(int sz, const float* data, float* x, float* y);
void  orc_gauss_fmaxmin_2d(int sz, float sigma, const float* data, float* maxn, float* minbg);
void  orc_mat_inv_n(float* m, float* minv, float* diag, int sz);

/* ---- FindROI (cuda_FindROI.cu:10-147, findROI.m:60-225) ------------------ */
void orc_gauss_coeffs(float sigma, float* b0, float* b1, float* b2, float* b3, float* B);
void orc_gauss_x(float* d, int xsize, int ysize, int nframes,
                 float b0, float b1, float b2, float b3, float B);
void orc_gauss_y(float* d, int xsize, int ysize, int nframes,
                 float b0, float b1, float b2, float b3, float B);
void orc_subtract(float* a, const float* b, size_t n);
void orc_max_x(const float* a, float* b, int xsize, int ysize, int nframes,
               int kernelsz, float minval);
void orc_max_y(float* a, const float* b, int xsize, int ysize, int nframes,
               int kernelsz, float minval);
/* Whole findROI: flag map (0/1 per pixel, float like the reference's Res)
 * and the ordered detections (1-based row, col, frame + box corners).
 * Returns the number of detections (<= cap written). */
int64_t orc_find_roi(const float* stack, int xsize, int ysize, int nframes,
                     float sigma1, float sigma2, int boxsize, float minval,
                     float* res, int64_t cap, int64_t* lin_idx,
                     int32_t* start_row, int32_t* start_col, int32_t* frame);

/* ---- RJMCMC chain (LibRJMCMC.cpp, cRJMCMC.cpp) --------------------------- */
#include "oracle_rjmcmc.h"

#ifdef __cplusplus
}
#endif
#endif
