/*
 * bamf_gpu.h — C ABI of libbamf_gpu.so, the MI355X (gfx950) implementation of
 * BAMF's data-parallel hot path:
 *   - the single-emitter Newton MLE that feeds the intensity prior,
 *   - the difference-of-Gaussian ROI finder,
 *   - the per-ROI reversible-jump MCMC chain.
 *
 * Conventions
 *   - Every entry point returns int: BAMF_OK (0) or a negative BAMF_E_* code;
 *     bamf_last_error() gives the message for the calling thread.  Nothing
 *     throws across the ABI.
 *   - d_* arguments are device pointers owned by the caller; h_* are host
 *     pointers.  `stream` is a hipStream_t passed as void* (NULL = default
 *     stream).  Calls are asynchronous on `stream` unless stated otherwise.
 *   - No torch or C++ types appear in any signature.
 */
#ifndef BAMF_GPU_H
#define BAMF_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BAMF_ABI_VERSION 1

#define BAMF_OK 0
#define BAMF_E_INVALID_ARG (-1)
#define BAMF_E_HIP (-2)
#define BAMF_E_UNSUPPORTED (-3)
#define BAMF_E_CAPACITY (-4)

const char* bamf_last_error(void);
int bamf_version(void);
int bamf_device_info(int device, int* cu_count, int* clock_khz, char* name, int name_len);

/* ------------------------------------------------------------------------
 * gaussMLE, fit type 'Basic' (theta = {x, y, N, bg}).
 * Replaces kernel_MLEFit_XYNB_ (source/cuda/cuda_gaussMLEv2/cuda_gaussMLEv2.cu:16-17),
 * launched by @RJ/gaussMLE.m:119,154.  Same argument meaning and layouts:
 *   d_data          nfits boxes of sz*sz floats, box-contiguous, pixel (ii, jj) at sz*jj + ii;
 *   d_params/d_crlb parameter-major: entry k of fit i at [nfits*k + i], k = x, y, N, bg;
 *   d_ll            one log-likelihood per fit.
 * gaussMLE.m:183-190 maps k = 0,1,2,3 to Y, X, Photons, BG.
 * ------------------------------------------------------------------------ */
int bamf_gauss_mle_xynb(const float* d_data, float psf_sigma, int sz, int iterations,
                        float* d_params, float* d_crlb, float* d_ll, int nfits, void* stream);

/* ------------------------------------------------------------------------
 * FindROI.  A frame is `ysize` runs of `xsize` contiguous floats (the
 * kernels' Xsize/Ysize, cuda_FindROI.cu:14-21; findROI.m passes
 * Xsize = size(Image,2), Ysize = size(Image,1)); frames are back to back.
 *
 * The reference's five kernels, one for one (cuda_FindROI.cu:10,48,85,94,119):
 * ------------------------------------------------------------------------ */
/* Young-van Vliet coefficients as findROI.m:169-182 computes them (double,
 * then single): out = {b0, b1, b2, b3, B}.  Host only. */
void bamf_gauss_coeffs(float sigma, float out[5]);
int bamf_find_roi_gauss_x(float* d, int xsize, int ysize, int nframes,
                          float b0, float b1, float b2, float b3, float B, void* stream);
int bamf_find_roi_gauss_y(float* d, int xsize, int ysize, int nframes,
                          float b0, float b1, float b2, float b3, float B, void* stream);
int bamf_find_roi_subtract(float* d_a, const float* d_b, int64_t n, void* stream);
int bamf_find_roi_max_x(const float* d_a, float* d_b, int xsize, int ysize, int nframes,
                        int kernelsz, float minval, void* stream);
int bamf_find_roi_max_y(float* d_a, const float* d_b, int xsize, int ysize, int nframes,
                        int kernelsz, float minval, void* stream);

/* The whole of findROI.m:60-163 on the device: gauss(sigma1) and
 * gauss(sigma2) of the stack, their difference, the separable local maximum
 * (boxsize half-window, minval floor), then the detections in MATLAB find()
 * order (frame-major, then column-major linear index), each with the
 * 1-based (row, col, frame) box corner of findROI.m:130-148 (SMD.YBoxCorner,
 * SMD.XBoxCorner, SMD.FrameNum).
 *   d_workspace    >= bamf_find_roi_workspace_bytes(xsize, ysize, nframes) bytes;
 *   d_res_bits     optional (may be NULL): the Res map as bits, bit (i & 31)
 *                  of word i >> 5 for linear pixel i; ceil(n/32) words;
 *   d_lin_idx etc. capacity entries each (any may be NULL);
 *   d_count        receives the number of detections (it may exceed capacity;
 *                  then only the first capacity entries are written).
 * The input stack is not modified. */
size_t bamf_find_roi_workspace_bytes(int xsize, int ysize, int nframes);
int bamf_find_roi(const float* d_stack, int xsize, int ysize, int nframes,
                  float sigma1, float sigma2, int boxsize, float minval,
                  void* d_workspace, size_t workspace_bytes, uint32_t* d_res_bits,
                  int64_t* d_lin_idx, int32_t* d_start_row, int32_t* d_start_col,
                  int32_t* d_frame, int64_t capacity, int64_t* d_count, void* stream);
/* ROIStack crop (findROI.m:157-160): box b is boxsize x boxsize, column-major
 * (row r, col c at r + boxsize*c), cut from frame d_frame[b] at the 1-based corner. */
int bamf_find_roi_crop(const float* d_stack, int xsize, int ysize,
                       const int32_t* d_start_row, const int32_t* d_start_col,
                       const int32_t* d_frame, int64_t n, int boxsize,
                       float* d_roi_stack, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* BAMF_GPU_H */
