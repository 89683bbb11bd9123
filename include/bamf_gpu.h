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
/* gaussMLE fit type 'Sigma' (theta = {x, y, N, bg, sigma}): replaces
 * kernel_MLEFit_XYNBS_ (cuda_gaussMLEv2.cu:144-145), bound by gaussMLE.m:98-100.
 * Same layouts with five parameters; psf_sigma is the starting width, and the
 * fitted width is kept in [0.5, sz/2] (cuda_gaussMLEv2.cu:237-238). */
int bamf_gauss_mle_xynbs(const float* d_data, float psf_sigma, int sz, int iterations,
                         float* d_params, float* d_crlb, float* d_ll, int nfits, void* stream);
/* gaussMLE fit type 'SigmaYX' (theta = {x, y, N, bg, sigma along ii, sigma
 * along jj}): replaces kernel_MLEFit_XYNBSXSY_ (cuda_gaussMLEv2.cu:414-415),
 * bound by gaussMLE.m:106-109.  Six parameters; both widths start at psf_sigma
 * and are kept in [0.5, sz/2] (cuda_gaussMLEv2.cu:510-513). */
int bamf_gauss_mle_xynbsxsy(const float* d_data, float psf_sigma, int sz, int iterations,
                            float* d_params, float* d_crlb, float* d_ll, int nfits, void* stream);
/* gaussMLE fit type 'Z' (theta = {x, y, N, bg, z}, astigmatic PSF): replaces
 * kernel_MLEFit_XYNBZ_ (cuda_gaussMLEv2.cu:280-282), called by gaussMLE.m:155-158
 * with Sigma(1), ZStruct.Ax, Ay, Bx, By, Gamma, D, Sigma(2).  The widths are
 * psf_sigma_x sqrt(alpha(z - gamma; ax, bx, d)) along the box rows and
 * psf_sigma_y sqrt(alpha(z + gamma; ay, by, d)) along its columns, alpha(t) =
 * 1 + (t/d)^2 + A (t/d)^3 + B (t/d)^4; z starts at 0 and moves at most 0.1 per
 * iteration. */
int bamf_gauss_mle_xynbz(const float* d_data, float psf_sigma_x, float ax, float ay, float bx, float by,
                         float gamma, float d, float psf_sigma_y, int sz, int iterations, float* d_params,
                         float* d_crlb, float* d_ll, int nfits, void* stream);

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
/* bamf_find_roi without a workspace argument.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The workspace is a library-owned scratch array of
 * BAMF_FIND_ROI_NOALLOC_SCRATCH_BYTES per device, held from the first call on
 * that device until the process ends; the call allocates nothing.  Stacks
 * whose bamf_find_roi_workspace_bytes() exceeds it are refused with
 * BAMF_E_INVALID_ARG (64 MiB: up to 115 frames of 256 x 256 per call).  Every
 * call on one device shares that device's scratch: calls that overlap on one
 * GPU (other streams, other threads) race and must be serialised by the
 * caller.  Arguments and results are those of bamf_find_roi. */
#define BAMF_FIND_ROI_NOALLOC_SCRATCH_BYTES (64u << 20)
int bamf_find_roi_noalloc(const float* d_stack, int xsize, int ysize, int nframes,
                          float sigma1, float sigma2, int boxsize, float minval, uint32_t* d_res_bits,
                          int64_t* d_lin_idx, int32_t* d_start_row, int32_t* d_start_col,
                          int32_t* d_frame, int64_t capacity, int64_t* d_count, void* stream);
/* ROIStack crop (findROI.m:157-160): box b is boxsize x boxsize, column-major
 * (row r, col c at r + boxsize*c), cut from frame d_frame[b] at the 1-based corner. */
int bamf_find_roi_crop(const float* d_stack, int xsize, int ysize,
                       const int32_t* d_start_row, const int32_t* d_start_col,
                       const int32_t* d_frame, int64_t n, int boxsize,
                       float* d_roi_stack, void* stream);

/* ------------------------------------------------------------------------
 * Prior construction (@RJ/findPriors.m:35-80).
 * bamf_fit_keep: @RJ/threhsSMA.m for fit type 'Basic' on gaussMLE outputs
 *   (params/crlb parameter-major [4][nfits]: Y, X, Photons, BG; SE = sqrt(CRLB)):
 *   keep[i] = X_SE, Y_SE < max_xy_se, p-value 1 - gammainc(-LL, (box^2 - 4)/2)
 *   > min_pvalue, Photons > min_photons, 0 < X, Y < box, BG > 0.  (The
 *   whole-array isreal() tests of threhsSMA.m, which reject every fit when one
 *   X or Y CRLB is negative, are the caller's: bamf_find_priors_host applies them.)
 * bamf_kde_pdf: pdf(fitdist(x, 'Kernel', 'BandWidth', h), x0 + dx*(0:ngrid-1)),
 *   the normal kernel density in double.
 * bamf_find_priors_host: the whole of findPriors.m:35-80 on host frames in
 *   MATLAB layout (Data(y, x, f)): findROI (Boxsize 8, sigma and 2 sigma,
 *   Minval = MeanPhotons/(4 pi sigma^2)/4), gaussMLE 'Basic', threhsSMA, the
 *   0.98-quantile trim (above 200 fits), the kernel density on 0:p_dp:max+1000
 *   p_dp, the second-mode cut and refit, the Gamma fit of the backgrounds, and
 *   P_EmPhotons (signal_pdf), P_Offset (offset_pdf = gampdf) and P_BgEmPhotons
 *   (bg_pdf = exppdf with mean mean(Photons)/bg_prior_range, or zeros when
 *   bg_prior is 0).  *len receives the grid length (entries beyond capacity are
 *   not written); info = {ROIs, fits kept, photons used, Gamma a, Gamma b, refit,
 *   mean photons, Minval}.
 * ------------------------------------------------------------------------ */
int bamf_fit_keep(const float* d_params, const float* d_crlb, const float* d_ll, int64_t nfits, int box,
                  float max_xy_se, float min_pvalue, float min_photons, uint8_t* d_keep, void* stream);
int bamf_kde_pdf(const float* d_samples, int64_t n, double bandwidth, double x0, double dx, int ngrid,
                 double* d_pdf, void* stream);
int bamf_find_priors_host(const float* frames, int ysize, int xsize, int nframes, float psf_sigma,
                          float mean_photons, float max_xy_se, float min_pvalue, float min_photons, float p_dp,
                          float bg_prior_range, int bg_prior, double* signal_pdf, double* bg_pdf, double* offset_pdf,
                          int64_t capacity, int64_t* len, double info[8]);

/* ------------------------------------------------------------------------
 * Frame tiling and the stitched posterior image.
 * bamf_make_subregions replaces @RJ/makeSubRegions.m:42-102 (called at
 * RJ.m:192,227,249): frames in MATLAB layout (Data(y, x, f) at
 * y + ypix (x + xpix f)) are cut into Xsub x Ysub tiles of box + 2 overlap
 * pixels (less at the frame edges), numbered frame-major then column-major
 * over positions (Ind, makeSubRegions.m:90).  Outputs (any may be NULL):
 *   d_pixels, d_pix_offset, d_xsize, d_ysize  the chain batch (bamf_roi_batch):
 *                  tile b compact, column-major, at d_pix_offset[b];
 *                  bamf_subregion_pixels() floats in all;
 *   d_frame, d_roi_ind   ROI_Struct.FrameNum and ROI_Struct.ROI_Ind (1-based);
 *   d_roi_data     ROI_Data: (box+2overlap)^2 floats per tile, zero padded;
 *   d_grid_offset, d_grid_place (with zoom)  the tiles' placement in one
 *                  stitched (zoom ypix) x (zoom xpix) column-major image, for
 *                  bamf_rjmcmc_out: the chains then add their counts straight
 *                  into it, which is removeOverlapping.m:29-54 followed by
 *                  stitchROIs.m:28-65 with ImSize(3) = 1 (RJ.m:231-234), with
 *                  the per-position grids of analyzeROIs.m:153 summed over
 *                  frames (all counts are integers, so the sum is exact).
 * Tilings the MATLAB cannot index (a tile reaching past the frame) are rejected.
 * bamf_subregion_count / _pixels return -1 for them.
 * ------------------------------------------------------------------------ */
int64_t bamf_subregion_count(int ypix, int xpix, int nframes, int box, int overlap);
int64_t bamf_subregion_pixels(int ypix, int xpix, int nframes, int box, int overlap);
int bamf_make_subregions(const float* d_data, int ypix, int xpix, int nframes, int box, int overlap,
                         float* d_pixels, int64_t* d_pix_offset, int32_t* d_xsize, int32_t* d_ysize,
                         int32_t* d_frame, int32_t* d_roi_ind, float* d_roi_data, int zoom,
                         int64_t* d_grid_offset, int32_t* d_grid_place, void* stream);

/* ------------------------------------------------------------------------
 * Reversible-jump MCMC chains, one per ROI.
 * Replaces the cRJMCMC mex entry (source/c/cRJMCMC.cpp:52, driving
 * RJMCMC::start_Chain, source/c/LibRJMCMC.cpp:685-750) for a whole batch of
 * ROIs per call.  Chain semantics, draw order and the posterior-ratio
 * formulas are the reference's; see bamf_rjmcmc_run for the modes.
 * ------------------------------------------------------------------------ */

/* The MCMC struct fields read by make_Params (cRJMCMC.cpp:14-50).  p_burnin
 * and p_trials are the raw 8-entry jump probabilities in the reference's
 * index order [move, split, merge, birth, death, gSplit, gMerge, backFore];
 * the cumulative vectors are formed inside, as make_Params does. */
typedef struct bamf_mcmc_params {
    float psf_sigma;
    int32_t n_trials;
    int32_t n_burnin;
    float i_std_fg, i_std_bg, x_std_fg, x_std_bg;
    float p_burnin[8];
    float p_trials[8];
    float rho;        /* emitters per pixel (MeanN = rho * Xsize * Ysize) */
    float bnd_out;
    int32_t grid_zoom;
    float split_std;
    float icutoff;
    float bg_std;
    float abbg_std;
    float dx;         /* grid step of the three prior tables */
    float drift_x, drift_y;
    int32_t is_backg;
} bamf_mcmc_params;

/* One record per jump call (burn-in and trials), for replay checks. */
typedef struct bamf_step_record {
    int32_t kind;      /* dispatch: 0 move 1 split 2 merge 3 birth 4 death 5 gSplit 6 gMerge 7 backFore */
    int32_t jump_type; /* the reference's recorded JumpType (1,2,3,10,4,5,6,7), 0 when the step records none */
    int32_t flags;     /* bit0 accepted, bit1 chain record, bit2 grid update, bit3 bounds test failed,
                          bit4 decided on the single-precision PR (saturated), bit5 replay: own decision differed */
    int32_t n;         /* N after the step */
    float llr, pr;     /* single-precision LLR = exp(sum) and PR, formed as the reference forms them */
    float bg;          /* BG after the step */
    float photons;     /* sum of I after the step */
    double log_pr;     /* log PR in double precision from the single-precision factors */
    int32_t draws;     /* uniforms consumed by the step */
    int32_t n_prop;    /* N of the proposed state */
    int32_t pix;       /* birth: sampled pixel index, else -1 */
    int32_t aux;       /* reserved */
} bamf_step_record;

/* A batch of ROIs (device pointers).  ROI r has xsize[r] x ysize[r] pixels
 * (each 1..32), column-major with y fastest (data[y + x*ysize], the layout of
 * LibRJMCMC.cpp:52), starting at element pix_offset[r] of d_data. */
typedef struct bamf_roi_batch {
    int32_t n_roi;
    const int64_t* d_pix_offset;
    const int32_t* d_xsize;
    const int32_t* d_ysize;
    const float* d_data;
    const float* d_scmos;        /* same layout as d_data, or NULL for zeros */
    const float* d_rho;          /* per-ROI emitters per pixel, or NULL for params.rho */
    const float* d_bg0;          /* per-ROI initial BG, or NULL for 1 (BG = 0 -> 1 as cRJMCMC.cpp:132) */
    const int32_t* d_n0;         /* per-ROI initial emitter count, or NULL for 0 */
    const int64_t* d_em0_offset; /* offsets into d_i0/d_x0/d_y0 when d_n0 is given */
    const float* d_i0;
    const float* d_x0;
    const float* d_y0;
    const uint32_t* d_key;       /* per-ROI random-stream key, or NULL for the ROI index */
    const int32_t* d_order;      /* work-queue order (a permutation of 0..n_roi-1), or NULL */
} bamf_roi_batch;

/* Replay input: each chain consumes its recorded draws and follows the
 * recorded accept decisions and birth pixels, so its state stays on the
 * recorded trajectory; its own decisions and log ratios are reported in the
 * step records for comparison. */
typedef struct bamf_replay {
    const uint32_t* d_tape;              /* raw 31-bit draws; U = (float)r / 2^31 */
    const int64_t* d_tape_offset;        /* [n_roi] */
    const int64_t* d_tape_len;           /* [n_roi] */
    const bamf_step_record* d_ref_steps; /* [n_roi][n_burnin + n_trials] */
} bamf_replay;

/* One entry of the Chain / ProposedChain records (RJMCMCChain::addEmitters,
 * LibRJMCMC.cpp:290-322): the state after a jump that reached its accept
 * test or failed its bounds test, and the proposed state of that jump. */
typedef struct bamf_chain_rec {
    int32_t n;          /* emitters in the state */
    int32_t jump_type;  /* JumpType */
    int32_t accepted;
    int32_t em_index;   /* index of its first emitter entry among this ROI's entries */
    float llr, pr;      /* the jump's LLR and PR (both records carry the same values) */
    float bg, abg, bbg;
} bamf_chain_rec;

/* Outputs (device pointers; NULL skips an output). */
typedef struct bamf_rjmcmc_out {
    int32_t* d_status;       /* [n_roi]: 0 ok, 1 emitter cap reached, 2 unsupported jump, 3 replay tape exhausted,
                                4 internal draw-window overrun (a defect; never expected),
                                5 ROI side outside 1..32 */
    int32_t* d_counters;     /* [n_roi][30]: N_* A_* N_*Fg A_*Fg N_*Bg A_*Bg for jump, split, merge, birth, death */
    float* d_final;          /* [n_roi][4 + 4*final_cap]: N, BG, ABG, BBG, then I, X, Y, Signal per emitter */
    int32_t final_cap;
    int32_t* d_n_hist;       /* [n_roi][hist_len]: post-burn-in histogram of N (last bin collects N >= hist_len-1) */
    int32_t hist_len;
    double* d_logpost;       /* [n_roi] or NULL: sum of the log PR of accepted jumps (no counterpart in the
                                reference; its double-precision log per accepted jump is formed only when
                                this output or the step records are requested) */
    int16_t* d_trace_n;      /* [n_roi][trace_len]: N at every trace_thin-th post-burn-in step */
    float* d_trace_photons;  /* [n_roi][trace_len]: total photons at the same steps */
    int32_t trace_len;
    int32_t trace_thin;
    bamf_step_record* d_steps;   /* [n_roi][n_burnin + n_trials] */
    float* d_pgrid;          /* zoomed posterior images (update_grids, LibRJMCMC.cpp:1629-1648) */
    float* d_pback;
    const int64_t* d_grid_offset; /* [n_roi] offset of ROI r's (ysize*zoom) x (xsize*zoom) grid */
    /* [n_roi][5] or NULL: where ROI r's grid lands in a larger image (the stitched
     * posterior image of removeOverlapping.m + stitchROIs.m): the image's row count
     * (the grid's column stride), then the core window x_lo, x_hi, y_lo, y_hi in grid
     * coordinates; counts outside the window are dropped, as removeOverlapping.m:38-52
     * zeroes them.  bamf_make_subregions writes it.  NULL: the ROI's own
     * (ysize*zoom) x (xsize*zoom) grid, whole. */
    const int32_t* d_grid_place;
    /* Chain records (all four arrays or none): ROI r's records at
     * [r * rec_cap ..), its emitter entries (x + DriftX, y + DriftY, I, Signal; one
     * (-100, -100, 0, -1) entry for an empty state) at [r * em_cap ..) in units
     * of 4 floats.  d_rec_count[3 r + 0..2] = records, accepted-state entries,
     * proposed-state entries (totals; entries beyond the capacities are not written). */
    bamf_chain_rec* d_acc_rec;
    bamf_chain_rec* d_prop_rec;
    float* d_acc_em;
    float* d_prop_em;
    int64_t* d_rec_count;
    int32_t rec_cap;
    int32_t em_cap;
    uint64_t* d_alg_flops;   /* [n_roi]: algorithmic floating-point operations of the chain: per jump
                                6P (likelihood ratio) + 2P + 8(X+Y) per changed emitter + 5P for a
                                background move + 4P for a birth pixel draw; 4P + N(2P + 8(X+Y)) per
                                model rebuild (P = X*Y pixels); with a sampled PSF 10P in place of
                                2P + 8(X+Y) */
} bamf_rjmcmc_out;

#define BAMF_RJ_PLAIN 0   /* counter-based Philox draws keyed by (seed, key[r], step) */
#define BAMF_RJ_REPLAY 1  /* recorded draws and decisions (bamf_replay) */
#define BAMF_RJ_LEARNED 2 /* PLAIN with births drawn from a learned proposal (bamf_rjmcmc_set_proposal) */

typedef struct bamf_rjmcmc_plan bamf_rjmcmc_plan;

/* Builds the three prior tables (Stat, LibRJMCMC.cpp:176-217) from host pdf
 * arrays on the grid 0, dx, 2dx, ... and keeps them on the current device. */
int bamf_rjmcmc_plan_create(const bamf_mcmc_params* params, const float* h_off_pdf, int off_len,
                            const float* h_sig_pdf, int sig_len, const float* h_back_pdf, int back_len,
                            bamf_rjmcmc_plan** plan);
int bamf_rjmcmc_plan_destroy(bamf_rjmcmc_plan* plan);
/* Runs every chain of the batch to completion: one wavefront per chain, ROIs
 * split into two size classes (up to 4 and up to 16 pixel slots per lane).
 * By default one persistent launch drains both classes through one work
 * queue, the large ROIs first; the other launch settings
 * (bamf_rjmcmc_set_tuning) give each class its own launch and queue, the two
 * side by side (the small class on an internal stream that `stream` joins)
 * or one after the other.  The call is asynchronous on `stream` as a whole. */
int bamf_rjmcmc_run(bamf_rjmcmc_plan* plan, const bamf_roi_batch* batch, uint64_t seed, int mode,
                    const bamf_replay* replay, const bamf_rjmcmc_out* out, void* stream);
/* Segmented runs (no counterpart in the reference, whose chain is one loop,
 * LibRJMCMC.cpp:685-750): later bamf_rjmcmc_run calls on this plan split
 * every chain into n_segments launches of consecutive steps (1..64), each
 * chain's state kept in plan-owned device memory between them (about 7 KB
 * per ROI); results equal the one-launch run bit for bit.  With use_graph
 * the whole launch sequence is captured into a HIP graph on the first run
 * and replayed by later runs with the same arguments (other than the seed);
 * a run with other arguments captures anew.  Each launch then takes a
 * bounded time, so other work on the device can interleave. */
int bamf_rjmcmc_set_segments(bamf_rjmcmc_plan* plan, int n_segments, int use_graph);

/* Launch configuration of the chain kernels (no counterpart in the
 * reference).  concurrent: 2 (the default) runs both classes in one launch
 * over one work queue, at 4 chains per SIMD (plain, segmented and learned
 * chains; runs with gSplit/gMerge or the sampled PSF then run the classes
 * side by side); 1 gives each class its own launch, the two side by side;
 * 0 one after the other.  chains_per_simd[c] (settings 0 and 1, plain
 * one-launch-per-segment runs): resident chains per SIMD that size class c's
 * kernel is compiled for (3, 4 or 5: 168, 128 or 96 VGPRs for the small
 * class, so more registers and fewer spills against more latency hiding;
 * the large class fits 4 work-groups per CU in LDS, so its 5 compiles
 * without a register limit and runs at 2).  Segmented runs and runs with
 * gSplit/gMerge use 4 chains per SIMD.  small_max_slots: the ROI tile each
 * ROI runs in, by its pixel slots per lane (64-lane slots of whole columns,
 * column height padded to a power of two): ROIs of at most this many slots
 * (1..4) run in the 4-slot class, the rest in the 16-slot class; 0 in a
 * struct passed to bamf_rjmcmc_set_tuning means the default, 4.  The
 * autotuner tries 4, 2 and 1. */
typedef struct bamf_rjmcmc_tuning {
    int32_t chains_per_simd[2];
    int32_t concurrent;
    float ms;  /* autotune: the chosen setting's time on the sample run */
    int32_t small_max_slots;
} bamf_rjmcmc_tuning;
/* Times every setting on the caller's sample batch (its ROIs with chains of
 * sample_jumps jumps, plain mode), keeps the fastest in the plan and returns
 * it.  Synchronous. */
int bamf_rjmcmc_autotune(bamf_rjmcmc_plan* plan, const bamf_roi_batch* batch, int sample_jumps,
                         bamf_rjmcmc_tuning* best, void* stream);
int bamf_rjmcmc_set_tuning(bamf_rjmcmc_plan* plan, const bamf_rjmcmc_tuning* t);
int bamf_rjmcmc_get_tuning(const bamf_rjmcmc_plan* plan, bamf_rjmcmc_tuning* t);
/* Learned birth proposal for BAMF_RJ_LEARNED runs of this plan (no
 * counterpart in the reference): d_q[pixel] is each ROI's birth-pixel
 * probability (summing to 1 over the ROI, the batch's pixel layout; see
 * bamf_net_proposal_maps) and d_mu[pixel] a log-intensity centre.  A birth
 * draws the pixel from q (PixProb = P q[pixel] in the birth ratio,
 * LibRJMCMC.cpp:1302, so a uniform q reproduces the reference's PixProb = 1)
 * and the intensity from alpha * prior + (1 - alpha) * log-uniform on
 * [mu - halfwidth, mu + halfwidth], the ratio gaining prior(I) / q(I) with the
 * signal prior normalised by its table's trapezoid integral.  With
 * probability beta a single-emitter move (LibRJMCMC.cpp:882-909) redraws
 * the intensity from the log-uniform window at the emitter's pixel instead
 * of the random walk (an independence Metropolis-Hastings kernel: the ratio
 * gains q(I | x') / q(I' | x) = I' / I, or 0 when I lies outside the window at
 * the new position); the position still walks.  The maps are caller-owned and
 * must stay valid while runs use them.  Learned runs with maps keep each
 * ROI's running sums of q in a plan-owned buffer (4 KB per ROI, allocated by
 * the first such run of a larger batch), formed when a chain starts its ROI. */
int bamf_rjmcmc_set_proposal(bamf_rjmcmc_plan* plan, const float* d_q, const float* d_mu, float alpha,
                             float halfwidth, float beta);
/* The independence intensity proposal of BAMF_RJ_LEARNED single-emitter
 * moves (taken with probability beta of bamf_rjmcmc_set_proposal; no
 * counterpart in the reference): kind 0 draws log I uniformly in the
 * network's window at the emitter's pixel; kind 1 draws I from a Gaussian
 * fitted to the emitter's conditional likelihood (the rest of the model held
 * fixed: weighted least squares, one Fisher-scoring step, sd multiplied by
 * `inflation`), scored in the Metropolis-Hastings ratio by the same fit at
 * the proposed position.  single_share: the share of moves that are
 * single-emitter moves (the reference's 0.25, LibRJMCMC.cpp:882; up to 0.5,
 * taken from the group random walk, jump type 3). */
int bamf_rjmcmc_set_intensity_proposal(bamf_rjmcmc_plan* plan, int kind, float inflation, float single_share);
/* The learned split of BAMF_RJ_LEARNED chains (no counterpart in the
 * reference; its split and merge are LibRJMCMC.cpp:972-1028 and :1111-1217).
 * With probability `share` a split places its first child (the part u1 of
 * the intensity, at the parent's index) uniformly inside a pixel drawn from
 * the occupancy map restricted to the square window of half-width `radius`
 * pixels around the parent; otherwise (u2, u3) are the reference's
 * Split_std-Gaussian offsets.  Both the split ratio and the merge ratio then
 * use the mixture density q(u2, u3) = share * q_map(pixel of the child) / Z +
 * (1 - share) * N(u2) N(u3) (Z: the map's mass in the window around the
 * parent, or around the merged emitter in the merge) in place of the
 * reference's N(u2) N(u3), so each split ratio stays the reciprocal of the
 * ratio of the merge that undoes it.  share = 0 (the default) is the
 * reference's split.  The reference's pair is not reversible (a merge may
 * pick either child as jj), so its stationary law depends on the split
 * density: share > 0, like another Split_std, moves the stationary N law
 * (tests/test_gpu_rjmcmc.py measures both); experimental, off by default. */
int bamf_rjmcmc_set_split_proposal(bamf_rjmcmc_plan* plan, float share, float radius);
/* Sampled-PSF model for later runs of this plan: SampledPSF with OptModel = 1
 * (cRJMCMC.cpp:73-74; likelihood_ratio LibRJMCMC.cpp:1650-1686 and
 * find_Birth :1693-1714 then read NumModel, :62-170, instead of Model).
 * h_psf is the PSFsize(1) x PSFsize(2) x PSFsize(3) x PSFsize(4) array in
 * MATLAB's column-major order (entry (y, x, phase_y, phase_x), PSFsize(4) =
 * the oversampling); psf_size = MCMC.PSFsize (cRJMCMC.cpp:35).  The table is
 * copied to plan-owned device memory (synchronous).  h_psf = NULL restores the
 * erf model (OptModel = 0).  Not with BAMF_RJ_LEARNED runs. */
int bamf_rjmcmc_set_sampled_psf(bamf_rjmcmc_plan* plan, const float* h_psf, const int32_t psf_size[4]);
/* The work-queue counters are reset by bamf_rjmcmc_run itself; this returns
 * how many chains the last run on the plan claimed (synchronous). */
int bamf_rjmcmc_last_claimed(bamf_rjmcmc_plan* plan, int64_t* claimed);

/* ------------------------------------------------------------------------
 * Learned birth proposal (no counterpart in the reference, whose birth
 * draws the pixel from the data residual and fixes PixProb = 1,
 * LibRJMCMC.cpp:1688-1736, 1302).
 *
 * A convolutional network (3x3 convolutions 1->32, 32->32, 32->32, 32->32
 * with ReLU, then a 1x1 head 32->2: occupancy logit and log-intensity) runs
 * on bfloat16 matrix cores with fp32 accumulation, forward and backward.
 * Parameters are one fp32 array of bamf_net_param_count() values:
 *   W1[32][9], b1[32], then for layers 2..4 W[32][288] (index tap*32 + c,
 *   tap = 3*(dh+1) + (dw+1)), b[32], then Wh[2][32], bh[2].
 * Images are batch x h x w floats (row = ROI x, column = ROI y, i.e. the
 * ROI layout data[y + x*Ysize] for an h = Xsize, w = Ysize image); the
 * network normalises each image as (sqrt(D + 3/8) - mean) / 2.
 * ------------------------------------------------------------------------ */
typedef struct bamf_net bamf_net;
int bamf_net_param_count(void);
/* h_params: initial fp32 parameters (host).  The network keeps fp32 masters,
 * Adam moments and bfloat16 operand copies on the current device. */
int bamf_net_create(int max_batch, int h, int w, const float* h_params, bamf_net** net);
int bamf_net_destroy(bamf_net* net);
int bamf_net_get_params(bamf_net* net, float* h_params);
/* d_out: batch x h x w x 2 floats (occupancy logit, log-intensity). */
int bamf_net_forward(bamf_net* net, const float* d_images, int batch, float* d_out, void* stream);
/* Loss L = mean over pixels of BCE(logit, occ) + occ * (mu - logI)^2 and its
 * gradient with respect to every parameter (d_grad: param_count floats,
 * d_loss: the two loss terms); nothing is updated. */
int bamf_net_gradients(bamf_net* net, const float* d_images, const float* d_occ, const float* d_logi, int batch,
                       float* d_grad, float* d_loss, void* stream);
/* One Adam step (beta 0.9 / 0.999) on the same loss. */
int bamf_net_train_step(bamf_net* net, const float* d_images, const float* d_occ, const float* d_logi, int batch,
                        float lr, float* d_loss, void* stream);
/* Per-pixel birth proposal of every ROI of a batch (network built for a
 * 32 x 32 canvas): d_q (ROI layout, each ROI summing to 1) =
 * (1 - eps) sigmoid(logit) / sum + eps / P, d_mu = predicted log-intensity. */
int bamf_net_proposal_maps(bamf_net* net, const bamf_roi_batch* batch, float eps, float* d_q, float* d_mu,
                           void* stream);

/* ------------------------------------------------------------------------
 * Post-passes over the Chain records of a run (bamf_rjmcmc_out d_acc_rec,
 * d_acc_em, d_rec_count with their capacities); one call for all ROIs.
 * Chain(nn), nn = 1..NLoop, is record nn - 1 for nn < NLoop, and Chain(NLoop)
 * the empty entry cRJMCMC leaves (cRJMCMC.cpp:159-174).
 * ------------------------------------------------------------------------ */
/* findMap(Chain, FB) (@RJ/findMap.m:1-112) per ROI: over nn = round(NLoop/2)..NLoop
 * the most frequent number of emitters with Signal == fb (the first of equal
 * maxima; the empty entry and N = 0 states count 0), then k-means of those
 * emitters' positions over the states with that number, and the background
 * terms over the same states.
 *   d_n_map  [n_roi]           MostFrequent;
 *   d_bg     [n_roi][4]        mean(BG), std(BG), mean(ABG), mean(BBG);
 *   d_em     [n_roi][cap][6]   per cluster: X, Y, Photons means, then their
 *                              standard deviations (SMD.X .. SMD.I_SE);
 *   d_status [n_roi]           0, or 1 when records exceeded the capacities or
 *                              MostFrequent > cap (clusters truncated to cap).
 * k-means (kmeans(Points, MostFrequent), findMap.m:79) is seeded with the
 * positions of the last state in the range with that number (the reference's
 * kmeans seeds at random), then Lloyd iterations (at most 100) until no
 * centre moves; an empty cluster keeps its centre. */
int bamf_find_map(const bamf_chain_rec* d_acc_rec, const float* d_acc_em, const int64_t* d_rec_count,
                  int32_t rec_cap, int32_t em_cap, int n_roi, int fb, int cap, int32_t* d_status, int32_t* d_n_map,
                  float* d_bg, float* d_em, void* stream);
/* The refinement summary of mcmc.m:144-166 (a fixed-dimension chain, P =
 * [1 0 0 0 0 0 0 0], mcmc.m:103): over nn = round(NLoop/2)..NLoop, per
 * emitter index the mean and standard deviation of X, Y and Photons
 * (FoundMAP.X .. FoundMAP.I_SE).  d_n [n_roi]: emitters summarised;
 * d_status 2 when the states of the range differ in N (the reference's
 * cat(1, X, x) fails then), 1 when records exceeded the capacities. */
int bamf_chain_summary(const bamf_chain_rec* d_acc_rec, const float* d_acc_em, const int64_t* d_rec_count,
                       int32_t rec_cap, int32_t em_cap, int n_roi, int cap, int32_t* d_status, int32_t* d_n,
                       float* d_em, void* stream);

/* cRJMCMC on host arrays: the argument and output layout of the mex entry
 * (source/c/cRJMCMC.cpp:52-305) for one ROI, run on the device.
 *   mcmc                the MCMC struct (make_Params fields);
 *   data                single Ysize x Xsize, column-major (data[y + x*Ysize]);
 *   n0, i0, x0, y0, bg0 E_active (N, I, X, Y, BG; BG = 0 becomes 1, :132-133);
 *   off/sig/back_pdf    OffSetPDF, SignalPDF, BackPDF with their lengths;
 *   scmos               sCMOScorr (same layout as data) or NULL for zeros;
 *   sampled_psf,        SampledPSF and MCMC.PSFsize (4 entries), read when
 *   psf_size            opt_model = 1 (see bamf_rjmcmc_set_sampled_psf), else ignored;
 *   opt_model           0 the erf PSF model, 1 the sampled PSF (NumModel);
 *   seed                the Philox stream replacing the process-global rand();
 *   pgrid, pback        (Ysize*Zoom) x (Xsize*Zoom) floats, column-major, or NULL;
 *   acc_rec/prop_rec,   Chain and ProposedChain records and their emitter
 *   acc_em/prop_em      entries (capacities rec_cap, em_cap; counts[0..2] =
 *                       records, accepted entries, proposed entries), or NULL;
 *                       the mex layer drops the last record as :174,229 do;
 *   stat                the 16 Stat fields (NTrials, Accept_* ratios, :283-305).
 * Synchronous; returns BAMF_E_CAPACITY when the chain reached the emitter cap. */
int bamf_crjmcmc(const bamf_mcmc_params* mcmc, const float* data, int xsize, int ysize, int n0, const float* i0,
                 const float* x0, const float* y0, float bg0, const float* off_pdf, int off_len,
                 const float* sig_pdf, int sig_len, const float* back_pdf, int back_len, const float* scmos,
                 const float* sampled_psf, const int32_t* psf_size, int opt_model, uint64_t seed, float* pgrid, float* pback, bamf_chain_rec* acc_rec,
                 bamf_chain_rec* prop_rec, float* acc_em, float* prop_em, int32_t rec_cap, int32_t em_cap,
                 int64_t counts[3], double stat[16]);

/* gaussMLE(Data, 'Basic', 'CCD', Sigma) on host arrays (@RJ/gaussMLE.m:64-211,
 * the fit type and camera the hot path uses):
 *   data        single (sz, sz, nfits), column-major: the layout gaussMLE.m
 *               hands to kernel_MLEFit_XYNB_ (gaussMLE.m:154);
 *   y .. bg     Results.Y, X, Photons, BG (gaussMLE.m:183-186), nfits each;
 *   *_se        Results.*_SE = sqrt(CRLB) (:187-190), or NULL;
 *   ll          Results.LogLikelihood;
 *   fit_seconds Statistics.FitTime (host wall clock), or NULL.
 * Synchronous. */
int bamf_gauss_mle_host(const float* data, int sz, int64_t nfits, float sigma, int iterations, float* y, float* x,
                        float* photons, float* bg, float* y_se, float* x_se, float* photons_se, float* bg_se,
                        float* ll, double* fit_seconds);
/* gaussMLE(Data, FitType, 'CCD', Sigma[, ZStruct]) for FitType 'Basic'
 * (BAMF_FIT_BASIC, kernel _XYNB_), 'Sigma' (BAMF_FIT_SIGMA, _XYNBS_), 'SigmaYX'
 * (BAMF_FIT_SIGMAYX, _XYNBSXSY_) or 'Z' (BAMF_FIT_Z, _XYNBZ_), host arrays in
 * and out: params and crlb are NP x nfits parameter-major (NP = 4, 5 or 6: Y,
 * X, Photons, BG[, Sigma | SigmaY, SigmaX | Z] as gaussMLE.m:183-202 maps them;
 * SE = sqrt(crlb)).  sigma is mean(Sigma), or Sigma(1) for 'Z'; zcal (only for
 * 'Z', else may be NULL) = {Ax, Ay, Bx, By, Gamma, D, Sigma(2)}. */
#define BAMF_FIT_BASIC 0
#define BAMF_FIT_SIGMA 1
#define BAMF_FIT_SIGMAYX 2
#define BAMF_FIT_Z 3
int bamf_gauss_mle_fit_host(const float* data, int sz, int64_t nfits, float sigma, int iterations, int fit_type,
                            const float* zcal, float* params, float* crlb, float* ll, double* fit_seconds);

/* The sCMOS kernels kernel_MLEFit_SCMOSXYNB_ / _SCMOSXYNBS_ / _SCMOSXYNBSXSY_ /
 * _SCMOSXYNBZ_ (cuda_gaussMLEv2.cu:559-561, 697-699, 980-982, 839-842), called by
 * gaussMLE.m:162-173 for CameraType 'SCMOS', with their argument lists: d_coords
 * holds the box corner of fit b at [2b], [2b+1] (read as the reference reads
 * d_Coords), d_gain_ratio the mapsz x mapsz map of variance/gain^2 (the
 * VarianceImage argument), pixel (ii, jj) of fit b reads
 * d_gain_ratio[(int)((d_coords[2b+1] + jj) * mapsz + d_coords[2b] + ii)]; d_x0
 * one starting z per fit.  The model adds the ratio g to model and data: cf =
 * (data - model)/(model + g), df = (data + g)/(model + g)^2; like the reference,
 * the CRLB and log-likelihood use the ratio of the box's last pixel (ii = jj =
 * sz-1) for every pixel (the reference forms the ratio only in its Newton loop).
 * Contract (the kernels do not check it; like the reference, they read the map
 * at the caller's corners): every corner satisfies 0 <= x, 0 <= y,
 * x + sz <= mapsz and y + sz <= mapsz, so each box lies on the map.
 * bamf_gauss_mle_scmos_fit_host checks this on the host and refuses the call
 * (BAMF_E_INVALID_ARG) otherwise. */
int bamf_gauss_mle_scmos_xynb(const float* d_data, const float* d_coords, const float* d_gain_ratio,
                              float psf_sigma, int sz, int mapsz, int iterations, float* d_params,
                              float* d_crlb, float* d_ll, int nfits, void* stream);
int bamf_gauss_mle_scmos_xynbs(const float* d_data, const float* d_coords, const float* d_gain_ratio,
                               float psf_sigma, int sz, int mapsz, int iterations, float* d_params,
                               float* d_crlb, float* d_ll, int nfits, void* stream);
int bamf_gauss_mle_scmos_xynbsxsy(const float* d_data, const float* d_coords, const float* d_gain_ratio,
                                  float psf_sigma, int sz, int mapsz, int iterations, float* d_params,
                                  float* d_crlb, float* d_ll, int nfits, void* stream);
int bamf_gauss_mle_scmos_xynbz(const float* d_data, const float* d_coords, const float* d_gain_ratio,
                               const float* d_x0, float psf_sigma_x, float ax, float ay, float bx, float by,
                               float gamma, float d, float psf_sigma_y, int sz, int mapsz, int iterations,
                               float* d_params, float* d_crlb, float* d_ll, int nfits, void* stream);
/* bamf_gauss_mle_fit_host for CameraType 'SCMOS': coords (2 per fit), the gain
 * map (mapsz^2) and, for 'Z', z0 (one per fit; gaussMLE.m:167 passes zeros). */
int bamf_gauss_mle_scmos_fit_host(const float* data, int sz, int64_t nfits, const float* coords,
                                  const float* gain_ratio, int mapsz, float sigma, int iterations, int fit_type,
                                  const float* zcal, const float* z0, float* params, float* crlb, float* ll,
                                  double* fit_seconds);

/* findROI(Image, Sigma1, Sigma2, Boxsize, Minval) on host arrays
 * (@RJ/findROI.m:1-163):
 *   image       single (ysize, xsize, nframes), column-major (MATLAB memory);
 *   x_corner,   SMD.XBoxCorner, SMD.YBoxCorner, SMD.FrameNum (1-based), in
 *   y_corner,   find() order; capacity entries each (any may be NULL);
 *   frame_num
 *   roi_stack   ROIStack, (boxsize, boxsize, capacity) column-major, or NULL;
 *   res         Res as 0/1 bytes, ysize*xsize*nframes, or NULL;
 *   count       receives the number of detections; when it exceeds capacity
 *               only the first capacity are written, and the caller calls
 *               again with a larger capacity.
 * The stack goes to the device in frame chunks sized to free device memory
 * (findROI.m:98-123 chunks the same way).  Synchronous. */
int bamf_find_roi_host(const float* image, int ysize, int xsize, int nframes, float sigma1, float sigma2,
                       int boxsize, float minval, int64_t capacity, int32_t* x_corner, int32_t* y_corner,
                       int32_t* frame_num, float* roi_stack, uint8_t* res, int64_t* count);

/* The branch-free erf the chain's PSF stencils use, evaluated elementwise
 * (a self-check of the device math; no counterpart in the reference). */
int bamf_math_erf(const float* d_x, float* d_y, int64_t n, void* stream);

/* ------------------------------------------------------------------------
 * Sampled PSF from isolated emitters (@RJ/findPSF.m:30-119).
 *
 * bamf_psf_shift_average replaces findPSF.m:38-56: for the first
 * min(n, 1000) ROIs (BoxSize x BoxSize, BoxSize = 8, MATLAB layout, already
 * (ROIStack - Offset) / Gain) the centred spectrum fftshift(fft2(fftshift(R /
 * 10))) with the phase ramp that moves the emitter at (d_x, d_y) (1-based
 * pixel-centre coordinates) to the common position, summed in ROI order, and
 * |fftshift(ifft2(.))| of that sum zero-padded by 12 per side into d_image
 * ((BoxSize + 24)^2 floats, column-major).  Workspace:
 * bamf_psf_workspace_bytes(n). */
int bamf_psf_shift_average(const float* d_rois, const float* d_x, const float* d_y, int64_t n, int box,
                           void* d_workspace, float* d_image, void* stream);
int64_t bamf_psf_workspace_bytes(int64_t n);
/* SampledPSFm = findPSF(ROIStack, Gain, Offset, PSFSigma, BoxSize) on host
 * arrays: the 'Sigma' fits of findPSF.m:31 (this library's gaussMLE 'Sigma',
 * 20 iterations, in place of the un-vendored SMA_Core.gaussMLE; positions
 * converted to 1-based pixel centres, the fit's + 0.5), bamf_psf_shift_average,
 * then findPSF.m:58-117
 * (edge-mean offset, peak 1, padded to 256 x 256, split into the 4 x 4 phase
 * images of 64 x 64, the middle third moved by one for nonzero phases, each
 * summing to 1).  sampled_psf: 64 x 64 x 4 x 4 floats (SampledPSF(y, x,
 * phase_y, phase_x), MATLAB order); psf_size = {64, 64, 4, 4}; fit_x / fit_y
 * (n floats each, or NULL) the positions used.  BoxSize must be 8, the size
 * for which the reference's fixed split is consistent.  Synchronous. */
int bamf_find_psf_host(const float* roi_stack, int box, int64_t n, float gain, float offset, float psf_sigma,
                       float* sampled_psf, int32_t psf_size[4], float* fit_x, float* fit_y);


/* ------------------------------------------------------------------------
 * Synthetic data from the forward model on the device (no counterpart in the
 * reference, whose examples load recorded data): the frames and ROIs the
 * benchmarks and the proposal network's training use.  Pixel (y, x) covers
 * [x, x+1) x [y, y+1) and has the expected count
 *   bg + sum_e I_e dPhi_x(e) dPhi_y(e)  (the pixel-integrated Gaussian of
 * LibRJMCMC.cpp:15-59), observed as a Poisson count (noise = 1) or as the
 * expected count itself (noise = 0).  Emitters form a Poisson point process
 * of `density` per pixel^2 with Gamma(gamma_k, gamma_theta) intensities.
 * Deterministic for a given seed.
 *
 * bamf_synth_frames: nframes frames of ysize rows of xsize contiguous pixels
 *   (the FindROI kernels' layout).  The process is drawn per 16 x 16 cell
 *   (bamf_synth_cells(ysize, xsize) cells per frame, row-major over cells):
 *   d_truth[((f * cells + c) * cap + k) * 3 + {0, 1, 2}] = x, y, I of emitter k
 *   of cell c (frame coordinates), d_count[f * cells + c] = its emitters (a
 *   Poisson count cut at cap, 1..128).  PSF tails beyond 16 pixels are
 *   dropped.
 * bamf_synth_roi_batch: n ROIs of side x side (1..32) in the ROI layout
 *   [roi][x][y] (data[y + x * side]); per ROI a density ~ U[density_lo,
 *   density_hi], N ~ Poisson(density side^2) (at most 128), uniform positions,
 *   a background ~ U[bg_lo, bg_hi].  d_occ / d_logi (may be NULL): 1 at pixel
 *   (floor(x + 1/2), floor(y + 1/2)) of each emitter and the log of the
 *   intensity summed there (the proposal network's training targets);
 *   d_nem: N per ROI; d_truth (may be NULL): per ROI 1 + 3 * 128 floats, the
 *   background, then x, y, I per emitter.
 * ------------------------------------------------------------------------ */
int64_t bamf_synth_cells(int ysize, int xsize);
int bamf_synth_frames(int nframes, int ysize, int xsize, float density, float gamma_k, float gamma_theta, float bg,
                      float sigma, uint64_t seed, int cap, int noise, float* d_frames, float* d_truth,
                      int32_t* d_count, void* stream);
int bamf_synth_roi_batch(int n, int side, float density_lo, float density_hi, float bg_lo, float bg_hi,
                         float gamma_k, float gamma_theta, float sigma, uint64_t seed, int noise, float* d_img,
                         float* d_occ, float* d_logi, int32_t* d_nem, float* d_truth, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* BAMF_GPU_H */
