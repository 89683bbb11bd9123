// Device-side argument block of the chain kernel (shared by rjmcmc.hip and
// the host launcher in rjmcmc_host.cpp).  Not part of the public ABI.
#pragma once
#include <cstdint>

#include "../../include/bamf_gpu.h"

namespace bamf {
namespace rj {

constexpr int kNMax = 128;       // emitters per chain (LDS state capacity)
constexpr int kWavesPerBlock = 4;
constexpr int kMaxPPL = 16;      // pixels per lane: ROIs up to 32 x 32
constexpr int kMaxSide = 32;
// ROI size classes: each has its own kernel instantiation with PPL pixel
// slots per lane (fewer registers, more resident chains for small ROIs).
constexpr int kClasses = 2;
constexpr int kClassPPL[kClasses] = {4, 16};
constexpr int kClassMinBlocks[kClasses] = {4, 4};  // work-groups per CU the register budget targets
// Register budgets the autotuner chooses from for the one-launch plain
// kernels (work-groups of kWavesPerBlock waves per CU = chains per SIMD):
// 168, 128 and 96 VGPRs.  The large class fits 4 work-groups per CU in LDS
// (39 KB each), not 5.
constexpr int kTuneChoices = 3;
constexpr int kTuneMinBlocks[kTuneChoices] = {3, 4, 5};
constexpr int kTuneDefault = 1;
// Feature sets (bit 0 gSplit/gMerge, bit 1 segmented, bit 2 learned, bit 3
// sampled PSF) with a one-launch kernel over both classes: plain, segmented
// and learned.
constexpr bool one_launch_features(int f) { return f == 0 || f == 2 || f == 4; }
// How a run launches the classes (bamf_rjmcmc_tuning.concurrent): one after
// the other, side by side on two streams, or one launch over both.
constexpr int kLaunchSequential = 0, kLaunchSideBySide = 1, kLaunchOne = 2;

// One prior table (Stat, LibRJMCMC.cpp:176-217) on the device.
struct StatDev {
    const float* x;      // Xarray (0, dx, 2dx, ... by repeated addition)
    const float* pdf;    // PDFarray
    const float* ux;     // UniqXarray
    const float* uc;     // UniqCDFarray
    const int32_t* guide;  // guide[k] = last index with uc <= k / kGuide (genRand search start)
    int32_t len;
    int32_t nu;
    int32_t x_exact;     // 1 when x[i] == (float)i * dx for every i (O(1) findPDF index)
    float x_last;        // x[len - 1]
    float inv_dx;        // 1 / dx, correctly rounded
};
constexpr int kGuide = 256;

// Segmented runs: a launch runs steps [seg_begin, seg_end) of every chain;
// between launches each chain's state waits in HBM.  Everything a later
// step reads is kept, so a segmented run equals the one-launch run bit for bit.
constexpr int kMaxSegments = 64;
struct SegState {
    double logpost;
    uint64_t aflops;
    int64_t tcur;
    int64_t rec[3];
    int32_t status, n, e_ok, nacc;
    float bg, abg, bbg, pad;
    int32_t cnt[32];
    int32_t hist[kNMax + 1];
    int32_t pad2[3];
    float I[kNMax], X[kNMax], Y[kNMax];
    int32_t S[kNMax];
    float m[64 * kMaxPPL];  // the model, in the lane-slot layout
};

struct ChainArgs {
    bamf_mcmc_params p;
    float pvb[8];  // cumulative P_Burnin (make_Params)
    float pvt[8];  // cumulative P_Trials
    StatDev off, sig, back;
    int32_t mode;
    uint32_t seed_lo, seed_hi;
    // batch
    int32_t n_roi;
    const int64_t* pix_off;
    const int32_t* xs;
    const int32_t* ys;
    const float* data;
    const float* scmos;
    const float* rho;
    const float* bg0;
    const int32_t* n0;
    const int64_t* em0_off;
    const float* i0;
    const float* x0;
    const float* y0;
    const uint32_t* key;
    const int32_t* order;       // caller's ROI order (read by the classifier only)
    const int32_t* list;        // this launch's class list of ROI indices
    const int32_t* list_count;  // its length (device)
    const int32_t* list2;       // one launch over both classes: the small-ROI list, drained after `list`
    const int32_t* list2_count;
    uint32_t* queue;            // work-queue counter (reset by the classifier)
    // replay
    const uint32_t* tape;
    const int64_t* tape_off;
    const int64_t* tape_len;
    const bamf_step_record* ref_steps;
    // learned birth proposal: per-ROI per-pixel probabilities and log-intensity
    // centres (same layout as data), the prior weight of the intensity mixture,
    // the log-uniform half-width and 1 / (integral of the signal pdf table)
    const float* prop_map;
    const float* prop_mu;
    // learned births: each ROI's running sums of prop_map in pixel order, in
    // the lane-slot layout ([roi][64 * kMaxPPL], -inf where the map is not
    // positive), written by the chain when it starts the ROI
    float* qcum;
    float l_alpha, l_halfwidth, l_zinv;
    float l_beta;  // probability of the independence intensity proposal in a single-emitter move
    int32_t l_ikind;  // that proposal: 0 the network's log-uniform window, 1 the conditional-likelihood fit
    float c_infl;     // sd inflation of the conditional fit
    float c_mean, c_sd;  // signal prior mean and sd (the fit's fallback)
    float l_single_from;  // learned mode: single-emitter moves take Ub >= this (the reference: 0.75)
    float l_split;        // learned mode: probability that a split places its first child from the map
    float l_split_r;      // half-width (px) of the square window around the parent the map draw uses
    // outputs
    int32_t* status;
    int32_t* counters;
    float* fin;
    int32_t fin_cap;
    int32_t* hist;
    int32_t hist_len;
    double* logpost;
    int16_t* trace_n;
    float* trace_ph;
    int32_t trace_len;
    int32_t trace_thin;
    bamf_step_record* steps;
    float* pgrid;
    float* pback;
    const int64_t* grid_off;
    const int32_t* grid_place;  // [n_roi][5]: column stride, x_lo, x_hi, y_lo, y_hi, or null
    bamf_chain_rec* acc_rec;
    bamf_chain_rec* prop_rec;
    float* acc_em;
    float* prop_em;
    int64_t* rec_count;
    int32_t rec_cap, em_cap;
    uint64_t* alg_flops;
    int32_t recompute_every;  // full model rebuild after this many accepted jumps
    // sampled-PSF model (OptModel = 1; kernels with the kNUM feature): the
    // PSFsize(1) x .. x PSFsize(4) table in column-major order and its size
    const float* psf;
    int32_t psf_dims[4];
    int32_t psf_numel;
    // segments (seg_begin = 0 and seg_end >= total: the whole chain in one launch)
    int32_t seg_begin, seg_end;
    SegState* seg_state;       // one per ROI; read when seg_begin > 0, written when seg_end < total
    const uint32_t* seed_dev;  // when set, {seed_lo, seed_hi} are read from here (graph replays)
};

}  // namespace rj
}  // namespace bamf
