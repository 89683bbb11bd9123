// Host side of the chain ABI: prior tables, argument checks, launch sizing.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "rjmcmc_dev.hpp"

namespace bamf {
namespace rj {
int launch_classify(int n_roi, const int32_t* order, const int32_t* xs, const int32_t* ys, int32_t* list0,
                    int32_t* list1, int32_t* counts, uint32_t* queues, int nq, int small_max, hipStream_t s);
int launch_set_seed(uint32_t* d, uint32_t lo, uint32_t hi, hipStream_t s);
// features: bit 0 gSplit/gMerge, bit 1 segmented, bit 2 learned (the kernel variants of rjmcmc.hip)
int launch_chain(const ChainArgs& a, int cls, int features, int tune, int grid, hipStream_t s);
int chain_occupancy_blocks(int cls, int features, int tune, int* blocks_per_cu);
int launch_erf(const float* x, float* y, int64_t n, hipStream_t s);
}  // namespace rj
}  // namespace bamf

using namespace bamf;
using namespace bamf::rj;

namespace {

// Stat::Stat (LibRJMCMC.cpp:176-217): trapezoid CDF clamped at 1, Xarray by
// repeated addition of DX, and the strictly increasing unique-CDF arrays with
// the first entry forced to 0 and the last to 1.
struct HostStat {
    std::vector<float> x, pdf, ux, uc;
    void build(const float* p, int len, float dx) {
        pdf.assign(p, p + len);
        std::vector<float> cdf(len), xm(len);
        x.assign(len, 0.0f);
        cdf[0] = 0;
        xm[0] = 0;
        for (int ii = 1; ii < len; ii++) {
            cdf[ii] = cdf[ii - 1] + (pdf[ii] + pdf[ii - 1]) * dx / 2.0f;
            if (cdf[ii] > 1.0f) cdf[ii] = 1.0f;
            xm[ii] = cdf[ii];
            x[ii] = x[ii - 1] + dx;
        }
        for (int ii = len - 2; ii >= 0; ii--)
            if (cdf[ii] == cdf[ii + 1]) xm[ii + 1] = -1;
        ux.clear();
        uc.clear();
        for (int ii = 0; ii < len; ii++)
            if (xm[ii] != -1) {
                ux.push_back(x[ii]);
                uc.push_back(cdf[ii]);
            }
        if (uc.back() != 1) uc.back() = 1;
        if (uc.front() != 0) uc.front() = 0;
    }
    // guide[k]: the last index whose uc <= k / kGuide (uc[0] = 0, strictly increasing)
    std::vector<float> guide() const {
        std::vector<float> g(kGuide);
        int i = 0;
        for (int k = 0; k < kGuide; k++) {
            const float t = (float)k / (float)kGuide;
            while (i + 1 < (int)uc.size() && uc[i + 1] <= t) i++;
            int32_t v = i;
            float f;
            std::memcpy(&f, &v, 4);
            g[k] = f;
        }
        return g;
    }
    bool exact(float dx) const {
        for (size_t i = 0; i < x.size(); i++)
            if (x[i] != (float)i * dx) return false;
        return true;
    }
};

// The correctly rounded reciprocal of b (the float nearest 1/b): the nearest
// float to 1/(double)b or a neighbour, judged by |1 - y b| exactly in double.
float rcp_rn(float b) {
    const float y = (float)(1.0 / (double)b);
    float best = y;
    double err = std::fabs(1.0 - (double)y * (double)b);
    for (float z : {std::nextafter(y, 0.0f), std::nextafter(y, 2.0f * y)}) {
        const double e = std::fabs(1.0 - (double)z * (double)b);
        if (e < err) {
            err = e;
            best = z;
        }
    }
    return best;
}

}  // namespace

struct bamf_rjmcmc_plan {
    bamf_mcmc_params params;
    float pvb[8], pvt[8];
    HostStat off, sig, back;
    float* d_tables = nullptr;
    // work-queue counters: class c of segment k at [2k + c]; then the two
    // class counts (int32) at kCountsAt and the graph seed at kSeedAt
    uint32_t* d_queue = nullptr;
    int32_t* d_lists = nullptr;   // class lists, list_cap entries each
    int64_t list_cap = 0;
    hipStream_t side = nullptr;   // second stream: the classes run concurrently
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    StatDev s_off, s_sig, s_back;
    double sig_integral = 0.0;  // trapezoid integral of the signal pdf table
    double sig_mean = 0.0, sig_sd = 0.0;  // its mean and sd (trapezoid moments of the normalised table)
    int l_ikind = 0;            // learned mode's independence intensity proposal
    float c_infl = 1.2f;
    float l_single_from = 0.75f;  // learned mode's share of single-emitter moves among moves: 1 - this
    const float* d_q = nullptr;   // learned proposal maps (caller-owned)
    float* d_qcum = nullptr;      // the maps' running sums per ROI (learned births), qcum_cap ROIs
    int64_t qcum_cap = 0;
    const float* d_mu = nullptr;
    bool proposal_set = false;    // bamf_rjmcmc_set_proposal was called (maps may be NULL)
    float l_alpha = 1.0f, l_halfwidth = 0.5f, l_beta = 0.0f;
    float l_split = 0.0f, l_split_r = 2.6f;  // learned split (bamf_rjmcmc_set_split_proposal)
    int device = 0;
    int cu = 0;
    int blocks_per_cu[kClasses][16] = {};  // [class][features] at the default register budget
    int blocks_per_cu_all[16] = {};        // [features]: the one launch over both classes
    // sampled PSF (OptModel = 1), plan-owned; null: the erf model
    float* d_psf = nullptr;
    int32_t psf_dims[4] = {0, 0, 0, 0};
    int32_t psf_numel = 0;
    int blocks_per_cu_tune[kClasses][kTuneChoices] = {};  // one-launch plain kernels per register budget
    // autotuned (or set) launch configuration of the one-launch plain kernels
    int tune[kClasses] = {kTuneDefault, kTuneDefault};
    int launch = kLaunchOne;  // how the classes are launched (kLaunch*)
    int small_max = kClassPPL[0];  // ROIs of at most this many pixel slots run in the small class
    int32_t* d_tune_status = nullptr;
    int64_t tune_cap = 0;
    // segmented runs and their graph
    int n_segments = 1;
    bool use_graph = false;
    SegState* d_seg = nullptr;
    int64_t seg_cap = 0;
    int last_nseg = 1;
    hipStream_t cap = nullptr;  // capture stream
    hipGraphExec_t gexec = nullptr;
    ChainArgs graph_key;
    int graph_nseg = 0;
};

constexpr int kCountsAt = 2 * kMaxSegments;
constexpr int kSeedAt = kCountsAt + 4;

namespace {
void plan_release(bamf_rjmcmc_plan* pl) {
    if (pl->d_tables) (void)hipFree(pl->d_tables);
    if (pl->d_queue) (void)hipFree(pl->d_queue);
    if (pl->d_lists) (void)hipFree(pl->d_lists);
    if (pl->ev_fork) (void)hipEventDestroy(pl->ev_fork);
    if (pl->ev_join) (void)hipEventDestroy(pl->ev_join);
    if (pl->side) (void)hipStreamDestroy(pl->side);
    if (pl->gexec) (void)hipGraphExecDestroy(pl->gexec);
    if (pl->cap) (void)hipStreamDestroy(pl->cap);
    if (pl->d_seg) (void)hipFree(pl->d_seg);
    if (pl->d_tune_status) (void)hipFree(pl->d_tune_status);
    if (pl->d_psf) (void)hipFree(pl->d_psf);
    if (pl->d_qcum) (void)hipFree(pl->d_qcum);
    delete pl;
}
void drop_graph(bamf_rjmcmc_plan* pl) {
    if (pl->gexec) (void)hipGraphExecDestroy(pl->gexec);
    pl->gexec = nullptr;
    pl->graph_nseg = 0;
}

// The launch sequence of one run: the classifier (which also resets the work
// queues), then either one launch per segment over both classes, or for each
// size class its segments in order on its own stream (the small-ROI class on
// the side stream beside the caller's when the classes run side by side).
int enqueue_run(bamf_rjmcmc_plan* plan, const ChainArgs& a, int nseg, bool segmented, const int tune[kClasses],
                int launch, hipStream_t s) {
    const int total = a.p.n_burnin + a.p.n_trials;
    int32_t* counts = reinterpret_cast<int32_t*>(plan->d_queue + kCountsAt);
    int rc = launch_classify(a.n_roi, a.order, a.xs, a.ys, plan->d_lists, plan->d_lists + plan->list_cap, counts,
                             plan->d_queue, 2 * nseg, plan->small_max, s);
    if (rc != BAMF_OK) return rc;
    const int cu = plan->cu > 0 ? plan->cu : 256;
    const int need = (int)ceil_div(a.n_roi, kWavesPerBlock);
    // gSplit/gMerge are dispatched on P_Burnin's cumulative vector in both
    // phases (LibRJMCMC.cpp:705-710, 737-742): reachable only when its
    // entries 5 or 6 are nonzero
    const bool gj = a.p.p_burnin[5] != 0.0f || a.p.p_burnin[6] != 0.0f;
    const int features =
        (gj ? 1 : 0) | (segmented ? 2 : 0) | (a.mode == BAMF_RJ_LEARNED ? 4 : 0) | (a.psf ? 8 : 0);
    if (launch == kLaunchOne && one_launch_features(features)) {
        const int bpc = plan->blocks_per_cu_all[features];
        int grid = (bpc > 0 ? bpc : 1) * cu;
        if (grid > need) grid = need;
        for (int k = 0; k < nseg; k++) {
            ChainArgs ac = a;
            ac.list = plan->d_lists + plan->list_cap;  // the large ROIs first
            ac.list_count = counts + 1;
            ac.list2 = plan->d_lists;
            ac.list2_count = counts;
            ac.queue = plan->d_queue + 2 * k + 1;
            ac.seg_begin = (int)((int64_t)total * k / nseg);
            ac.seg_end = k == nseg - 1 ? total : (int)((int64_t)total * (k + 1) / nseg);
            rc = launch_chain(ac, kClasses, features, kTuneDefault, grid, s);
            if (rc != BAMF_OK) return rc;
        }
        return BAMF_OK;
    }
    // features without a one-launch kernel run the classes side by side
    const bool concurrent = launch != kLaunchSequential;
    if (concurrent) {
        BAMF_HIP_TRY(hipEventRecord(plan->ev_fork, s));
        BAMF_HIP_TRY(hipStreamWaitEvent(plan->side, plan->ev_fork, 0));
    }
    for (int c = kClasses - 1; c >= 0; c--) {
        const bool tuned = features == 0;  // the register budget is tunable for the plain kernel only
        const int t = tuned ? tune[c] : kTuneDefault;
        const int bpc = tuned ? plan->blocks_per_cu_tune[c][t] : plan->blocks_per_cu[c][features];
        int grid = (bpc > 0 ? bpc : 1) * cu;
        if (grid > need) grid = need;
        for (int k = 0; k < nseg; k++) {
            ChainArgs ac = a;
            ac.list = plan->d_lists + (int64_t)c * plan->list_cap;
            ac.list_count = counts + c;
            ac.queue = plan->d_queue + 2 * k + c;
            ac.seg_begin = (int)((int64_t)total * k / nseg);
            ac.seg_end = k == nseg - 1 ? total : (int)((int64_t)total * (k + 1) / nseg);
            rc = launch_chain(ac, c, features, t, grid, (c == kClasses - 1 || !concurrent) ? s : plan->side);
            if (rc != BAMF_OK) return rc;
        }
    }
    if (concurrent) {
        BAMF_HIP_TRY(hipEventRecord(plan->ev_join, plan->side));
        BAMF_HIP_TRY(hipStreamWaitEvent(s, plan->ev_join, 0));
    }
    return BAMF_OK;
}

// Class lists for n_roi ROIs (hipFree waits for work that may still use them).
int ensure_lists(bamf_rjmcmc_plan* plan, int64_t n_roi) {
    if (plan->list_cap >= n_roi) return BAMF_OK;
    if (plan->d_lists) BAMF_HIP_TRY(hipFree(plan->d_lists));
    plan->d_lists = nullptr;
    plan->list_cap = 0;
    BAMF_HIP_TRY(hipMalloc(&plan->d_lists, sizeof(int32_t) * (size_t)kClasses * (size_t)n_roi));
    plan->list_cap = n_roi;
    drop_graph(plan);
    return BAMF_OK;
}

}  // namespace

extern "C" {

int bamf_rjmcmc_plan_create(const bamf_mcmc_params* params, const float* h_off_pdf, int off_len,
                            const float* h_sig_pdf, int sig_len, const float* h_back_pdf, int back_len,
                            bamf_rjmcmc_plan** plan) {
    BAMF_CHECK_ARG(params && plan, "null argument");
    BAMF_CHECK_ARG(h_off_pdf && h_sig_pdf && h_back_pdf, "prior tables are required");
    BAMF_CHECK_ARG(off_len >= 1 && sig_len >= 1 && back_len >= 1, "prior tables must be non-empty");
    BAMF_CHECK_ARG(params->dx > 0, "dx must be > 0");
    BAMF_CHECK_ARG(params->psf_sigma > 0, "psf_sigma must be > 0");
    BAMF_CHECK_ARG(params->n_burnin >= 0 && params->n_trials >= 0, "negative chain length");
    BAMF_CHECK_ARG(params->grid_zoom >= 1, "grid_zoom must be >= 1");
    auto* pl = new bamf_rjmcmc_plan();
    pl->params = *params;
    // make_Params (cRJMCMC.cpp:37-47): cumulative sums formed from the start for each entry
    for (int ii = 0; ii < 8; ii++) {
        float b = 0, t = 0;
        for (int mm = 0; mm <= ii; mm++) {
            b = b + params->p_burnin[mm];
            t = t + params->p_trials[mm];
        }
        pl->pvb[ii] = b;
        pl->pvt[ii] = t;
    }
    pl->off.build(h_off_pdf, off_len, params->dx);
    pl->sig.build(h_sig_pdf, sig_len, params->dx);
    pl->back.build(h_back_pdf, back_len, params->dx);
    double m1 = 0.0, m2 = 0.0;
    for (size_t i = 0; i + 1 < pl->sig.x.size(); i++) {
        const double x0 = pl->sig.x[i], x1 = pl->sig.x[i + 1], f0 = pl->sig.pdf[i], f1 = pl->sig.pdf[i + 1];
        pl->sig_integral += 0.5 * (f0 + f1) * (x1 - x0);
        m1 += 0.5 * (f0 * x0 + f1 * x1) * (x1 - x0);
        m2 += 0.5 * (f0 * x0 * x0 + f1 * x1 * x1) * (x1 - x0);
    }
    if (pl->sig_integral > 0.0) {
        pl->sig_mean = m1 / pl->sig_integral;
        pl->sig_sd = std::sqrt(std::max(m2 / pl->sig_integral - pl->sig_mean * pl->sig_mean, 1.0));
    }
    std::vector<float> flat;
    auto put = [&](const std::vector<float>& v) {
        size_t at = flat.size();
        flat.insert(flat.end(), v.begin(), v.end());
        return at;
    };
    size_t o[15];
    const HostStat* ss[3] = {&pl->off, &pl->sig, &pl->back};
    for (int k = 0; k < 3; k++) {
        o[5 * k + 0] = put(ss[k]->x);
        o[5 * k + 1] = put(ss[k]->pdf);
        o[5 * k + 2] = put(ss[k]->ux);
        o[5 * k + 3] = put(ss[k]->uc);
        o[5 * k + 4] = put(ss[k]->guide());
    }
    hipError_t e = hipGetDevice(&pl->device);
    if (e == hipSuccess) e = hipMalloc(&pl->d_tables, flat.size() * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(pl->d_tables, flat.data(), flat.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&pl->d_queue, sizeof(uint32_t) * (kSeedAt + 2));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        set_error("plan device setup failed: %s", hipGetErrorString(e));
        plan_release(pl);
        return BAMF_E_HIP;
    }
    StatDev* sd[3] = {&pl->s_off, &pl->s_sig, &pl->s_back};
    for (int k = 0; k < 3; k++) {
        sd[k]->x = pl->d_tables + o[5 * k + 0];
        sd[k]->pdf = pl->d_tables + o[5 * k + 1];
        sd[k]->ux = pl->d_tables + o[5 * k + 2];
        sd[k]->uc = pl->d_tables + o[5 * k + 3];
        sd[k]->guide = reinterpret_cast<const int32_t*>(pl->d_tables + o[5 * k + 4]);
        sd[k]->len = (int)ss[k]->x.size();
        sd[k]->nu = (int)ss[k]->ux.size();
        sd[k]->x_exact = ss[k]->exact(params->dx) ? 1 : 0;
        sd[k]->x_last = ss[k]->x.back();
        sd[k]->inv_dx = rcp_rn(params->dx);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, pl->device) == hipSuccess) pl->cu = prop.multiProcessorCount;
    for (int c = 0; c < kClasses; c++)
        for (int f = 0; f < 16; f++) {
            int bpc = 0;
            if (chain_occupancy_blocks(c, f, kTuneDefault, &bpc) == BAMF_OK) pl->blocks_per_cu[c][f] = bpc;
        }
    for (int c = 0; c < kClasses; c++)
        for (int t = 0; t < kTuneChoices; t++) {
            int bpc = 0;
            if (chain_occupancy_blocks(c, 0, t, &bpc) == BAMF_OK) pl->blocks_per_cu_tune[c][t] = bpc;
        }
    for (int f = 0; f < 16; f++) {
        int bpc = 0;
        if (one_launch_features(f) && chain_occupancy_blocks(kClasses, f, kTuneDefault, &bpc) == BAMF_OK)
            pl->blocks_per_cu_all[f] = bpc;
    }
    *plan = pl;
    return BAMF_OK;
}

int bamf_rjmcmc_plan_destroy(bamf_rjmcmc_plan* plan) {
    if (!plan) return BAMF_OK;
    (void)hipStreamSynchronize(plan->side);
    if (plan->gexec) (void)hipDeviceSynchronize();  // a replay may still be in flight on the caller's stream
    plan_release(plan);
    return BAMF_OK;
}

int bamf_math_erf(const float* d_x, float* d_y, int64_t n, void* stream) {
    BAMF_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_x && d_y, "null device pointer");
    return launch_erf(d_x, d_y, n, as_stream(stream));
}

int bamf_rjmcmc_set_proposal(bamf_rjmcmc_plan* plan, const float* d_q, const float* d_mu, float alpha,
                             float halfwidth, float beta) {
    BAMF_CHECK_ARG(plan, "null argument");
    BAMF_CHECK_ARG((d_q == nullptr) == (d_mu == nullptr), "give both maps or neither");
    BAMF_CHECK_ARG(alpha >= 0.0f && alpha <= 1.0f, "alpha must be in [0, 1]");
    BAMF_CHECK_ARG(halfwidth > 0.0f, "halfwidth must be > 0");
    BAMF_CHECK_ARG(beta >= 0.0f && beta <= 1.0f, "beta must be in [0, 1]");
    BAMF_CHECK_ARG(plan->sig_integral > 0.0, "the signal pdf table integrates to 0");
    plan->d_q = d_q;
    plan->d_mu = d_mu;
    plan->proposal_set = true;
    plan->l_alpha = alpha;
    plan->l_halfwidth = halfwidth;
    plan->l_beta = beta;
    return BAMF_OK;
}

int bamf_rjmcmc_set_intensity_proposal(bamf_rjmcmc_plan* plan, int kind, float inflation, float single_share) {
    BAMF_CHECK_ARG(plan, "null argument");
    BAMF_CHECK_ARG(kind == 0 || kind == 1, "kind must be 0 (network window) or 1 (conditional fit)");
    BAMF_CHECK_ARG(inflation > 0.0f, "inflation must be > 0");
    BAMF_CHECK_ARG(single_share >= 0.25f && single_share <= 0.5f, "single_share must be in [0.25, 0.5]");
    BAMF_CHECK_ARG(kind == 1 || plan->d_mu, "kind 0 draws from the network's window: set the maps first");
    plan->l_ikind = kind;
    plan->c_infl = inflation;
    plan->l_single_from = 1.0f - single_share;
    drop_graph(plan);
    return BAMF_OK;
}

int bamf_rjmcmc_set_split_proposal(bamf_rjmcmc_plan* plan, float share, float radius) {
    BAMF_CHECK_ARG(plan, "null argument");
    BAMF_CHECK_ARG(share >= 0.0f && share <= 1.0f, "share must be in [0, 1]");
    BAMF_CHECK_ARG(radius >= 0.5f && radius <= 16.0f, "radius must be in [0.5, 16] pixels");
    BAMF_CHECK_ARG(share == 0.0f || plan->d_q, "the learned split draws from the occupancy map: set the maps first");
    plan->l_split = share;
    plan->l_split_r = radius;
    drop_graph(plan);
    return BAMF_OK;
}

int bamf_rjmcmc_set_sampled_psf(bamf_rjmcmc_plan* plan, const float* h_psf, const int32_t psf_size[4]) {
    BAMF_CHECK_ARG(plan, "null argument");
    int64_t numel = 0;
    if (h_psf) {
        BAMF_CHECK_ARG(psf_size, "the sampled PSF needs its size");
        numel = 1;
        for (int k = 0; k < 4; k++) {
            BAMF_CHECK_ARG(psf_size[k] >= 1, "PSFsize entries must be >= 1");
            numel *= psf_size[k];
            BAMF_CHECK_ARG(numel <= 0x7fffffff, "sampled PSF too large");
        }
    }
    drop_graph(plan);
    if (plan->d_psf) BAMF_HIP_TRY(hipFree(plan->d_psf));  // waits for runs that read it
    plan->d_psf = nullptr;
    plan->psf_numel = 0;
    if (!h_psf) return BAMF_OK;
    BAMF_HIP_TRY(hipMalloc(&plan->d_psf, sizeof(float) * (size_t)numel));
    hipError_t e = hipMemcpy(plan->d_psf, h_psf, sizeof(float) * (size_t)numel, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(plan->d_psf);
        plan->d_psf = nullptr;
        set_error("sampled PSF upload failed: %s", hipGetErrorString(e));
        return BAMF_E_HIP;
    }
    std::memcpy(plan->psf_dims, psf_size, sizeof(plan->psf_dims));
    plan->psf_numel = (int32_t)numel;
    return BAMF_OK;
}

int bamf_rjmcmc_last_claimed(bamf_rjmcmc_plan* plan, int64_t* claimed) {
    BAMF_CHECK_ARG(plan && claimed, "null argument");
    uint32_t q[kCountsAt + 2];
    BAMF_HIP_TRY(hipMemcpy(q, plan->d_queue, sizeof(q), hipMemcpyDeviceToHost));
    int64_t n = 0;
    const uint32_t* last = q + 2 * (plan->last_nseg - 1);  // the last segment's queues
    for (int c = 0; c < kClasses; c++) n += last[c] < q[kCountsAt + c] ? last[c] : q[kCountsAt + c];
    *claimed = n;
    return BAMF_OK;
}

int bamf_rjmcmc_run(bamf_rjmcmc_plan* plan, const bamf_roi_batch* batch, uint64_t seed, int mode,
                    const bamf_replay* replay, const bamf_rjmcmc_out* out, void* stream) {
    BAMF_CHECK_ARG(plan && batch && out, "null argument");
    BAMF_CHECK_ARG(mode == BAMF_RJ_PLAIN || mode == BAMF_RJ_REPLAY || mode == BAMF_RJ_LEARNED, "unknown mode %d",
                   mode);
    if (mode == BAMF_RJ_LEARNED) {
        BAMF_CHECK_ARG(plan->proposal_set, "learned mode needs bamf_rjmcmc_set_proposal first");
        BAMF_CHECK_ARG(plan->params.is_backg == 0, "learned mode supports IsBackg = 0 only");
        BAMF_CHECK_ARG(!plan->d_psf, "learned mode supports the erf PSF model only");
    }
    BAMF_CHECK_ARG(batch->n_roi >= 0, "n_roi must be >= 0");
    if (batch->n_roi == 0) return BAMF_OK;
    BAMF_CHECK_ARG(batch->d_pix_offset && batch->d_xsize && batch->d_ysize && batch->d_data,
                   "batch pixel arrays are required");
    BAMF_CHECK_ARG(!batch->d_n0 || (batch->d_em0_offset && batch->d_i0 && batch->d_x0 && batch->d_y0),
                   "initial emitters need offsets and I/X/Y");
    if (mode == BAMF_RJ_REPLAY)
        BAMF_CHECK_ARG(replay && replay->d_tape && replay->d_tape_offset && replay->d_tape_len && replay->d_ref_steps,
                       "replay mode needs the tape and the recorded steps");
    BAMF_CHECK_ARG(!out->d_final || out->final_cap >= 0, "final_cap must be >= 0");
    BAMF_CHECK_ARG(!out->d_n_hist || out->hist_len >= 1, "hist_len must be >= 1");
    BAMF_CHECK_ARG(!out->d_trace_n || (out->d_trace_photons && out->trace_len >= 0 && out->trace_thin >= 1),
                   "trace needs photons, trace_len >= 0 and trace_thin >= 1");
    BAMF_CHECK_ARG(!out->d_pgrid || (out->d_pback && out->d_grid_offset), "grids need pback and offsets");
    const bamf_mcmc_params& p = plan->params;
    ChainArgs a;
    std::memset(&a, 0, sizeof(a));
    a.p = p;
    std::memcpy(a.pvb, plan->pvb, sizeof(a.pvb));
    std::memcpy(a.pvt, plan->pvt, sizeof(a.pvt));
    a.off = plan->s_off;
    a.sig = plan->s_sig;
    a.back = plan->s_back;
    a.mode = mode;
    a.seed_lo = (uint32_t)seed;
    a.seed_hi = (uint32_t)(seed >> 32);
    a.n_roi = batch->n_roi;
    a.pix_off = batch->d_pix_offset;
    a.xs = batch->d_xsize;
    a.ys = batch->d_ysize;
    a.data = batch->d_data;
    a.scmos = batch->d_scmos;
    a.rho = batch->d_rho;
    a.bg0 = batch->d_bg0;
    a.n0 = batch->d_n0;
    a.em0_off = batch->d_em0_offset;
    a.i0 = batch->d_i0;
    a.x0 = batch->d_x0;
    a.y0 = batch->d_y0;
    a.key = batch->d_key;
    a.order = batch->d_order;
    if (mode == BAMF_RJ_REPLAY) {
        a.tape = replay->d_tape;
        a.tape_off = replay->d_tape_offset;
        a.tape_len = replay->d_tape_len;
        a.ref_steps = replay->d_ref_steps;
    }
    a.status = out->d_status;
    a.counters = out->d_counters;
    a.fin = out->d_final;
    a.fin_cap = out->final_cap;
    a.hist = out->d_n_hist;
    a.hist_len = out->hist_len;
    a.logpost = out->d_logpost;
    a.trace_n = out->d_trace_n;
    a.trace_ph = out->d_trace_photons;
    a.trace_len = out->trace_len;
    a.trace_thin = out->trace_thin > 0 ? out->trace_thin : 1;
    a.steps = out->d_steps;
    a.pgrid = out->d_pgrid;
    a.pback = out->d_pback;
    a.grid_off = out->d_grid_offset;
    a.grid_place = out->d_grid_place;
    a.alg_flops = out->d_alg_flops;
    BAMF_CHECK_ARG((!out->d_acc_rec && !out->d_prop_rec && !out->d_acc_em && !out->d_prop_em) ||
                       (out->d_acc_rec && out->d_prop_rec && out->d_acc_em && out->d_prop_em && out->d_rec_count &&
                        out->rec_cap >= 0 && out->em_cap >= 0),
                   "chain records need all four arrays, the counts and non-negative capacities");
    a.acc_rec = out->d_acc_rec;
    a.prop_rec = out->d_prop_rec;
    a.acc_em = out->d_acc_em;
    a.prop_em = out->d_prop_em;
    a.rec_count = out->d_rec_count;
    a.rec_cap = out->rec_cap;
    a.em_cap = out->em_cap;
    if (mode == BAMF_RJ_LEARNED) {
        BAMF_CHECK_ARG(plan->d_mu || plan->l_ikind == 1 || plan->l_beta == 0.0f,
                       "the network-window intensity proposal needs the maps (bamf_rjmcmc_set_proposal)");
        a.prop_map = plan->d_q;
        a.prop_mu = plan->d_mu;
        a.l_alpha = plan->l_alpha;
        a.l_halfwidth = plan->l_halfwidth;
        a.l_beta = plan->l_beta;
        a.l_zinv = (float)(1.0 / plan->sig_integral);
        a.l_ikind = plan->l_ikind;
        a.c_infl = plan->c_infl;
        a.c_mean = (float)plan->sig_mean;
        a.c_sd = (float)plan->sig_sd;
        a.l_single_from = plan->l_single_from;
        a.l_split = plan->d_q ? plan->l_split : 0.0f;
        a.l_split_r = plan->l_split_r;
        if (plan->d_q && plan->qcum_cap < batch->n_roi) {
            if (plan->d_qcum) BAMF_HIP_TRY(hipFree(plan->d_qcum));
            plan->d_qcum = nullptr;
            plan->qcum_cap = 0;
            BAMF_HIP_TRY(hipMalloc(&plan->d_qcum, sizeof(float) * 64 * kMaxPPL * (size_t)batch->n_roi));
            plan->qcum_cap = batch->n_roi;
        }
        a.qcum = plan->d_q ? plan->d_qcum : nullptr;
    }
    if (plan->d_psf) {
        a.psf = plan->d_psf;
        std::memcpy(a.psf_dims, plan->psf_dims, sizeof(a.psf_dims));
        a.psf_numel = plan->psf_numel;
    }
    // rebuilding every 16 accepted jumps bounds the float32 drift of the
    // incrementally updated model to a few roundings: a death's LLR then stays
    // within twice the error of a full float32 re-evaluation (about 2 % of
    // chain time; every 128 jumps leaves 3.7x on the worst steps)
    a.recompute_every = 16;
    hipStream_t s = as_stream(stream);
    const int total = p.n_burnin + p.n_trials;
    const int nseg = std::max(1, std::min(plan->n_segments, total));
    int rc0 = ensure_lists(plan, batch->n_roi);
    if (rc0 != BAMF_OK) return rc0;
    if (nseg > 1 && plan->seg_cap < batch->n_roi) {
        if (plan->d_seg) BAMF_HIP_TRY(hipFree(plan->d_seg));
        plan->d_seg = nullptr;
        plan->seg_cap = 0;
        BAMF_HIP_TRY(hipMalloc(&plan->d_seg, sizeof(SegState) * (size_t)batch->n_roi));
        plan->seg_cap = batch->n_roi;
        drop_graph(plan);
    }
    a.seg_state = nseg > 1 ? plan->d_seg : nullptr;
    plan->last_nseg = nseg;
    if (!plan->use_graph) return enqueue_run(plan, a, nseg, nseg > 1, plan->tune, plan->launch, s);
    // graph replay: the launch sequence is captured once per argument set and
    // replayed; the seed travels through device memory, written ahead of it
    a.seed_lo = a.seed_hi = 0;
    a.seed_dev = plan->d_queue + kSeedAt;
    if (!plan->gexec || plan->graph_nseg != nseg || std::memcmp(&plan->graph_key, &a, sizeof(a)) != 0) {
        drop_graph(plan);
        BAMF_HIP_TRY(hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_run(plan, a, nseg, true, plan->tune, plan->launch, plan->cap);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(plan->cap, &g);
        if (rc != BAMF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        BAMF_HIP_TRY(e);
        const hipError_t ei = hipGraphInstantiate(&plan->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            plan->gexec = nullptr;
            set_error("graph instantiation failed: %s", hipGetErrorString(ei));
            return BAMF_E_HIP;
        }
        plan->graph_key = a;
        plan->graph_nseg = nseg;
    }
    const int rc = launch_set_seed(plan->d_queue + kSeedAt, (uint32_t)seed, (uint32_t)(seed >> 32), s);
    if (rc != BAMF_OK) return rc;
    BAMF_HIP_TRY(hipGraphLaunch(plan->gexec, s));
    return BAMF_OK;
}

int bamf_rjmcmc_get_tuning(const bamf_rjmcmc_plan* plan, bamf_rjmcmc_tuning* t) {
    BAMF_CHECK_ARG(plan && t, "null argument");
    for (int c = 0; c < kClasses; c++) t->chains_per_simd[c] = kTuneMinBlocks[plan->tune[c]] * kWavesPerBlock / 4;
    t->concurrent = plan->launch;
    t->ms = 0.0f;
    t->small_max_slots = plan->small_max;
    return BAMF_OK;
}

int bamf_rjmcmc_set_tuning(bamf_rjmcmc_plan* plan, const bamf_rjmcmc_tuning* t) {
    BAMF_CHECK_ARG(plan && t, "null argument");
    int idx[kClasses];
    for (int c = 0; c < kClasses; c++) {
        idx[c] = -1;
        for (int k = 0; k < kTuneChoices; k++)
            if (kTuneMinBlocks[k] * kWavesPerBlock / 4 == t->chains_per_simd[c]) idx[c] = k;
        BAMF_CHECK_ARG(idx[c] >= 0, "chains_per_simd[%d] = %d is not one of the compiled budgets (3, 4, 5)", c,
                       t->chains_per_simd[c]);
    }
    BAMF_CHECK_ARG(t->concurrent >= kLaunchSequential && t->concurrent <= kLaunchOne, "concurrent must be 0, 1 or 2");
    BAMF_CHECK_ARG(t->concurrent != kLaunchOne || (idx[0] == kTuneDefault && idx[1] == kTuneDefault),
                   "the one launch over both classes runs 4 chains per SIMD");
    BAMF_CHECK_ARG(t->small_max_slots >= 0 && t->small_max_slots <= kClassPPL[0],
                   "small_max_slots must be in 0..%d (0: the default)", kClassPPL[0]);
    for (int c = 0; c < kClasses; c++) plan->tune[c] = idx[c];
    plan->launch = t->concurrent;
    plan->small_max = t->small_max_slots > 0 ? t->small_max_slots : kClassPPL[0];
    drop_graph(plan);
    return BAMF_OK;
}

int bamf_rjmcmc_autotune(bamf_rjmcmc_plan* plan, const bamf_roi_batch* batch, int sample_jumps,
                         bamf_rjmcmc_tuning* best, void* stream) {
    BAMF_CHECK_ARG(plan && batch && best, "null argument");
    BAMF_CHECK_ARG(sample_jumps >= 2, "sample_jumps must be >= 2");
    BAMF_CHECK_ARG(batch->n_roi >= 1, "the sample batch is empty");
    hipStream_t s = as_stream(stream);
    if (plan->tune_cap < batch->n_roi) {
        if (plan->d_tune_status) BAMF_HIP_TRY(hipFree(plan->d_tune_status));
        plan->d_tune_status = nullptr;
        plan->tune_cap = 0;
        BAMF_HIP_TRY(hipMalloc(&plan->d_tune_status, sizeof(int32_t) * (size_t)batch->n_roi));
        plan->tune_cap = batch->n_roi;
    }
    // the sample runs use this plan with short chains, one launch, no graph
    const bamf_mcmc_params saved = plan->params;
    const int saved_tune[kClasses] = {plan->tune[0], plan->tune[1]};
    const int saved_launch = plan->launch, saved_small = plan->small_max;
    const bool saved_graph = plan->use_graph;
    const int saved_nseg = plan->n_segments;
    plan->params.n_burnin = sample_jumps / 2;
    plan->params.n_trials = sample_jumps - sample_jumps / 2;
    plan->use_graph = false;
    plan->n_segments = 1;
    bamf_rjmcmc_out out;
    std::memset(&out, 0, sizeof(out));
    out.d_status = plan->d_tune_status;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = BAMF_OK;
    float best_ms = -1.0f;
    int best_t[kClasses] = {kTuneDefault, kTuneDefault};
    int best_launch = kLaunchOne, best_small = kClassPPL[0];
    // the ROI tiles: the small class takes ROIs of at most 4, 2 or 1 pixel slots
    constexpr int kSmallMax[3] = {kClassPPL[0], 2, 1};
    // group jumps and segments use the default budget: only the concurrency is tunable then
    const bool gj = saved.p_burnin[5] != 0.0f || saved.p_burnin[6] != 0.0f;
    const int nt = gj ? 1 : kTuneChoices;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("autotune: hipEventCreate failed");
        rc = BAMF_E_HIP;
    }
    for (int lm = kLaunchOne; lm >= kLaunchSequential && rc == BAMF_OK; lm--)
        for (int sm = 0; sm < 3 && rc == BAMF_OK; sm++)
        for (int t1 = 0; t1 < nt && rc == BAMF_OK; t1++)
            for (int t0 = 0; t0 < nt && rc == BAMF_OK; t0++) {
                // the one launch over both classes has the default budget only
                if (lm == kLaunchOne && (t0 != kTuneDefault || t1 != kTuneDefault) && !gj) continue;
                plan->tune[0] = gj ? kTuneDefault : t0;
                plan->tune[1] = gj ? kTuneDefault : t1;
                plan->launch = lm;
                plan->small_max = kSmallMax[sm];
                float ms = 0.0f;
                for (int rep = 0; rep < 2 && rc == BAMF_OK; rep++) {  // the first run loads the code objects
                    if (hipEventRecord(e0, s) != hipSuccess) rc = BAMF_E_HIP;
                    if (rc == BAMF_OK) rc = bamf_rjmcmc_run(plan, batch, 0x5eed, BAMF_RJ_PLAIN, nullptr, &out, s);
                    if (rc == BAMF_OK && hipEventRecord(e1, s) != hipSuccess) rc = BAMF_E_HIP;
                    if (rc == BAMF_OK && hipEventSynchronize(e1) != hipSuccess) rc = BAMF_E_HIP;
                    if (rc == BAMF_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = BAMF_E_HIP;
                }
                if (rc == BAMF_OK && (best_ms < 0.0f || ms < best_ms)) {
                    best_ms = ms;
                    best_t[0] = plan->tune[0];
                    best_t[1] = plan->tune[1];
                    best_launch = plan->launch;
                    best_small = plan->small_max;
                }
            }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    plan->params = saved;
    plan->use_graph = saved_graph;
    plan->n_segments = saved_nseg;
    if (rc != BAMF_OK) {
        plan->tune[0] = saved_tune[0];
        plan->tune[1] = saved_tune[1];
        plan->launch = saved_launch;
        plan->small_max = saved_small;
        if (rc == BAMF_E_HIP) set_error("autotune: a sample run failed on the device");
        return rc;
    }
    plan->tune[0] = best_t[0];
    plan->tune[1] = best_t[1];
    plan->launch = best_launch;
    plan->small_max = best_small;
    drop_graph(plan);
    rc = bamf_rjmcmc_get_tuning(plan, best);
    best->ms = best_ms;
    return rc;
}

int bamf_rjmcmc_set_segments(bamf_rjmcmc_plan* plan, int n_segments, int use_graph) {
    BAMF_CHECK_ARG(plan, "null argument");
    BAMF_CHECK_ARG(n_segments >= 1 && n_segments <= kMaxSegments, "n_segments must be in 1..%d", kMaxSegments);
    if (use_graph && !plan->cap) BAMF_HIP_TRY(hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking));
    drop_graph(plan);
    plan->n_segments = n_segments;
    plan->use_graph = use_graph != 0;
    return BAMF_OK;
}

}  // extern "C"
