// Post-passes over the Chain records of a run, on the device:
//   k_find_map       findMap (@RJ/findMap.m:1-112): the most frequent number
//                    of Signal == FB emitters over the second half of the
//                    chain, k-means of their positions in the states with
//                    that number, per-cluster means and standard errors, and
//                    the background terms' means;
//   k_chain_summary  the refinement summary of mcmc.m:144-166: per-emitter
//                    means and standard deviations of X, Y and photons over
//                    the second half of a fixed-dimension chain.
// One wavefront per ROI.  Chain(nn), nn = 1..NLoop, is record nn - 1 for
// nn < NLoop; Chain(NLoop) is the empty entry cRJMCMC leaves
// (cRJMCMC.cpp:159-174), which the loops over round(NLoop/2):NLoop visit.
#include <cfloat>

#include "common.hpp"

// squared distances and statistics exactly as written (no fused multiply-add)
#pragma clang fp contract(off)

namespace bamf {
namespace {

constexpr int kWaves = 4;
constexpr int kMaxN = 128;        // emitters per state (the chain's capacity)
constexpr int kMaxIter = 100;     // k-means iterations (kmeans' MaxIter default)
constexpr double kFix = 16777216.0;  // 2^24: fixed-point scale of the centroid sums

struct MapLds {
    int hist[kMaxN + 2];
    float cx[kMaxN], cy[kMaxN];
    unsigned long long sx[kMaxN], sy[kMaxN];
    int cnt[kMaxN];
    double st[6][kMaxN];  // sums of X, Y, I and of their squares per cluster
    int changed;
};

// Makes this wavefront's LDS writes visible to its other lanes.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// Emitters of record k with Signal == fb; 0 for an N = 0 placeholder state
// (Signal -1: `if Chain(nn).Signal ~= -1` fails) and for the empty entry.
__device__ __forceinline__ int count_fb(const bamf_chain_rec* rec, const float* em, int k, float fb) {
    const bamf_chain_rec r = rec[k];
    if (r.n < 1) return 0;
    int c = 0;
    for (int i = 0; i < r.n; i++) c += em[4 * ((int64_t)r.em_index + i) + 3] == fb ? 1 : 0;
    return c;
}

__device__ __forceinline__ int nearest(const MapLds& L, int k, float x, float y) {
    int best = 0;
    float bd = FLT_MAX;
    for (int c = 0; c < k; c++) {
        const float dx = x - L.cx[c], dy = y - L.cy[c];
        const float d = dx * dx + dy * dy;
        if (d < bd) {
            bd = d;
            best = c;
        }
    }
    return best;
}

__global__ __launch_bounds__(64 * kWaves) void k_find_map(int n_roi, const bamf_chain_rec* acc_rec,
                                                          const float* acc_em, const int64_t* rec_count,
                                                          int32_t rec_cap, int32_t em_cap, float fb, int cap,
                                                          int32_t* status, int32_t* n_map, float* bgo, float* emo) {
    __shared__ MapLds lds[kWaves];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    MapLds& L = lds[w];
    for (int r = blockIdx.x * kWaves + w; r < n_roi; r += gridDim.x * kWaves) {
        const int64_t nloop = rec_count[3 * (int64_t)r];
        const bamf_chain_rec* rec = acc_rec + (int64_t)r * rec_cap;
        const float* em = acc_em + 4 * (int64_t)r * em_cap;
        float* eo = emo + (int64_t)r * cap * 6;
        if (nloop > rec_cap || rec_count[3 * (int64_t)r + 1] > em_cap || nloop < 1) {
            if (lane == 0) {
                status[r] = nloop < 1 ? 0 : 1;
                n_map[r] = 0;
                for (int j = 0; j < 4; j++) bgo[4 * (int64_t)r + j] = __builtin_nanf("");
            }
            continue;
        }
        const int it = (int)nloop;
        const int lo = (it + 1) / 2;  // round(Iter/2), 1-based
        // ---- histogram of the fb count over Chain(lo..it) ----
        for (int i = lane; i < kMaxN + 2; i += 64) L.hist[i] = 0;
        wsync();
        int mx = 0;
        for (int nn = lo + lane; nn <= it; nn += 64) {
            const int c = nn < it ? count_fb(rec, em, nn - 1, fb) : 0;
            atomicAdd(&L.hist[min(c, kMaxN + 1)], 1);
            mx = max(mx, c);
        }
        mx = wave_max_i(mx);
        wsync();
        // most frequent count, the first of equal maxima (find(...)(1) - 1)
        int mf = 0;
        if (lane == 0) {
            int best = -1;
            for (int c = 0; c <= mx; c++)
                if (L.hist[c] > best) {
                    best = L.hist[c];
                    mf = c;
                }
        }
        mf = __shfl(mf, 0);
        const int k = min(mf, cap);
        // ---- background terms over the states with that count ----
        double sb = 0.0, sb2 = 0.0, sa = 0.0, sbb = 0.0, nb = 0.0;
        int last = -1;
        for (int nn = lo + lane; nn < it; nn += 64) {
            if (count_fb(rec, em, nn - 1, fb) != mf) continue;
            const bamf_chain_rec q = rec[nn - 1];
            sb += (double)q.bg;
            sb2 += (double)q.bg * (double)q.bg;
            sa += (double)q.abg;
            sbb += (double)q.bbg;
            nb += 1.0;
            last = max(last, nn);
        }
        sb = wave_sum_d(sb);
        sb2 = wave_sum_d(sb2);
        sa = wave_sum_d(sa);
        sbb = wave_sum_d(sbb);
        nb = wave_sum_d(nb);
        last = wave_max_i(last);
        if (lane == 0) {
            n_map[r] = mf;
            status[r] = mf > cap ? 1 : 0;
            const double m = sb / nb;
            bgo[4 * (int64_t)r + 0] = nb > 0 ? (float)m : __builtin_nanf("");
            bgo[4 * (int64_t)r + 1] = nb > 1 ? (float)sqrt(fmax((sb2 - nb * m * m) / (nb - 1.0), 0.0)) : (nb > 0 ? 0.0f : __builtin_nanf(""));
            bgo[4 * (int64_t)r + 2] = nb > 0 ? (float)(sa / nb) : __builtin_nanf("");
            bgo[4 * (int64_t)r + 3] = nb > 0 ? (float)(sbb / nb) : __builtin_nanf("");
        }
        if (k < 1 || last < 0) continue;
        // ---- k-means: centres from the last such state's emitters, then Lloyd ----
        {
            const bamf_chain_rec q = rec[last - 1];
            if (lane == 0) {
                int c = 0;
                for (int i = 0; i < q.n && c < k; i++) {
                    const float* e = em + 4 * ((int64_t)q.em_index + i);
                    if (e[3] == fb) {
                        L.cx[c] = e[0];
                        L.cy[c] = e[1];
                        c++;
                    }
                }
            }
        }
        wsync();
        for (int iter = 0; iter < kMaxIter; iter++) {
            for (int c = lane; c < k; c += 64) {
                L.sx[c] = 0ull;
                L.sy[c] = 0ull;
                L.cnt[c] = 0;
            }
            if (lane == 0) L.changed = 0;
            wsync();
            for (int nn = lo + lane; nn < it; nn += 64) {
                if (count_fb(rec, em, nn - 1, fb) != mf) continue;
                const bamf_chain_rec q = rec[nn - 1];
                for (int i = 0; i < q.n; i++) {
                    const float* e = em + 4 * ((int64_t)q.em_index + i);
                    if (e[3] != fb) continue;
                    const int c = nearest(L, k, e[0], e[1]);
                    atomicAdd(&L.sx[c], (unsigned long long)(long long)llrint((double)e[0] * kFix));
                    atomicAdd(&L.sy[c], (unsigned long long)(long long)llrint((double)e[1] * kFix));
                    atomicAdd(&L.cnt[c], 1);
                }
            }
            wsync();
            for (int c = lane; c < k; c += 64) {
                if (L.cnt[c] > 0) {  // an empty cluster keeps its centre
                    const float nx = (float)((double)(long long)L.sx[c] / kFix / (double)L.cnt[c]);
                    const float ny = (float)((double)(long long)L.sy[c] / kFix / (double)L.cnt[c]);
                    if (nx != L.cx[c] || ny != L.cy[c]) L.changed = 1;
                    L.cx[c] = nx;
                    L.cy[c] = ny;
                }
            }
            wsync();
            if (!L.changed) break;
        }
        // ---- per-cluster means and standard deviations of X, Y, photons ----
        for (int c = lane; c < k; c += 64) {
            L.cnt[c] = 0;
            for (int j = 0; j < 6; j++) L.st[j][c] = 0.0;
        }
        wsync();
        for (int nn = lo + lane; nn < it; nn += 64) {
            if (count_fb(rec, em, nn - 1, fb) != mf) continue;
            const bamf_chain_rec q = rec[nn - 1];
            for (int i = 0; i < q.n; i++) {
                const float* e = em + 4 * ((int64_t)q.em_index + i);
                if (e[3] != fb) continue;
                const int c = nearest(L, k, e[0], e[1]);
                atomicAdd(&L.cnt[c], 1);
                for (int j = 0; j < 3; j++) {
                    const double v = (double)e[j];
                    atomicAdd(&L.st[j][c], v);
                    atomicAdd(&L.st[3 + j][c], v * v);
                }
            }
        }
        wsync();
        for (int c = lane; c < k; c += 64) {
            const double n = (double)L.cnt[c];
            for (int j = 0; j < 3; j++) {
                const double m = L.st[j][c] / n;
                const double var = n > 1 ? fmax((L.st[3 + j][c] - n * m * m) / (n - 1.0), 0.0) : 0.0;
                eo[6 * c + j] = n > 0 ? (float)m : __builtin_nanf("");
                eo[6 * c + 3 + j] = n > 0 ? (float)sqrt(var) : __builtin_nanf("");
            }
        }
        wsync();
    }
}

__global__ __launch_bounds__(64 * kWaves) void k_chain_summary(int n_roi, const bamf_chain_rec* acc_rec,
                                                               const float* acc_em, const int64_t* rec_count,
                                                               int32_t rec_cap, int32_t em_cap, int cap,
                                                               int32_t* status, int32_t* n_out, float* emo) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    for (int r = blockIdx.x * kWaves + w; r < n_roi; r += gridDim.x * kWaves) {
        const int64_t nloop = rec_count[3 * (int64_t)r];
        const bamf_chain_rec* rec = acc_rec + (int64_t)r * rec_cap;
        const float* em = acc_em + 4 * (int64_t)r * em_cap;
        float* eo = emo + (int64_t)r * cap * 6;
        if (nloop > rec_cap || rec_count[3 * (int64_t)r + 1] > em_cap || nloop < 2) {
            if (lane == 0) {
                status[r] = nloop < 2 ? 0 : 1;
                n_out[r] = 0;
            }
            continue;
        }
        const int it = (int)nloop;
        const int lo = (it + 1) / 2;
        // entries per state: N, or the placeholder's one; all states of the
        // range must have the same number (cat(1, X, x) needs it)
        const int n0 = max(rec[lo - 1].n, 1);
        int bad = 0;
        for (int nn = lo + lane; nn < it; nn += 64) bad |= max(rec[nn - 1].n, 1) != n0 ? 1 : 0;
        bad = wave_max_i(bad);
        if (bad || n0 > cap) {
            if (lane == 0) {
                status[r] = 2;
                n_out[r] = 0;
            }
            continue;
        }
        // lane j sums column j in chain order
        for (int j = lane; j < n0; j += 64) {
            double s[3] = {0.0, 0.0, 0.0}, s2[3] = {0.0, 0.0, 0.0};
            double n = 0.0;
            for (int nn = lo; nn < it; nn++) {
                const float* e = em + 4 * ((int64_t)rec[nn - 1].em_index + j);
                for (int q = 0; q < 3; q++) {
                    const double v = (double)e[q];
                    s[q] += v;
                    s2[q] += v * v;
                }
                n += 1.0;
            }
            for (int q = 0; q < 3; q++) {
                const double m = s[q] / n;
                eo[6 * j + q] = (float)m;
                eo[6 * j + 3 + q] = n > 1 ? (float)sqrt(fmax((s2[q] - n * m * m) / (n - 1.0), 0.0)) : 0.0f;
            }
        }
        if (lane == 0) {
            status[r] = 0;
            n_out[r] = n0;
        }
    }
}

}  // namespace
}  // namespace bamf

using namespace bamf;

extern "C" {

int bamf_find_map(const bamf_chain_rec* d_acc_rec, const float* d_acc_em, const int64_t* d_rec_count,
                  int32_t rec_cap, int32_t em_cap, int n_roi, int fb, int cap, int32_t* d_status, int32_t* d_n_map,
                  float* d_bg, float* d_em, void* stream) {
    BAMF_CHECK_ARG(n_roi >= 0, "n_roi must be >= 0");
    if (n_roi == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_acc_rec && d_acc_em && d_rec_count && d_status && d_n_map && d_bg && d_em, "null argument");
    BAMF_CHECK_ARG(rec_cap >= 1 && em_cap >= 1, "capacities must be >= 1");
    BAMF_CHECK_ARG(fb == 0 || fb == 1, "FB must be 0 or 1");
    BAMF_CHECK_ARG(cap >= 1 && cap <= kMaxN, "cap must be in 1..%d", kMaxN);
    const int grid = (int)ceil_div(n_roi, kWaves);
    hipLaunchKernelGGL(k_find_map, dim3(grid), dim3(64 * kWaves), 0, as_stream(stream), n_roi, d_acc_rec, d_acc_em,
                       d_rec_count, rec_cap, em_cap, (float)fb, cap, d_status, d_n_map, d_bg, d_em);
    BAMF_LAUNCH_CHECK("k_find_map");
    return BAMF_OK;
}

int bamf_chain_summary(const bamf_chain_rec* d_acc_rec, const float* d_acc_em, const int64_t* d_rec_count,
                       int32_t rec_cap, int32_t em_cap, int n_roi, int cap, int32_t* d_status, int32_t* d_n,
                       float* d_em, void* stream) {
    BAMF_CHECK_ARG(n_roi >= 0, "n_roi must be >= 0");
    if (n_roi == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_acc_rec && d_acc_em && d_rec_count && d_status && d_n && d_em, "null argument");
    BAMF_CHECK_ARG(rec_cap >= 1 && em_cap >= 1, "capacities must be >= 1");
    BAMF_CHECK_ARG(cap >= 1 && cap <= kMaxN, "cap must be in 1..%d", kMaxN);
    const int grid = (int)ceil_div(n_roi, kWaves);
    hipLaunchKernelGGL(k_chain_summary, dim3(grid), dim3(64 * kWaves), 0, as_stream(stream), n_roi, d_acc_rec,
                       d_acc_em, d_rec_count, rec_cap, em_cap, cap, d_status, d_n, d_em);
    BAMF_LAUNCH_CHECK("k_chain_summary");
    return BAMF_OK;
}

}  // extern "C"
