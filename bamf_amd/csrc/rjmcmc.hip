// Reversible-jump MCMC chains on gfx950: one wavefront per ROI chain,
// persistent work queue over ROIs.
//
// Chain semantics follow class RJMCMC (source/c/LibRJMCMC.cpp:657-1736):
// the same jump set (move with its four sub-moves, split, merge, birth,
// death, backFore), the same uniform draw order (urand/randn,
// LibRJMCMCreturn.hpp:17-26), the same proposal arithmetic in single
// precision without contraction, the same prior-table lookups and the same
// posterior-ratio factors, formed in the reference's order.
//
// gfx950 mapping
//   - pixel p of the ROI (column-major, y fastest) lives in lane p % 64, slot
//     p / 64: data, current model M and proposed model M' stay in VGPRs;
//   - the emitter sets (current and proposed) are double-buffered in LDS;
//     accepting a jump swaps the buffers;
//   - a proposal changes M by a short list of emitter stencils
//     (+new / -old): lanes 0..31 evaluate the x stencil, lanes 32..63 the y
//     stencil (two erf per lane), and every lane updates its pixels;
//   - the log-likelihood ratio sum(M - M' + D log(M'/M)) is reduced across
//     the wavefront with DPP row scans; the full model is rebuilt from the
//     emitter set after every `recompute_every` accepted jumps so that the
//     incremental updates cannot drift;
//   - draws live in a 64-position register window: lane l holds uniform
//     wbase + l of the step (a counter-based Philox4x32-10 word, or the tape
//     in replay mode) and the Irwin-Hall sum of the twelve uniforms starting
//     there, so urand() and randn() are one readlane each.  Every jump but
//     the group moves fits in one window; the group moves re-centre it;
//   - accept: the test U < PR decided in double precision (log U < log PR)
//     with a single-precision pre-test outside a narrow band around PR, so
//     the double-precision logs are formed for near-ties, accepted jumps
//     (the running log posterior) and step records only.
#include <cmath>

#include "common.hpp"
#include "erf_fast.hpp"
#include "philox.hpp"
#include "rjmcmc_dev.hpp"

#pragma clang fp contract(off)

namespace bamf {
namespace rj {
namespace {

constexpr int kMaxDelta = 2 * kNMax + 8;
static_assert(kNMax <= 256, "neighbour indices are stored as bytes");
constexpr double kPiD = 3.141592;  // LibRJMCMCreturn.hpp:2 (a double literal)

enum { C_N = 0, C_A = 5, C_NFG = 10, C_AFG = 15, C_NBG = 20, C_ABG = 25 };
enum { J_MOVE = 0, J_SPLIT = 1, J_MERGE = 2, J_BIRTH = 3, J_DEATH = 4 };
enum { ST_OK = 0, ST_NMAX = 1, ST_UNSUPPORTED = 2, ST_TAPE = 3, ST_DRAWS = 4, ST_SIZE = 5 };

struct Em {
    float I[kNMax];
    float X[kNMax];
    float Y[kNMax];
    int S[kNMax];
};

template <int PPL>
struct WaveLds {
    Em e[2];
    int16_t delta[kMaxDelta];  // stencil list codes: emitter | set << 8 | sign << 9
    float sten[64];
    uint8_t nb[kNMax + 1];     // neighbour indices (< kNMax)
    int hist[kNMax + 1];
    int cnt[32];
    float dpx[64 * PPL];  // the ROI's data in the pixel-slot layout: slot s at [64 s + lane]
    int64_t rec[3];       // chain-record counts: records, accepted-state entries, proposed-state entries
    // factors of the ratios that are fixed for a ROI or a phase, formed once
    // with the reference's operations (so the same floats): 1 / Area, the
    // phase's P[1] / P[2] and P[4] / P[3], and normpdf's 1 / (sqrt(2 pi) s)
    // and 2 s s for s = Split_std
    float kc[5];
};

__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int bcasti(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float bcastf(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Pixel slots of an X x Y ROI: columns padded to a power of two, 64 / Ypad per slot.
__host__ __device__ constexpr int slots_for(int X, int Y) {
    return Y <= 1 ? (X + 63) / 64 : Y <= 2 ? (X + 31) / 32 : Y <= 4 ? (X + 15) / 16 : Y <= 8 ? (X + 7) / 8
         : Y <= 16 ? (X + 3) / 4 : Y <= 32 ? (X + 1) / 2 : X * ((Y + 63) / 64);
}

// DPP move of a float (lanes whose source row is masked off or out of range read 0).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
// Inclusive prefix sum over the wavefront: Hillis-Steele within each row of
// 16 lanes (row_shr 1, 2, 4, 8), then the row totals carried forward with
// row_bcast:15 and row_bcast:31.
__device__ __forceinline__ float wave_scan(float v) {
    v = v + dpp_f<0x111, 0xf>(v);
    v = v + dpp_f<0x112, 0xf>(v);
    v = v + dpp_f<0x114, 0xf>(v);
    v = v + dpp_f<0x118, 0xf>(v);
    v = v + dpp_f<0x142, 0xa>(v);
    v = v + dpp_f<0x143, 0xc>(v);
    return v;
}
__device__ __forceinline__ int kth_bit(uint64_t m, int k) {
    for (int i = 0; i < k; i++) m &= m - 1;
    return __builtin_ctzll(m);
}

// LibRJMCMCreturn.hpp:28
__device__ __forceinline__ int kround(float v) { return (int)floor((double)v + 0.5); }

// LibRJMCMC.cpp:219-246: linear interpolation of the pdf table at I between
// the first grid point above I and the one before it.  When the grid is exact
// multiples of dx, that index is floor(I/dx) + 1 corrected by one comparison
// with the exact product; otherwise it is found by search from the same guess.
__device__ __forceinline__ float findpdf(const StatDev& s, float dx, float I) {
    if (!(I > 0.0f) || I >= s.x_last) return 0.0f;  // NaN also yields 0, as the reference does
    int j = (int)(I * s.inv_dx) + 1;  // a guess within one of floor(I / dx) + 1, corrected below
    j = j < 1 ? 1 : (j > s.len - 1 ? s.len - 1 : j);
    float xj, xj1;
    if (s.x_exact) {
        if ((float)(j - 1) * dx > I) j--;
        else if (j < s.len - 1 && (float)j * dx <= I) j++;
        xj = (float)j * dx;
        xj1 = (float)(j - 1) * dx;
    } else {
        while (j > 1 && s.x[j - 1] > I) j--;
        while (j < s.len - 1 && s.x[j] <= I) j++;
        xj = s.x[j];
        xj1 = s.x[j - 1];
    }
    const float num = s.pdf[j] * (I - xj1) + s.pdf[j - 1] * (xj - I);
    if (s.x_exact) {
        // num / dx correctly rounded from the correctly rounded 1/dx and one
        // residual step (Markstein)
        const float q0 = num * s.inv_dx;
        return fmaf(fmaf(-q0, dx, num), s.inv_dx, q0);
    }
    return num / (s.x[1] - s.x[0]);
}

// LibRJMCMC.cpp:248-282 with the uniform already drawn; the first uc > u is
// found by a forward scan from the guide entry of u, 64 entries per step
// (lane l tests entry hi + l; the first lane that stops the scan gives the
// index).  Called by the whole wavefront with a wave-uniform u.
__device__ __forceinline__ float genrand(const StatDev& s, float u) {
    if (u <= 0) return 0;
    if (u >= 1) return s.ux[s.nu - 1];
    if (s.nu < 2) return 0;
    int g = (int)(u * (float)kGuide);
    g = g > kGuide - 1 ? kGuide - 1 : g;
    int hi = s.guide[g] + 1;  // uc[guide[g]] <= g / kGuide <= u
    const int lane = (int)__lane_id();
    while (true) {
        const int k = hi + lane;
        const uint64_t stop = __ballot(k >= s.nu - 1 || s.uc[k] - u > 0);
        if (stop) {
            hi += (int)__builtin_ctzll(stop);
            break;
        }
        hi += 64;
    }
    return (s.ux[hi - 1] * (s.uc[hi] - u) + s.ux[hi] * (u - s.uc[hi - 1])) / (s.uc[hi] - s.uc[hi - 1]);
}

// LibRJMCMCreturn.hpp:19-21 (sqrt(2*pi) evaluated in double, then single).
__device__ __forceinline__ float normpdf(float x, float u, float s) {
    const float a = (float)sqrt(2.0f * kPiD);
    const float e = expf(-(x - u) * (x - u) / (2 * s * s));
    return 1.0f / (a * s) * e;
}


// Proposal kinds that need the likelihood ratio, with their accept counters.
enum { EV_NONE = 0, EV_MOVE, EV_SPLIT, EV_MERGE, EV_BIRTH, EV_DEATH, EV_GSPLIT, EV_GMERGE };

// Feature bits of a kernel variant; the plain one-launch kernel (F = 0)
// carries none of their code.  kGJ: the gSplit/gMerge jumps (reachable only
// when P_Burnin entries 5 or 6 are nonzero).  kSEG: the launch runs a segment
// of the chain (state loaded from and saved to HBM, seed read from device
// memory when given).  kLRN: the learned proposals of BAMF_RJ_LEARNED.
// kNUM: the sampled-PSF model (NumModel) in place of the erf stencils.
constexpr int kGJ = 1, kSEG = 2, kLRN = 4, kNUM = 8;

template <int PPL, int F>
struct Chain {
    static constexpr bool GJ = (F & kGJ) != 0, SEG = (F & kSEG) != 0, LRN = (F & kLRN) != 0;
    static constexpr bool NUM = (F & kNUM) != 0;
    // prior lookups paired over the lanes in the plain, learned and segmented
    // kernels: in the group-jump and sampled-PSF variants the divergent lookup
    // has not been measured
    static constexpr bool PAIR = F == 0 || F == kLRN || F == kSEG;
    static_assert(!(NUM && LRN), "the learned proposals use the erf stencil");
    const ChainArgs* a;  // the kernel argument block (in the kernarg segment)
    WaveLds<PPL>& L;
    const int lane;
    int roi, X, Y, P, ppl;
    float C, meanN;
    int bnd;
    float xb, yb;  // Xsize + Bnd_out, Ysize + Bnd_out
    int ci;        // current emitter buffer
    int cN, tN;
    float cBG, cABG, cBBG, tBG, tABG, tBBG;
    float m[PPL], mt[PPL];
    // Pixel layout: the column height is padded to Ypad = 2^lgY >= Y and a
    // slot holds cps = 64 / Ypad whole columns: lane l of slot s is pixel
    // (x = s * cps + (l >> lgY), y = l & (Ypad - 1)), present when x < X and
    // y < Y.  Lane order within a slot is then the reference's pixel order
    // (x * Y + y), and a lane's y never changes between slots.
    int lgY, cps;
    int xsub, ysub;
    // draws: window of 64 positions starting at wbase; dj = next position
    int wbase, dj;
    float wu, wr;
    uint32_t pw[4];  // Philox block of window position `lane` for steps 4 pblk .. 4 pblk + 3
    int pblk;  // uniform at wbase + lane, Irwin-Hall sum of the 12 uniforms from there
    uint32_t key, sd_lo, sd_hi;
    int step;
    int64_t tbase, tlen, tcur;
    // phase
    float P1, P2, P3, P4, P56;  // P[1..4] and P[5] / P[6] of the phase
    bool burn;
    int status;
    int nacc;
    double logpost;
    int ndelta;
    bool bgdelta;
    float cbg_pdf, tbg_pdf;  // offset-prior pdf at the current / proposed BG
    int sxa, sxb;      // x range of the last stencil's nonzero columns (sxa > sxb: none)
    int lo_s, hi_s;    // pixel slots the proposal changes
    bool e_ok;         // the current set passes Emitters::test
    bool t_bad;        // a value written into the proposed set fails it
    // proposal bookkeeping for the shared evaluation
    int ev, jt, acc_c1, acc_c2;
    bool need_test;
    float f_pri, f_u1, f_u2, f_u3, f_ijp, f_pixprob, f_iratio;
    float f_mz;  // learned split/merge: q_map(pixel of the first child) / Z of the window, 0 outside it
    int f_jj, f_jm;
    // step record
    int r_jt, r_flags, r_nprop, r_pix;
    float r_llr, r_pr;
    double r_logpr;
    int ref_flags, ref_pix;
    bool applied;                     // the decision the step applied (the recorded one in replay mode)
    double lp_prod;  // product of the posterior-ratio factors other than the LLR, in double
    bool lp_neg;     // merge divides by the product
    uint64_t aflops;  // algorithmic work of this chain

    __device__ Chain(const ChainArgs* a_, WaveLds<PPL>& L_, int lane_) : a(a_), L(L_), lane(lane_) {}

    __device__ __forceinline__ Em& cur() { return L.e[ci]; }
    __device__ __forceinline__ Em& tst() { return L.e[ci ^ 1]; }
    // The lane coordinates pass through an empty asm at each use: without it
    // the compiler hoists every per-slot value derived from them (the slot's
    // column, its pixel address, its validity mask) out of the chain loop and
    // keeps all of them live across it, which spills; re-deriving one where
    // it is used costs an instruction or two.
    __device__ __forceinline__ static int opq(int v) {
        asm volatile("" : "+v"(v));
        return v;
    }
    __device__ __forceinline__ int px(int s) const { return s * cps + opq(xsub); }
    __device__ __forceinline__ bool has(int s) const { return opq(ysub) < Y && px(s) < X; }
    __device__ __forceinline__ int pidx(int s) const { return px(s) * Y + opq(ysub); }

    // ---------------- draws ----------------
    // Window of draws wbase .. wbase + 63 of this step.  Draw j of step s is
    // word s % 4 of Philox4x32-10 with counter (s / 4, j, seed_hi, 0) and key
    // (roi key, seed_lo), shifted to 31 bits (the range of rand()): the block
    // lane l computes for window position l serves four consecutive steps and
    // is kept for the three after it.  Replay mode reads the tape instead (0
    // past its end, flagged at the step end).
    // urand = (float)raw / 2^31 (LibRJMCMCreturn.hpp:17); the twelve-term
    // sum of randn (LibRJMCMCreturn.hpp:23-26) is formed left to right from
    // each position, so randn() is that sum minus 6, bit for bit.
    __device__ __forceinline__ void fill_window(int base) {
        wbase = base;
        const uint32_t q = (uint32_t)(base + lane);
        uint32_t w;
        if (a->mode == BAMF_RJ_REPLAY) {
            w = (tcur + (int64_t)q < tlen) ? a->tape[tbase + tcur + q] : 0u;
        } else {
            const uint32_t blk = (uint32_t)step >> 2;
            uint32_t o[4];
            if (base == 0 && blk == (uint32_t)pblk) {
                o[0] = pw[0];
                o[1] = pw[1];
                o[2] = pw[2];
                o[3] = pw[3];
            } else {
                if constexpr (SEG) philox(blk, q, sd_hi, 0u, key, sd_lo, o);
                else philox(blk, q, a->seed_hi, 0u, key, a->seed_lo, o);
                if (base == 0) {
                    pw[0] = o[0];
                    pw[1] = o[1];
                    pw[2] = o[2];
                    pw[3] = o[3];
                    pblk = (int)blk;
                }
            }
            const uint32_t c = (uint32_t)step & 3u;
            w = (c == 0 ? o[0] : c == 1 ? o[1] : c == 2 ? o[2] : o[3]) >> 1;
        }
        wu = (float)w / 2147483648.0f;
        // lane l: ((u_l + u_{l+1}) + u_{l+2}) + ... + u_{l+11}, the reference's
        // left-to-right order; each term arrives by a one-lane wavefront shift
        // (DPP wave_shl:1, lane l reads lane l + 1).  Lanes above 52 sum
        // shifted-in zeros and are never read (randn() takes positions <= 52).
        float t = wu, v = wu;
#pragma unroll
        for (int i = 1; i < 12; i++) {
            v = dpp_f<0x130, 0xf>(v);
            t = t + v;
        }
        wr = t;
    }
    // Re-centres the window when fewer than n positions remain in it.
    __device__ __forceinline__ void need(int n) {
        if (dj - wbase + n > 64) fill_window(dj);
    }
    __device__ __forceinline__ float urand() {
        const int k = dj - wbase;
        if (k > 63) status = ST_DRAWS;
        dj++;
        return bcastf(wu, k & 63);
    }
    __device__ __forceinline__ float randn() {
        const int k = dj - wbase;
        if (k > 52) status = ST_DRAWS;
        dj += 12;
        return bcastf(wr, k & 63) - 6.0f;
    }
    __device__ __forceinline__ void cnt_inc(int k) {
        if (lane == 0) L.cnt[k]++;
    }

    // ---------------- emitter sets ----------------
    __device__ __forceinline__ void copy_to_test() {
        Em& c = cur();
        Em& t = tst();
        for (int n = lane; n < cN; n += 64) {
            t.I[n] = c.I[n];
            t.X[n] = c.X[n];
            t.Y[n] = c.Y[n];
            t.S[n] = c.S[n];
        }
        tN = cN;
        t_bad = false;
        tBG = cBG;
        tABG = cABG;
        tBBG = cBBG;
        ndelta = 0;
        bgdelta = false;
        wsync();
    }
    // One emitter's part of Emitters::test (LibRJMCMC.cpp:492-514).
    __device__ __forceinline__ bool bad_val(float I, float x, float y) const {
        return (I < 0) || (x < (float)(-bnd)) || (y < (float)(-bnd)) || (x > (float)(X + bnd)) ||
               (y > (float)(Y + bnd));
    }
    __device__ __forceinline__ void set_t(int k, float I, float x, float y) {
        t_bad |= bad_val(I, x, y);
        if (lane == 0) {
            tst().I[k] = I;
            tst().X[k] = x;
            tst().Y[k] = y;
        }
    }
    __device__ __forceinline__ bool add_t(float I, float x, float y, int s) {
        if (tN >= kNMax) {
            status = ST_NMAX;
            return false;
        }
        t_bad |= bad_val(I, x, y);
        if (lane == 0) {
            tst().I[tN] = I;
            tst().X[tN] = x;
            tst().Y[tN] = y;
            tst().S[tN] = s;
        }
        tN++;
        wsync();
        return true;
    }
    __device__ __forceinline__ void remove_t(int id) {
        Em& t = tst();
        wsync();
        float vi[2], vx[2], vy[2];
        int vs[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int n = lane + 64 * h;
            if (n >= id && n < tN - 1) {
                vi[h] = t.I[n + 1];
                vx[h] = t.X[n + 1];
                vy[h] = t.Y[n + 1];
                vs[h] = t.S[n + 1];
            }
        }
        wsync();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int n = lane + 64 * h;
            if (n >= id && n < tN - 1) {
                t.I[n] = vi[h];
                t.X[n] = vx[h];
                t.Y[n] = vy[h];
                t.S[n] = vs[h];
            }
        }
        tN--;
        wsync();
    }
    __device__ __forceinline__ void delta_push(int which, int k, bool neg) {
        if (lane == 0) L.delta[ndelta] = (int16_t)(k | (which << 8) | ((neg ? 1 : 0) << 9));
        ndelta++;
    }
    // Emitters::test (LibRJMCMC.cpp:492-514) on the proposed set: only the
    // values the proposal wrote can fail it once the current set is known to
    // pass; otherwise every emitter is checked.
    __device__ __forceinline__ bool test_t() {
        if (t_bad) return false;
        if (e_ok) return true;
        const Em& t = tst();
        bool bad = false;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int n = lane + 64 * h;
            if (n < tN) {
                const float I = t.I[n], x = t.X[n], y = t.Y[n];
                bad |= (I < 0) || (x < (float)(-bnd)) || (y < (float)(-bnd)) || (x > (float)(X + bnd)) ||
                       (y > (float)(Y + bnd));
            }
        }
        return ballot(bad) == 0;
    }
    __device__ __forceinline__ int nem_cur() {
        const Em& c = cur();
        int k = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int n = lane + 64 * h;
            k += __popcll(ballot(n < cN && c.S[n] == 1));
        }
        return k;
    }
    // Neighbours of emitter jj in the current set closer than dmin, as two
    // 64-bit index masks.  Distance as closest()/closestMulti() compute it:
    // float differences, squared and summed in double, sqrt in double, then float.
    __device__ __forceinline__ void near_mask(int jj, float dmin, uint64_t& m0, uint64_t& m1) {
        const Em& c = cur();
        const float xj = c.X[jj], yj = c.Y[jj];
        bool in[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int n = lane + 64 * h;
            in[h] = false;
            if (n < cN && n != jj) {
                const double dx = (double)(c.X[n] - xj), dy = (double)(c.Y[n] - yj);
                const float dd = (float)sqrt(dx * dx + dy * dy);
                in[h] = dd < dmin;
            }
        }
        m0 = ballot(in[0]);
        m1 = ballot(in[1]);
    }
    __device__ __forceinline__ int pick(uint64_t m0, uint64_t m1, int k) {
        const int c0 = __popcll(m0);
        return k < c0 ? kth_bit(m0, k) : 64 + kth_bit(m1, k - c0);
    }
    // LibRJMCMC.cpp:525-558
    __device__ __forceinline__ int closest(int jj, float dmin) {
        uint64_t m0, m1;
        near_mask(jj, dmin, m0, m1);
        const int num = __popcll(m0) + __popcll(m1);
        if (num > 0) {
            const float Ua = urand();
            const int mm = min((int)(Ua * num), num - 1);
            return pick(m0, m1, mm);
        }
        const float Ua = urand();
        const int mm = min((int)(Ua * (cN - 1)), cN - 2);
        return mm < jj ? mm : mm + 1;
    }
    // LibRJMCMC.cpp:615-650: writes the chosen indices to L.nb, returns how many.
    __device__ __forceinline__ int closest_multi(int jj, float dmin) {
        uint64_t m0, m1;
        near_mask(jj, dmin, m0, m1);
        const int num = __popcll(m0) + __popcll(m1);
        int np = 0;
        if (num > 0) {
            float Ua = urand();
            if (Ua < 1.0f / (float)num) np = num;
            else np = min((int)(Ua * num), num - 1);
            for (int n = 0; n < np; n++) {
                need(1);
                Ua = urand();
                const int k = min((int)(Ua * (num - n)), num - n - 1);
                const int idx = pick(m0, m1, k);
                if (idx < 64) m0 &= ~(1ull << idx);
                else m1 &= ~(1ull << (idx - 64));
                if (lane == 0) L.nb[n] = (uint8_t)idx;
            }
        }
        wsync();
        return np;
    }

    // ---------------- model ----------------
    // Stencil of one emitter (LibRJMCMC.cpp:15-36): N_X[x] = I*0.5*(erf(e_x+1) - erf(e_x)) into
    // sten[x] (lanes x < X), and this lane's N_Y[y] = 0.5*(erf(e_y+1) - erf(e_y)) returned, with
    // pixel edges e_k = (k - pos)*C.  erf is the branch-free one, two edges per lane.
    __device__ __forceinline__ float stencil(float I, float xe, float ye) {
        const float l = (float)lane;
        // each edge's erf once: the next x edge arrives by a lane shift, the
        // lane's two y edges by permutes from the lanes that evaluated them
        float vx, ey0, ey1;
        if (X + Y + 2 <= 64) {
            // one erf per lane: x edges 0..X on lanes 0..X, y edges 0..Y on lanes X+1..X+Y+1
            const int by = X + 1;
            const float e = erf_nr(lane < by ? (l - xe) * C : ((float)(lane - by) - ye) * C);
            vx = 0.5f * (__shfl_down(e, 1) - e);
            const int src = by + (ysub < Y ? ysub : 0);
            ey0 = __shfl(e, src);
            ey1 = __shfl(e, src + 1);
        } else if (15.0f * C >= 5.0f) {
            // one erf per lane over a window of 32 edges around each coordinate
            // (x edges on lanes 0..31, y edges on 32..63): an edge outside it
            // lies 15 px or more from the emitter, |e| >= 15 C >= 5, where erf
            // is exactly -1 or +1, so those edges are filled in without evaluation
            const int wx = (int)floorf(xe) - 15, wy = (int)floorf(ye) - 15;
            const bool isx = lane < 32;
            const int k = isx ? wx + lane : wy + (lane - 32);
            const float e = erf_nr(((float)k - (isx ? xe : ye)) * C);
            auto edge = [&](int kk, int w, int base) {
                const int d = kk - w;
                const float v = __shfl(e, base + (d < 0 ? 0 : (d > 31 ? 31 : d)));
                return d < 0 ? -1.0f : (d > 31 ? 1.0f : v);
            };
            vx = 0.5f * (edge(lane + 1, wx, 0) - edge(lane, wx, 0));
            ey0 = edge(ysub, wy, 32);
            ey1 = edge(ysub + 1, wy, 32);
        } else {
            const float ex = erf_nr((l - xe) * C), ey = erf_nr((l - ye) * C);
            vx = 0.5f * (__shfl_down(ex, 1) - ex);
            ey0 = __shfl(ey, ysub);
            ey1 = __shfl(ey, ysub + 1);
        }
        const float vy = 0.5f * (ey1 - ey0);
        // columns x >= X and rows y >= Y (padding lanes) get exact zeros, so
        // the pixel loops need no lane masks
        L.sten[lane] = lane < X ? I * vx : 0.0f;
        // columns where the x factor (before the intensity) is nonzero; a
        // zero factor makes the whole column's contribution exactly zero
        const uint64_t nz = ballot(lane < X && vx != 0.0f);
        sxa = nz ? (int)__builtin_ctzll(nz) : 1;
        sxb = nz ? 63 - (int)__builtin_clzll(nz) : 0;
        wsync();
        return ysub < Y ? vy : 0.0f;
    }
    // ---------------- sampled PSF (NumModel, LibRJMCMC.cpp:67-169) ----------------
    // Table entry (y, x, phase_y, phase_x) at y + P0 x + P01 phase_y + P012
    // phase_x; OS = PSFsize(4) phases per pixel.  Per emitter (wave-uniform,
    // in double as the reference forms them): the bases of the four
    // interpolation corners (phase and next phase along each axis, the phase
    // after the last being phase 0 one pixel on) and the two ratios; pixel
    // (x, y) then reads corner k at nb[k] + y + P0 x, a linear index outside
    // the table reading the table's last entry (LibRJMCMC.cpp:131-158).  No
    // contraction, so each term rounds as the reference's does.
    int nb[4];
    float nxr, nyr;
    __device__ __forceinline__ void num_setup(float xe, float ye) {
#pragma clang fp contract(off)
        const int p0 = a->psf_dims[0], p01 = p0 * a->psf_dims[1], p012 = p01 * a->psf_dims[2];
        const double os = (double)a->psf_dims[3];
        const int st0 = p0 / 2;
        const double xi = (double)xe, yi = (double)ye, fx = floor(xi), fy = floor(yi);
        const int stx = (int)ceil(st0 - xi), imx = (int)floor(os * (xi - fx));
        const int sty = (int)ceil(st0 - yi), imy = (int)floor(os * (yi - fy));
        const bool wx = imx == os - 1, wy = imy == os - 1;
        const int stx2 = wx ? stx - 1 : stx, imx2 = wx ? 0 : imx + 1;
        const int sty2 = wy ? sty - 1 : sty, imy2 = wy ? 0 : imy + 1;
        nb[0] = sty + stx * p0 + imy * p01 + imx * p012;
        nb[1] = sty + stx2 * p0 + imy * p01 + imx2 * p012;
        nb[2] = sty2 + stx * p0 + imy2 * p01 + imx * p012;
        nb[3] = sty2 + stx2 * p0 + imy2 * p01 + imx2 * p012;
        nyr = (float)(1 - ((yi - fy) - (double)imy / os) * os);
        nxr = (float)(1 - ((xi - fx) - (double)imx / os) * os);
    }
    // This lane's pixel of slot s in the unit image of the last num_setup
    // emitter (padding lanes read clamped, in-table entries).
    __device__ __forceinline__ float num_at(int s) const {
#pragma clang fp contract(off)
        const int n = a->psf_numel;
        const int q = ysub + px(s) * a->psf_dims[0];
        int i[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int v = nb[k] + q;
            i[k] = (unsigned)v >= (unsigned)n ? n - 1 : v;
        }
        const float* T = a->psf;
        const float A = nxr * T[i[0]] + (1.0f - nxr) * T[i[1]];
        const float B = nxr * T[i[2]] + (1.0f - nxr) * T[i[3]];
        return nyr * A + (1.0f - nyr) * B;
    }
    __device__ __forceinline__ static float num_acc(float m, float I, float v, bool neg) {
#pragma clang fp contract(off)
        const float pr = I * v;
        return neg ? m - pr : m + pr;
    }

    // ---------------- conditional intensity fit (learned mode) ----------------
    // R = M minus emitter (I, x, y) into mt[], leaving that emitter's unit
    // stencil in place for cond_fit.
    float c_vy;
    __device__ __forceinline__ void cond_prep(float I, float x, float y) {
        c_vy = stencil(1.0f, x, y);
        const int s0 = sxa >> (6 - lgY), s1 = sxb >> (6 - lgY);
#pragma unroll
        for (int s = 0; s < PPL; s++) {
            mt[s] = m[s];
            if (s >= s0 && s <= s1) mt[s] = m[s] - I * (L.sten[px(s)] * c_vy);
        }
    }
    __device__ __forceinline__ void cond_stencil(float x, float y) { c_vy = stencil(1.0f, x, y); }
    // Gaussian approximation of the Poisson likelihood of D = R + I g in I
    // (g the unit stencil in place): weighted least squares with weights 1/R,
    // then one Fisher-scoring step with the information F = sum g^2 / (R + I g)
    // at that start; sd = inflation / sqrt(F).  No weight inside the ROI: the
    // signal prior's mean and sd.
    __device__ __forceinline__ void cond_fit(float& mean, float& sd) {
        const int s0 = sxa >> (6 - lgY), s1 = sxb >> (6 - lgY);
        float a0 = 0.0f, b0 = 0.0f;
#pragma unroll
        for (int s = 0; s < PPL; s++)
            if (s >= s0 && s <= s1) {
                const float g = L.sten[px(s)] * c_vy, r = fmaxf(mt[s], 1e-3f), d = L.dpx[64 * s + lane];
                const float w = __builtin_amdgcn_rcpf(r);
                a0 = a0 + g * (d - r) * w;
                b0 = b0 + g * g * w;
            }
        a0 = bcastf(wave_scan(a0), 63);
        b0 = bcastf(wave_scan(b0), 63);
        if (!(b0 > 1e-12f)) {
            mean = a->c_mean;
            sd = a->c_sd;
            return;
        }
        const float m0 = fmaxf(a0 / b0, 1.0f);
        float sc = 0.0f, fi = 0.0f;
#pragma unroll
        for (int s = 0; s < PPL; s++)
            if (s >= s0 && s <= s1) {
                const float g = L.sten[px(s)] * c_vy, r = fmaxf(mt[s], 1e-3f), d = L.dpx[64 * s + lane];
                const float w = __builtin_amdgcn_rcpf(r + m0 * g);
                sc = sc + g * (d * w - 1.0f);
                fi = fi + g * g * w;
            }
        sc = bcastf(wave_scan(sc), 63);
        fi = bcastf(wave_scan(fi), 63);
        mean = m0 + sc / fi;
        sd = a->c_infl * __builtin_amdgcn_rsqf(fi);
    }

    // Full model of the current set: background plane + sCMOS term, then the
    // emitters in order (Model::Model, LibRJMCMC.cpp:42-59).
    __device__ __forceinline__ void rebuild_model() {
        const int64_t off = a->pix_off[roi];
#pragma unroll
        for (int s = 0; s < PPL; s++)
            if (s < ppl) {
                const bool ok = has(s);
                const float sc = (a->scmos && ok) ? a->scmos[off + pidx(s)] : 0.0f;
                m[s] = ok ? cBG + cABG * (float)px(s) + cBBG * (float)ysub + sc : 1.0f;
            }
        const Em& c = cur();
        aflops += (uint64_t)(4 * P) + (uint64_t)cN * (uint64_t)(NUM ? 10 * P : 2 * P + 8 * (X + Y));
        for (int n = 0; n < cN; n++) {
            if constexpr (NUM) {
                num_setup(c.X[n], c.Y[n]);
                const float I = c.I[n];
#pragma unroll
                for (int s = 0; s < PPL; s++)
                    if (s < ppl && has(s)) m[s] = num_acc(m[s], I, num_at(s), false);
            } else {
                const float vy = stencil(c.I[n], c.X[n], c.Y[n]);
#pragma unroll
                for (int s = 0; s < PPL; s++)
                    if (s < ppl) m[s] = m[s] + L.sten[px(s)] * vy;
                wsync();
            }
        }
        nacc = 0;
    }
    // M' = M + background change + the stencil list; returns the sum over
    // pixels of M - M' + D log(M'/M) (likelihood_ratio, LibRJMCMC.cpp:1677-1678).
    __device__ __forceinline__ float proposal_llr_sum() {
        aflops += (uint64_t)(6 * P) + (uint64_t)ndelta * (uint64_t)(NUM ? 10 * P : 2 * P + 8 * (X + Y)) +
                  (bgdelta ? 5 * P : 0);
#pragma unroll
        for (int s = 0; s < PPL; s++) mt[s] = m[s];
        lo_s = ppl;
        hi_s = -1;
        if (bgdelta) {
            const float db = tBG - cBG, da = tABG - cABG, dbb = tBBG - cBBG;
#pragma unroll
            for (int s = 0; s < PPL; s++)
                if (s < ppl && has(s)) mt[s] = mt[s] + (db + da * (float)px(s) + dbb * (float)ysub);
            lo_s = 0;
            hi_s = ppl - 1;
        }
        wsync();
        for (int e = 0; e < ndelta; e++) {
            const int code = L.delta[e];
            const Em& E = ((code >> 8) & 1) ? tst() : cur();
            const int k = code & 255;
            const bool neg = (code >> 9) & 1;
            if constexpr (NUM) {  // every pixel reads the table
                num_setup(E.X[k], E.Y[k]);
                const float I = E.I[k];
                lo_s = 0;
                hi_s = ppl - 1;
#pragma unroll
                for (int s = 0; s < PPL; s++)
                    if (s < ppl && has(s)) mt[s] = num_acc(mt[s], I, num_at(s), neg);
                continue;
            }
            const float vy = stencil(E.I[k], E.X[k], E.Y[k]);
            // slots holding columns sxa..sxb
            const int s0 = sxa >> (6 - lgY), s1 = sxb >> (6 - lgY);
            lo_s = min(lo_s, s0);
            hi_s = max(hi_s, s1);
#pragma unroll
            for (int s = 0; s < PPL; s++)
                if (s >= s0 && s <= s1) {
                    const float pr = L.sten[px(s)] * vy;
                    mt[s] = neg ? mt[s] - pr : mt[s] + pr;
                }
            wsync();
        }
        // M - M' + D log(M'/M); slots outside lo_s..hi_s, and padding lanes
        // (D = 0, M = M' = 1), have M' == M and add exactly 0
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < PPL; s++)
            if (s >= lo_s && s <= hi_s) {
                // M'/M from the hardware reciprocal and one residual correction
                // (the correctly rounded quotient up to rare last-bit ties)
                const float r = __builtin_amdgcn_rcpf(m[s]);
                const float q0 = mt[s] * r;
                const float q = fmaf(fmaf(-q0, m[s], mt[s]), r, q0);
                const float lq = __builtin_amdgcn_logf(q) * 0.693147180559945f;  // hardware log2
                acc = acc + ((m[s] - mt[s]) + L.dpx[64 * s + lane] * lq);
            }
        return bcastf(wave_scan(acc), 63);
    }

    // ---------------- accept ----------------
    // log PR in double: the LLR's exponent plus the log of the other factors'
    // product formed in double from their single-precision values.
    __device__ __forceinline__ double logpr_of(float sum) const {
        const double l = log(lp_prod);
        return (double)sum + (lp_neg ? -l : l);
    }
    // The accept test U < PR (LibRJMCMC.cpp:939 and the other jumps), decided
    // in double precision as log U < log PR.  A single-precision estimate of
    // both logs (hardware log2, error far below 1e-5 (1 + |log PR|)) settles
    // every case outside that band; the double-precision logs are formed only
    // inside it.  When the single-precision PR is saturated (inf, 0, NaN or
    // beyond 1e+-30) the reference's own test decides, keeping its overflow
    // semantics.
    __device__ __forceinline__ bool decide(float pr, float sum) {
        need(1);
        const float U = urand();
        bool acc;
        if (isfinite(pr) && pr > 1e-30f && pr < 1e30f) {
            const float fp = (float)lp_prod;
            int v = -1;  // -1 undecided, else the decision
            if (fp > 1e-30f && fp < 1e30f) {
                const float lf = __logf(fp);
                const float lpf = sum + (lp_neg ? -lf : lf);
                const float lu = __logf(U);  // -inf for U = 0
                const float band = 1e-5f * (1.0f + fabsf(lpf));
                if (lu < lpf - band) v = 1;
                else if (lu > lpf + band) v = 0;
            }
            acc = v >= 0 ? v == 1 : log((double)U) < logpr_of(sum);
        } else {
            acc = U < pr;
            r_flags |= 16;
        }
        r_pr = pr;
        if (acc) r_flags |= 1;
        if (a->mode == BAMF_RJ_REPLAY) {
            const bool forced = (ref_flags & 1) != 0;
            if (forced != acc) r_flags |= 32;
            acc = forced;
        }
        return acc;
    }
    // Chain and Chain_Test entries (RJMCMCChain::addEmitters, LibRJMCMC.cpp:290-322):
    // the current state and the proposed one (the current buffer when the
    // jump was accepted), with the jump's LLR, PR, type and decision.
    __device__ __forceinline__ void chain_entry(bamf_chain_rec* recs, float* ems, int which, const Em& e, int n,
                                                float bg, float abg, float bbg) {
        const int64_t r0 = (int64_t)roi * a->rec_cap, e0 = (int64_t)roi * a->em_cap;
        const int64_t nrec = L.rec[0];
        int64_t nem = L.rec[which];
        if (lane == 0 && nrec < a->rec_cap) {
            bamf_chain_rec& h = recs[r0 + nrec];
            h.n = n;
            h.jump_type = jt;
            h.accepted = applied ? 1 : 0;
            h.em_index = (int32_t)nem;
            h.llr = r_llr;
            h.pr = r_pr;
            h.bg = bg;
            h.abg = abg;
            h.bbg = bbg;
        }
        if (n < 1) {
            if (lane == 0 && nem < a->em_cap) {
                float* q = ems + 4 * (e0 + nem);
                q[0] = -100.0f;
                q[1] = -100.0f;
                q[2] = 0.0f;
                q[3] = -1.0f;
            }
            nem += 1;
        } else {
            for (int k = lane; k < n; k += 64)
                if (nem + k < a->em_cap) {
                    float* q = ems + 4 * (e0 + nem + k);
                    q[0] = e.X[k] + a->p.drift_x;
                    q[1] = e.Y[k] + a->p.drift_y;
                    q[2] = e.I[k];
                    q[3] = (float)e.S[k];
                }
            nem += n;
        }
        wsync();
        if (lane == 0) L.rec[which] = nem;
        wsync();
    }
    __device__ __forceinline__ void record() {
        r_jt = jt;
        r_flags |= 2;
        r_nprop = tN;
        if (a->acc_rec) {
            const bool acc = applied;
            chain_entry(a->acc_rec, a->acc_em, 1, cur(), cN, cBG, cABG, cBBG);
            chain_entry(a->prop_rec, a->prop_em, 2, acc ? cur() : tst(), acc ? cN : tN, tBG, tABG, tBBG);
            if (lane == 0) L.rec[0] += 1;
            wsync();
        }
    }
    // update_grids (LibRJMCMC.cpp:1629-1648)
    __device__ __forceinline__ void update_grids() {
        r_flags |= 4;
        if (burn || !a->pgrid) return;
        const Em& c = cur();
        const int z = a->p.grid_zoom;
        const int64_t go = a->grid_off[roi];
        // the grid's column stride and kept window: the whole ROI grid, or its
        // core inside a stitched image (counts exact: integer-valued floats)
        int ld = Y * z, xl = 0, xh = X * z, yl = 0, yh = Y * z;
        if (a->grid_place) {
            const int32_t* gp = a->grid_place + 5 * (int64_t)roi;
            ld = gp[0]; xl = gp[1]; xh = gp[2]; yl = gp[3]; yh = gp[4];
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int n = lane + 64 * h;
            if (n < cN) {
                const int x = kround((float)z * (c.X[n] + a->p.drift_x));
                const int y = kround((float)z * (c.Y[n] + a->p.drift_y));
                const float i = c.I[n];
                if (x >= xl && x < xh && y >= yl && y < yh && i > a->p.icutoff) {
                    float* g = c.S[n] == 1 ? a->pgrid : a->pback;
                    atomicAdd(&g[go + (int64_t)x * ld + y], 1.0f);
                }
            }
        }
    }
    __device__ __forceinline__ float pdf_of(const Em& e, int k) {
        return e.S[k] == 1 ? findpdf(a->sig, a->p.dx, e.I[k]) : findpdf(a->back, a->p.dx, e.I[k]);
    }
    // normpdf(x, 0, Split_std) (LibRJMCMCreturn.hpp:19-21) with its two
    // constants formed once: -(x - 0)(x - 0) is -(x x), bit for bit
    __device__ __forceinline__ float normpdf_split(float x) const {
        const float e = expf(-x * x / L.kc[4]);
        return L.kc[3] * e;
    }
    // Lookups of one table in a single pass over the lanes: lane l looks up
    // the (l mod 2)-th (or (l mod 4)-th) intensity, read back by readlane;
    // each value is the one findpdf returns for it alone.
    __device__ __forceinline__ void findpdf2(const StatDev& s, float Ia, float Ib, float& ra, float& rb) const {
        if constexpr (PAIR) {
            const float v = findpdf(s, a->p.dx, (lane & 1) ? Ib : Ia);
            ra = bcastf(v, 0);
            rb = bcastf(v, 1);
        } else {
            ra = findpdf(s, a->p.dx, Ia);
            rb = findpdf(s, a->p.dx, Ib);
        }
    }
    __device__ __forceinline__ void findpdf3(const StatDev& s, float Ia, float Ib, float Ic, float& ra, float& rb,
                                             float& rc) const {
        if constexpr (PAIR) {
            const int q = lane & 3;
            const float v = findpdf(s, a->p.dx, q == 0 ? Ia : q == 1 ? Ib : Ic);
            ra = bcastf(v, 0);
            rb = bcastf(v, 1);
            rc = bcastf(v, 2);
        } else {
            ra = findpdf(s, a->p.dx, Ia);
            rb = findpdf(s, a->p.dx, Ib);
            rc = findpdf(s, a->p.dx, Ic);
        }
    }

    // The network's log-intensity at the pixel a birth would propose (x, y)
    // from (pixel (floor(x + 1/2), floor(y + 1/2))); -inf outside the ROI.
    __device__ __forceinline__ float map_mu(float x, float y) const {
        const int px = (int)floorf(x + 0.5f), py = (int)floorf(y + 0.5f);
        if (!(px >= 0 && px < X && py >= 0 && py < Y)) return -INFINITY;
        return a->prop_mu[a->pix_off[roi] + px * Y + py];
    }

    // ---------------- proposals ----------------
    // Each prop_* follows the reference jump up to the likelihood ratio and
    // returns the evaluation kind (EV_NONE when the jump ended early).
    // LibRJMCMC.cpp:752-916
    __device__ __forceinline__ int prop_move() {
        const bamf_mcmc_params& p = a->p;
        jt = 1;
        float PrI = 1, IPDF = 1, IPDFtest = 1;
        cnt_inc(C_N + J_MOVE);
        copy_to_test();
        if (cN < 1) return EV_NONE;
        const Em& E = cur();
        const Em& T = tst();
        const float Ua = urand();
        const int nn = min((int)(Ua * tN), tN - 1);
        const float Ub = urand();
        if (Ub < 0.25) {
            jt = 2;
            tBG = tBG + p.bg_std * randn();
            tABG = tABG + p.abbg_std * randn();
            tBBG = tBBG + p.abbg_std * randn();
            bgdelta = true;
        } else if (Ub < 0.50 && Ub >= 0.25) {
            if (cN < 2) return EV_NONE;
            jt = 10;
            const int np = closest_multi(nn, 2 * p.psf_sigma);
            if (np == 0) return EV_NONE;
            float I0 = E.I[nn];
            for (int ii = 0; ii < np; ii++) I0 = I0 + E.I[L.nb[ii]];
            float II0 = 0, XI0 = 0, YI0 = 0, XII0 = 0, YII0 = 0;
            for (int ii = 0; ii < np; ii++) {
                const int k = L.nb[ii];
                const bool fg = E.S[k] == 1;
                const float xs = fg ? p.x_std_fg : p.x_std_bg;
                need(36);
                const float tx = E.X[k] + xs * randn();
                const float ty = E.Y[k] + xs * randn();
                const float ti = (I0 / (float)(np + 1)) * (0.15f * randn() + 1);
                findpdf2(fg ? a->sig : a->back, ti, E.I[k], IPDFtest, IPDF);
                set_t(k, ti, tx, ty);
                PrI = PrI * IPDFtest / IPDF;
                XI0 = XI0 + E.I[k] * E.X[k];
                YI0 = YI0 + E.I[k] * E.Y[k];
                XII0 = XII0 + ti * tx;
                YII0 = YII0 + ti * ty;
                II0 = II0 + ti;
                delta_push(0, k, true);
                delta_push(1, k, false);
            }
            const float tin = I0 - II0;
            const float txn = (XI0 + E.I[nn] * E.X[nn] - XII0) / tin;
            const float tyn = (YI0 + E.I[nn] * E.Y[nn] - YII0) / tin;
            set_t(nn, tin, txn, tyn);
            const bool fg = E.S[nn] == 1;
            findpdf2(fg ? a->sig : a->back, tin, E.I[nn], IPDFtest, IPDF);
            PrI = PrI * IPDFtest / IPDF;
            delta_push(0, nn, true);
            delta_push(1, nn, false);
        } else if (Ub >= 0.5 && Ub < (LRN && a->mode == BAMF_RJ_LEARNED ? a->l_single_from : 0.75f)) {
            jt = 3;
            int np = closest_multi(nn, 2 * p.psf_sigma);
            if (lane == 0) L.nb[np] = (uint8_t)nn;
            np++;
            wsync();
            for (int ii = 0; ii < np; ii++) {
                const int k = L.nb[ii];
                const bool fg = E.S[k] == 1;
                cnt_inc(fg ? C_NFG + J_MOVE : C_NBG + J_MOVE);
                const float is = fg ? p.i_std_fg : p.i_std_bg, xs = fg ? p.x_std_fg : p.x_std_bg;
                need(36);
                const float ti = T.I[k] + is * randn();
                const float tx = T.X[k] + xs * randn();
                const float ty = T.Y[k] + xs * randn();
                findpdf2(fg ? a->sig : a->back, ti, E.I[k], IPDFtest, IPDF);
                wsync();
                set_t(k, ti, tx, ty);
                wsync();
                PrI = PrI * IPDFtest / IPDF;
                delta_push(0, k, true);
                delta_push(1, k, false);
            }
        } else {
            const bool fg = E.S[nn] == 1;
            cnt_inc(fg ? C_NFG + J_MOVE : C_NBG + J_MOVE);
            const float is = fg ? p.i_std_fg : p.i_std_bg, xs = fg ? p.x_std_fg : p.x_std_bg;
            // Learned mode: with probability beta the intensity is redrawn
            // from the network's log-uniform window at the emitter's pixel
            // (an independence proposal; its own Metropolis-Hastings kernel,
            // chosen independently of the state), the position walks as usual.
            const bool indep = LRN && a->mode == BAMF_RJ_LEARNED && urand() < a->l_beta;
            const bool cond = LRN && indep && a->l_ikind == 1;
            float ti, qratio = 1.0f, cm = 0.0f, cs = 1.0f;
            if constexpr (LRN) {
                if (cond) {
                    // intensity from the Gaussian fit of this emitter's conditional
                    // likelihood at its position, the rest of the model held fixed
                    cond_prep(T.I[nn], T.X[nn], T.Y[nn]);
                    cond_fit(cm, cs);
                }
            }
            if (cond) {
                ti = cm + cs * randn();
            } else if (indep) {
                const float w = a->l_halfwidth;
                const float mu0 = map_mu(T.X[nn], T.Y[nn]);
                ti = expf(mu0 + w * (2.0f * urand() - 1.0f));
                if (!(mu0 > -INFINITY)) return EV_NONE;  // outside the ROI: no proposal
            } else {
                ti = T.I[nn] + is * randn();
            }
            const float tx = T.X[nn] + xs * randn();
            const float ty = T.Y[nn] + xs * randn();
            if (cond) {
                // q(I | x') / q(I' | x): the fit at the new position scores the
                // way back (the rest of the model is the same in both directions)
                float rm = 0.0f, rs = 1.0f;
                if constexpr (LRN) {
                    cond_stencil(tx, ty);
                    cond_fit(rm, rs);
                }
                const float zf = (ti - cm) / cs, zr = (T.I[nn] - rm) / rs;
                qratio = (cs / rs) * expf(0.5f * (zf * zf - zr * zr));
            } else if (indep) {
                // q(I | x') / q(I' | x) of the log-uniform windows: I' / I when
                // the old intensity lies in the window at the new pixel, else 0
                const float mu1 = map_mu(tx, ty);
                const float I0 = T.I[nn];
                qratio = (I0 > 0.0f && fabsf(logf(I0) - mu1) <= a->l_halfwidth) ? ti / I0 : 0.0f;
            }
            findpdf2(fg ? a->sig : a->back, ti, E.I[nn], IPDFtest, IPDF);
            set_t(nn, ti, tx, ty);
            PrI = IPDFtest / IPDF * qratio;  // qratio = 1 outside the independence kernel
            delta_push(0, nn, true);
            delta_push(1, nn, false);
        }
        wsync();
        f_pri = PrI;
        acc_c1 = C_A + J_MOVE;
        acc_c2 = E.S[nn] == 1 ? C_AFG + J_MOVE : C_ABG + J_MOVE;
        need_test = true;
        return EV_MOVE;
    }
    // LibRJMCMC.cpp:918-936
    __device__ __forceinline__ float fin_move(float sum, float LLR) {
        const float BgPDF = cbg_pdf;  // findpdf(OffsetStat, BG) of the current state
        float PrBg;
        if (bgdelta) {
            const float BgPDFtest = findpdf(a->off, a->p.dx, tBG);
            tbg_pdf = BgPDFtest;
            PrBg = (BgPDFtest / BgPDF) * expf(-(tBBG * tBBG - cBBG * cBBG) / 2) *
                   expf(-(tABG * tABG - cABG * cABG) / 2);
        } else {
            // BG, ABG, BBG unchanged: both exponentials are expf(-0) = 1, so the
            // factor is BgPDF / BgPDF (1, or NaN for a zero pdf), bit for bit
            // (x / x: 1 for finite nonzero x, NaN for 0, infinities and NaN)
            PrBg = (fabsf(BgPDF) > 0.0f && fabsf(BgPDF) < INFINITY) ? 1.0f : __builtin_nanf("");
        }
        lp_prod = (double)f_pri * (double)PrBg;
        lp_neg = false;
        return LLR * f_pri * PrBg;
    }

    // LibRJMCMC.cpp:972-1028
    __device__ __forceinline__ int prop_split() {
        const bamf_mcmc_params& p = a->p;
        jt = 4;
        cnt_inc(C_N + J_SPLIT);
        copy_to_test();
        if (cN < 1) return EV_NONE;
        const Em& E = cur();
        const float Ua = urand();
        const int jj = min((int)floorf((float)cN * Ua), cN - 1);
        const float Ijp = E.I[jj], mux = E.X[jj], muy = E.Y[jj];
        const int BF = E.S[jj];
        cnt_inc(BF == 1 ? C_NFG + J_SPLIT : C_NBG + J_SPLIT);
        const float u1 = urand();
        float u2, u3;
        f_mz = 0.0f;
        if (learned_split()) {
            // the first child from the map inside the window around the parent,
            // with probability l_split; else the reference's Gaussian offsets
            const float r = a->l_split_r;
            const bool from_map = urand() < a->l_split;
            int pix = -1;
            const float Z = win_scan(mux, muy, r, from_map, urand(), pix);
            if (from_map && pix >= 0) {
                u2 = ((float)(pix / Y) + urand() - 0.5f) - mux;
                u3 = ((float)(pix - Y * (pix / Y)) + urand() - 0.5f) - muy;
            } else {
                u2 = p.split_std * randn();
                u3 = p.split_std * randn();
            }
            f_mz = win_density(mux + u2, muy + u3, mux, muy, r, Z);
        } else {
            u2 = p.split_std * randn();
            u3 = p.split_std * randn();
        }
        set_t(jj, u1 * Ijp, mux + u2, muy + u3);
        if (!add_t((1 - u1) * Ijp, mux - u1 * u2 / (1 - u1), muy - u1 * u3 / (1 - u1), BF)) return EV_NONE;
        delta_push(0, jj, true);
        delta_push(1, jj, false);
        delta_push(1, tN - 1, false);
        f_jj = jj;
        f_u1 = u1;
        f_u2 = u2;
        f_u3 = u3;
        f_ijp = Ijp;
        acc_c1 = C_A + J_SPLIT;
        acc_c2 = BF == 1 ? C_AFG + J_SPLIT : C_ABG + J_SPLIT;
        need_test = true;
        return EV_SPLIT;
    }
    // LibRJMCMC.cpp:1030-1070
    __device__ __forceinline__ float fin_split(float sum, float LLR) {
        const bamf_mcmc_params& p = a->p;
        const Em& E = cur();
        const Em& T = tst();
        // both children keep the parent's Signal label: one table for all three
        float tmp1 = 1, tmp2 = 1, pe, pt0, pt1;
        findpdf3(E.S[f_jj] == 1 ? a->sig : a->back, E.I[f_jj], T.I[f_jj], T.I[tN - 1], pe, pt0, pt1);
        tmp1 = tmp1 * pe;
        tmp2 = tmp2 * pt0;
        tmp2 = tmp2 * pt1;
        const float Norm1 = normpdf_split(f_u2);
        const float Norm2 = normpdf_split(f_u3);
        const float NEm = (float)nem_cur();
        const float PrM = meanN / (NEm + 1);
        const float PrI = tmp2 / tmp1;
        const float Area = xb * yb;
        const float PrX = L.kc[0];  // 1 / Area
        const float Us = split_us(Norm1 * Norm2);  // 1 / (Norm1 * Norm2) outside the learned split
        const double om = (double)(1 - f_u1);
        const float Jacob = (float)((double)f_ijp / (om * om));
        const float P12 = L.kc[1];  // P1 / P2
        lp_prod = (double)PrI * (double)PrX * (double)Us * (double)Jacob * (double)P12 * (double)PrM;
        lp_neg = false;
        return LLR * PrI * PrX * Us * Jacob * P12 * PrM;
    }

    // LibRJMCMC.cpp:1111-1164
    __device__ __forceinline__ int prop_merge() {
        const bamf_mcmc_params& p = a->p;
        jt = 5;
        cnt_inc(C_N + J_MERGE);
        copy_to_test();
        if (cN < 2) {
            update_grids();
            return EV_NONE;
        }
        const Em& E = cur();
        const float Ua = urand();
        const int jj = min((int)floorf((float)cN * Ua), cN - 1);
        const int ii = closest(jj, 3 * p.psf_sigma);
        cnt_inc(E.S[jj] == 1 ? C_NFG + J_MERGE : C_NBG + J_MERGE);
        {
            const double dx = (double)(E.X[ii] - E.X[jj]), dy = (double)(E.Y[ii] - E.Y[jj]);
            const float dd = (float)(dx * dx + dy * dy);
            if (sqrtf(dd) > 3 * p.psf_sigma) {
                update_grids();
                return EV_NONE;
            }
        }
        const float Ijp = E.I[jj] + E.I[ii];
        const float muxjp = (E.I[jj] * E.X[jj] + E.I[ii] * E.X[ii]) / Ijp;
        const float muyjp = (E.I[jj] * E.Y[jj] + E.I[ii] * E.Y[ii]) / Ijp;
        f_u1 = E.I[jj] / Ijp;
        f_u2 = E.X[jj] - muxjp;
        f_u3 = E.Y[jj] - muyjp;
        f_ijp = Ijp;
        f_mz = 0.0f;
        if (learned_split()) {  // the split that undoes this merge: window around the merged emitter
            int pix;
            const float r = a->l_split_r;
            const float Z = win_scan(muxjp, muyjp, r, false, 0.0f, pix);
            f_mz = win_density(E.X[jj], E.Y[jj], muxjp, muyjp, r, Z);
        }
        const float PDFfg = findpdf(a->sig, p.dx, Ijp);
        const float PDFbg = findpdf(a->back, p.dx, Ijp);
        set_t(jj, Ijp, muxjp, muyjp);
        if (lane == 0) tst().S[jj] = (PDFbg > PDFfg) ? 0 : 1;
        remove_t(ii);
        f_jj = jj;
        f_jm = ii > jj ? jj : jj - 1;  // the merged emitter's index in the proposed set
        L.nb[0] = (uint8_t)ii;          // keep ii for the prior ratio (same value from every lane)
        delta_push(0, jj, true);
        delta_push(0, ii, true);
        delta_push(1, f_jm, false);
        acc_c1 = C_A + J_MERGE;
        acc_c2 = E.S[jj] == 1 ? C_AFG + J_MERGE : C_ABG + J_MERGE;
        need_test = true;
        wsync();
        return EV_MERGE;
    }
    // LibRJMCMC.cpp:1166-1217
    __device__ __forceinline__ float fin_merge(float sum, float LLR) {
        const bamf_mcmc_params& p = a->p;
        const Em& E = cur();
        const Em& T = tst();
        const int ii = L.nb[0];
        float tmp1 = 1, tmp2 = 1, pj, pi, pm;
        if (E.S[f_jj] == E.S[ii] && T.S[f_jm] == E.S[f_jj]) {
            findpdf3(E.S[f_jj] == 1 ? a->sig : a->back, E.I[f_jj], E.I[ii], T.I[f_jm], pj, pi, pm);
        } else {
            pj = pdf_of(E, f_jj);
            pi = pdf_of(E, ii);
            pm = pdf_of(T, f_jm);
        }
        tmp1 = tmp1 * pj;
        tmp1 = tmp1 * pi;
        tmp2 = tmp2 * pm;
        const float Norm1 = normpdf_split(f_u2);
        const float Norm2 = normpdf_split(f_u3);
        const float NEm = (float)nem_cur();
        const float PrI = tmp1 / tmp2;
        const float Area = xb * yb;
        const float PrX = L.kc[0];  // 1 / Area
        const float Us = split_us(Norm1 * Norm2);
        const double om = (double)(1 - f_u1);
        const float Jacob = (float)((double)f_ijp / (om * om));
        const float PrM = meanN / NEm;
        lp_prod = (double)PrI * (double)PrX * (double)Jacob * (double)Us * (double)PrM * (double)P1 / (double)P2;
        lp_neg = true;
        return LLR / (PrI * PrX * Jacob * Us * PrM * P1 / P2);
    }

    // gSplit, LibRJMCMC.cpp:1389-1436: the picked emitter and a random subset
    // of its neighbours within 2 sigma give up the fraction u3 of their
    // intensity to a new emitter near their centroid.  No bounds test and no
    // acceptance counters.
    __device__ __forceinline__ int prop_gsplit() {
        const bamf_mcmc_params& p = a->p;
        jt = 8;
        copy_to_test();
        if (cN < 2) return EV_NONE;
        const Em& E = cur();
        const float Ua = urand();
        const int jj = min((int)floorf((float)cN * Ua), cN - 1);
        int np = closest_multi(jj, 2 * p.psf_sigma);
        if (lane == 0) L.nb[np] = (uint8_t)jj;
        np++;
        wsync();
        if (np == 1) return EV_NONE;
        float I0 = 0, X0 = 0, Y0 = 0;
        for (int n = 0; n < np; n++) {
            const int k = L.nb[n];
            I0 = I0 + E.I[k];
            X0 = X0 + E.X[k];
            Y0 = Y0 + E.Y[k];
        }
        need(25);
        const float u1 = 0.3f * randn();
        const float u2 = 0.3f * randn();
        const float u3 = (1 / (float)np) * urand();
        const float fnp = (float)np;
        // the intensity prior ratio needs no likelihood: formed here; the
        // proposed set's Signal is read at n, not at Closest[n] (LibRJMCMC.cpp:1447)
        float PrI = 1;
        for (int n = 0; n < np; n++) {
            const int k = L.nb[n];
            const float ti = (1 - u3) * E.I[k];
            const float tmpCurrent = pdf_of(E, k);
            const float tmpProposed = findpdf(E.S[n] == 1 ? a->sig : a->back, p.dx, ti);
            PrI = PrI * tmpProposed / tmpCurrent;
            set_t(k, ti, (E.X[k] - u3 * (X0 / fnp + u1)) / (1 - u3), (E.Y[k] - u3 * (Y0 / fnp + u2)) / (1 - u3));
            delta_push(0, k, true);
            delta_push(1, k, false);
        }
        wsync();
        const float In = u3 * I0;
        if (!add_t(In, X0 / fnp + u1, Y0 / fnp + u2, 1)) return EV_NONE;
        delta_push(1, tN - 1, false);
        f_pri = PrI * findpdf(a->sig, p.dx, In);
        f_u1 = u1;
        f_u2 = u2;
        f_u3 = u3;
        f_ijp = I0;
        f_jj = np;
        acc_c1 = acc_c2 = -1;
        need_test = false;
        return EV_GSPLIT;
    }
    // LibRJMCMC.cpp:1454-1472
    __device__ __forceinline__ float fin_gsplit(float sum, float LLR) {
        const float PDFu1 = normpdf(f_u1, 0.0f, 0.4f);
        const float PDFu2 = normpdf(f_u2, 0.0f, 0.4f);
        const float PDFu3 = (float)f_jj;
        const float Us = 1 / (PDFu1 * PDFu2 * PDFu3);
        const float NEm = (float)nem_cur();
        const float PrM = meanN / (NEm + 1);
        const float PrX = 1 / (xb * yb);
        const float Jacob = f_ijp / powf(1 - f_u3, (float)(f_jj + 1));
        lp_prod = (double)f_pri * (double)PrX * (double)Us * (double)Jacob * (double)PrM * (double)P56;
        lp_neg = false;
        return LLR * f_pri * PrX * Us * Jacob * PrM * P56;
    }

    // gMerge, LibRJMCMC.cpp:1487-1551: the picked emitter is absorbed by a
    // random subset of its neighbours within 3 sigma.  No bounds test and no
    // acceptance counters.
    __device__ __forceinline__ int prop_gmerge() {
        const bamf_mcmc_params& p = a->p;
        jt = 9;
        copy_to_test();
        if (cN < 3) return EV_NONE;
        const Em& E = cur();
        const float Ua = urand();
        const int jj = min((int)floorf((float)cN * Ua), cN - 1);
        const int np = closest_multi(jj, 3 * p.psf_sigma);
        if (np == 0) return EV_NONE;
        float I0 = E.I[jj];
        for (int n = 0; n < np; n++) I0 = I0 + E.I[L.nb[n]];
        const float U3 = E.I[jj] / I0;
        const float xj = E.X[jj], yj = E.Y[jj];
        float XX0 = 0, YY0 = 0, PrI = 1;
        delta_push(0, jj, true);
        for (int n = 0; n < np; n++) {
            const int k = L.nb[n];
            const float ti = E.I[k] / (1 - U3);
            const float tx = U3 * xj + (1 - U3) * E.X[k];
            const float ty = U3 * yj + (1 - U3) * E.Y[k];
            XX0 = XX0 + U3 * xj + (1 - U3) * E.X[k];
            YY0 = YY0 + U3 * yj + (1 - U3) * E.Y[k];
            const bool fg = E.S[k] == 1;
            const float tmpProposed = findpdf(fg ? a->sig : a->back, p.dx, ti);
            const float tmpCurrent = pdf_of(E, k);
            PrI = PrI * (tmpCurrent / tmpProposed);
            set_t(k, ti, tx, ty);
            delta_push(0, k, true);
            delta_push(1, k > jj ? k - 1 : k, false);  // its index once jj is removed
        }
        f_pri = PrI * pdf_of(E, jj);
        f_u1 = xj - XX0 / (float)np;
        f_u2 = yj - YY0 / (float)np;
        f_u3 = U3;
        f_ijp = I0;
        f_jj = np;
        remove_t(jj);
        acc_c1 = acc_c2 = -1;
        need_test = false;
        return EV_GMERGE;
    }
    // LibRJMCMC.cpp:1553-1571
    __device__ __forceinline__ float fin_gmerge(float sum, float LLR) {
        const float PDFu1 = normpdf(f_u1, 0.0f, 0.4f);
        const float PDFu2 = normpdf(f_u2, 0.0f, 0.4f);
        const float PDFu3 = (float)f_jj;
        const float Us = 1 / (PDFu1 * PDFu2 * PDFu3);
        const float NEm = (float)nem_cur();
        const float PrM = NEm / meanN;
        const float PrX = 1 / (xb * yb);
        const float Jacob = f_ijp / powf(1 - f_u3, (float)(f_jj + 1));
        lp_prod = (double)f_pri * (double)PrX * (double)Us * (double)Jacob * (double)PrM * (double)P56;
        lp_neg = true;
        if (PDFu1 == 0 || PDFu2 == 0) return -1.0f;
        return LLR / (f_pri * PrX * Us * Jacob * PrM * P56);
    }

    // find_Birth (LibRJMCMC.cpp:1688-1736): residual max(0, D - M), prefix
    // sum in pixel order, r = total * U, first pixel with r <= cumulative.
    __device__ __forceinline__ int birth_pixel(float u) {
        aflops += (uint64_t)(4 * P);
        // every slot's scan on its own (no dependence between slots, so the
        // scans overlap), then the slot totals carried in slot order: lane 63
        // of carry + scan is carry + total, the same sum as the running form
        float cum[PPL];
#pragma unroll
        for (int s = 0; s < PPL; s++) {
            const float r = L.dpx[64 * s + lane] - m[s];
            const float v = (0 > r) ? 0.0f : r;
            cum[s] = wave_scan(s < ppl ? v : 0.0f);
        }
        float carry = 0.0f;
#pragma unroll
        for (int s = 0; s < PPL; s++) {
            if (s < ppl) {
                const float tot = bcastf(cum[s], 63);
                cum[s] = carry + cum[s];
                carry = carry + tot;
            }
        }
        const float rr = carry * u;
        int pix = P;  // (padding lanes add 0: D - M = -1)
#pragma unroll
        for (int s = 0; s < PPL; s++) {
            if (s < ppl && pix == P) {
                const uint64_t mk = ballot(has(s) && rr <= cum[s]);
                if (mk) pix = bcasti(pidx(s), __builtin_ctzll(mk));
            }
        }
        return pix;
    }
    // Learned proposal: pixel drawn from the per-ROI proposal map q; returns
    // the pixel and its probability q[pix] / sum(q).  The running sums of q
    // in pixel order depend on the ROI only: map_sums forms them once when the
    // chain starts the ROI (slot scans carried in slot order) into qcum, a
    // pixel whose q is not positive holding -inf so that no draw selects it;
    // q_total is the sum.
    float q_total;
    __device__ __forceinline__ float* qcum_of() const { return a->qcum + (int64_t)roi * (64 * kMaxPPL); }
    __device__ __forceinline__ void map_sums() {
        const int64_t off = a->pix_off[roi];
        float* qc = qcum_of();
        float carry = 0.0f;
        for (int s = 0; s < ppl; s++) {
            const float v = has(s) ? a->prop_map[off + pidx(s)] : 0.0f;
            const float cum = carry + wave_scan(v);
            carry = bcastf(cum, 63);
            qc[64 * s + lane] = (has(s) && v > 0.0f) ? cum : -INFINITY;
        }
        q_total = carry;
    }
    __device__ __forceinline__ int map_pixel(float u, float& prob) {
        const float* qc = qcum_of();
        const float carry = q_total;
        const float rr = carry * u;
        int pix = P - 1;
        float qp = 0.0f;
        bool found = false;
        for (int s = 0; s < ppl && !found; s++) {
            const uint64_t mk = ballot(has(s) && rr <= qc[64 * s + lane]);
            if (mk) {
                pix = bcasti(pidx(s), __builtin_ctzll(mk));
                found = true;
            }
        }
        if (found) qp = a->prop_map[a->pix_off[roi] + pix];
        prob = carry > 0.0f ? qp / carry : 0.0f;
        return pix;
    }

    // Learned split window: the pixels whose centres lie within r of (cx, cy)
    // along both axes.  Returns Z, the occupancy map's mass over the window;
    // with pick, also draws a pixel of the window with probability q / Z
    // (u: the uniform) into pix.
    __device__ __forceinline__ float win_scan(float cx, float cy, float r, bool pick, float u, int& pix) {
        const int64_t off = a->pix_off[roi];
        float carry = 0.0f;
        float cum[PPL], q[PPL];
#pragma unroll
        for (int s = 0; s < PPL; s++) {
            cum[s] = 0.0f;
            q[s] = 0.0f;
            if (s < ppl) {
                const bool in = has(s) && fabsf((float)px(s) - cx) <= r && fabsf((float)ysub - cy) <= r;
                float v = in ? a->prop_map[off + pidx(s)] : 0.0f;
                q[s] = v;
                v = wave_scan(v);
                cum[s] = carry + v;
                carry = bcastf(cum[s], 63);
            }
        }
        pix = -1;
        if (pick && carry > 0.0f) {
            const float rr = carry * u;
            bool found = false;
#pragma unroll
            for (int s = 0; s < PPL; s++) {
                if (s < ppl && !found) {
                    const uint64_t mk = ballot(rr <= cum[s] && q[s] > 0.0f);
                    if (mk) {
                        pix = bcasti(pidx(s), __builtin_ctzll(mk));
                        found = true;
                    }
                }
            }
        }
        return carry;
    }
    // q_map(pixel of (xc, yc)) / Z for the window of half-width r around (cx, cy),
    // 0 outside it; -1 when the window holds no mass (the map draw then falls
    // back to the Gaussian offsets, whose density is then the whole proposal)
    __device__ __forceinline__ float win_density(float xc, float yc, float cx, float cy, float r, float Z) const {
        if (!(Z > 0.0f)) return -1.0f;
        const int qx = (int)floorf(xc + 0.5f), qy = (int)floorf(yc + 0.5f);
        if (qx < 0 || qx >= X || qy < 0 || qy >= Y) return 0.0f;
        if (!(fabsf((float)qx - cx) <= r && fabsf((float)qy - cy) <= r)) return 0.0f;
        return a->prop_map[a->pix_off[roi] + qx * Y + qy] / Z;
    }
    __device__ __forceinline__ bool learned_split() const {
        return LRN && a->mode == BAMF_RJ_LEARNED && a->prop_map != nullptr && a->l_split > 0.0f;
    }
    // 1 / q(u2, u3): the split's offsets scored by the proposal that drew them
    __device__ __forceinline__ float split_us(float nn) const {
        if (!learned_split() || f_mz < 0.0f) return 1 / nn;
        return 1 / (a->l_split * f_mz + (1 - a->l_split) * nn);
    }

    // LibRJMCMC.cpp:1258-1295
    __device__ __forceinline__ int prop_birth() {
        const bamf_mcmc_params& p = a->p;
        jt = 6;
        cnt_inc(C_N + J_BIRTH);
        copy_to_test();
        const float ur = urand();
        float pixprob = 1.0f;
        int pix;
        // learned births need the network's map; without it births are the reference's
        const bool learned = LRN && a->mode == BAMF_RJ_LEARNED && a->prop_map != nullptr;
        if constexpr (LRN) {
            if (learned) pix = map_pixel(ur, pixprob);
            else pix = birth_pixel(ur);
        } else {
            pix = birth_pixel(ur);
        }
        r_pix = pix;
        if (a->mode == BAMF_RJ_REPLAY) pix = ref_pix;
        const float y = (float)(pix - Y * (pix / Y)) + urand() - 0.5f;
        const float x = (float)(int)floorf(((float)pix) / Y) + urand() - 0.5f;
        int BF = 1;
        if (p.is_backg == 1) BF = kround(urand());
        cnt_inc(BF == 1 ? C_NFG + J_BIRTH : C_NBG + J_BIRTH);
        float I_new, iratio = 1.0f;
        if (learned) {
            // intensity from the mixture alpha * prior + (1 - alpha) *
            // log-uniform on [mu - w, mu + w] around the network's value at
            // the pixel; the birth ratio carries prior(I) / q(I) and PixProb
            // is the pixel's proposal probability relative to uniform (P q)
            const float um = urand(), ui = urand();
            const float mu = a->prop_mu[a->pix_off[roi] + pix];
            const float w = a->l_halfwidth;
            I_new = um < a->l_alpha ? genrand(a->sig, ui) : expf(mu + w * (2.0f * ui - 1.0f));
            const float prior = findpdf(a->sig, p.dx, I_new) * a->l_zinv;
            const float lg = logf(I_new);
            const float win = (I_new > 0.0f && fabsf(lg - mu) <= w) ? 1.0f / (2.0f * w * I_new) : 0.0f;
            const float q = a->l_alpha * prior + (1.0f - a->l_alpha) * win;
            iratio = prior / q;
            pixprob = (float)P * pixprob;
        } else {
            I_new = genrand(BF == 1 ? a->sig : a->back, urand());
        }
        if (!add_t(I_new, x, y, BF)) return EV_NONE;
        delta_push(1, tN - 1, false);
        f_pixprob = pixprob;
        f_iratio = iratio;
        acc_c1 = C_A + J_BIRTH;
        acc_c2 = BF == 1 ? C_AFG + J_BIRTH : C_ABG + J_BIRTH;
        need_test = true;
        return EV_BIRTH;
    }
    // LibRJMCMC.cpp:1297-1302
    __device__ __forceinline__ float fin_birth(float sum, float LLR) {
        const float NEm = (float)nem_cur();
        const float PrM = meanN / (NEm + 1);
        lp_prod = (double)P3 / ((double)P4 * (double)xb * (double)yb * (double)f_pixprob) * (double)PrM *
                  (double)f_iratio;
        lp_neg = false;
        return LLR * P3 / (P4 * xb * yb * f_pixprob) * PrM * f_iratio;  // f_iratio = 1 outside learned mode
    }

    // LibRJMCMC.cpp:1332-1355
    __device__ __forceinline__ int prop_death() {
        jt = 7;
        cnt_inc(C_N + J_DEATH);
        copy_to_test();
        if (cN < 1) return EV_NONE;
        const Em& E = cur();
        const float Ua = urand();
        const int jj = min((int)floorf((float)cN * Ua), cN - 1);
        const int sj = E.S[jj];
        cnt_inc(sj == 1 ? C_NFG + J_DEATH : C_NBG + J_DEATH);
        remove_t(jj);
        delta_push(0, jj, true);
        acc_c1 = C_A + J_DEATH;
        acc_c2 = sj == 1 ? C_AFG + J_DEATH : C_ABG + J_DEATH;
        need_test = false;  // death has no bounds test
        return EV_DEATH;
    }
    // LibRJMCMC.cpp:1356-1360
    __device__ __forceinline__ float fin_death(float sum, float LLR) {
        const float NEm = (float)nem_cur();
        const float PrM = NEm / meanN;
        const float P43 = L.kc[2];  // P4 / P3
        lp_prod = (double)P43 * (double)PrM;
        lp_neg = false;
        return LLR * P43 * PrM;
    }

    // LibRJMCMC.cpp:1586-1627 (changes the current set in place, no record)
    __device__ __forceinline__ void do_backfore() {
        if (cN == 0) return;
        if (a->p.is_backg == 0) return;
        Em& E = cur();
        const float Ua = urand();
        const int nn = min((int)floorf((float)cN * Ua), cN - 1);
        const float BackP = findpdf(a->back, a->p.dx, E.I[nn]);
        const float SignalP = findpdf(a->sig, a->p.dx, E.I[nn]);
        int s = E.S[nn];
        if (s == 1) {
            if (SignalP != 0) {
                if (BackP / SignalP > urand()) s = 0;
            } else {
                s = 0;
            }
        } else {
            if (BackP != 0) {
                if (SignalP / BackP > urand()) s = 1;
            } else {
                s = 1;
            }
        }
        wsync();
        if (lane == 0) E.S[nn] = s;
        wsync();
    }

    __device__ __forceinline__ int kind_of(float rtmp, const float* PV, const float* PVB) {
        if (rtmp < PV[0]) return 0;
        if (rtmp >= PV[0] && rtmp < PV[1]) return 1;
        if (rtmp >= PV[1] && rtmp < PV[2]) return 2;
        if (rtmp >= PV[2] && rtmp < PV[3]) return 3;
        if (rtmp >= PV[3] && rtmp < PV[4]) return 4;
        if (rtmp >= PVB[4] && rtmp < PVB[5]) return 5;
        if (rtmp >= PVB[5] && rtmp < PVB[6]) return 6;
        return 7;
    }

    // One jump: proposal, then (when the jump reaches it) the shared
    // likelihood evaluation, the posterior ratio, the accept test, the
    // records and the grid update, in the reference's order.
    __device__ __forceinline__ int one_step() {
        const float rtmp = urand();
        const int kind = burn ? kind_of(rtmp, a->pvb, a->pvb) : kind_of(rtmp, a->pvt, a->pvb);
        int evk = EV_NONE;
        switch (kind) {
            case 0: evk = prop_move(); break;
            case 1: evk = prop_split(); break;
            case 2: evk = prop_merge(); break;
            case 3: evk = prop_birth(); break;
            case 4: evk = prop_death(); break;
            case 7: do_backfore(); break;
            default:  // 5 gSplit, 6 gMerge
                if constexpr (GJ) evk = kind == 5 ? prop_gsplit() : prop_gmerge();
                else status = ST_UNSUPPORTED;
                break;
        }
        applied = false;
        if (evk != EV_NONE && status == ST_OK) {
            if (need_test && !test_t()) {
                r_flags |= 8;
            } else {
                const float sum = proposal_llr_sum();
                const float LLR = expf(sum);
                r_llr = LLR;
                float PR;
                switch (evk) {
                    case EV_MOVE: PR = fin_move(sum, LLR); break;
                    case EV_SPLIT: PR = fin_split(sum, LLR); break;
                    case EV_MERGE: PR = fin_merge(sum, LLR); break;
                    case EV_BIRTH: PR = fin_birth(sum, LLR); break;
                    default:
                        if constexpr (GJ) {
                            PR = evk == EV_GSPLIT   ? fin_gsplit(sum, LLR)
                                 : evk == EV_GMERGE ? fin_gmerge(sum, LLR)
                                                    : fin_death(sum, LLR);
                        } else {
                            PR = fin_death(sum, LLR);
                        }
                        break;
                }
                const bool acc = decide(PR, sum);
                applied = acc;
                // kept for step records and the accepted-jump total (each
                // formed only when its output is requested)
                double logpr = __builtin_nan("");
                if ((acc && a->logpost) || a->steps) logpr = logpr_of(sum);
                r_logpr = logpr;
                if (acc) {
                    if (!GJ || acc_c1 >= 0) {  // gSplit/gMerge keep no counters
                        cnt_inc(acc_c1);
                        cnt_inc(acc_c2);
                    }
                    ci ^= 1;
                    cN = tN;
                    if (bgdelta) cbg_pdf = tbg_pdf;  // only a background move changes BG
                    cBG = tBG;
                    cABG = tABG;
                    cBBG = tBBG;
#pragma unroll
                    for (int s = 0; s < PPL; s++)
                        if (s >= lo_s && s <= hi_s) m[s] = mt[s];
                    // the proposed set passed the test, or (no test: death, and
                    // gSplit/gMerge) passes it when the current set does and no
                    // value written into it fails it; death writes none
                    if constexpr (GJ) e_ok = need_test || (e_ok && !t_bad);
                    else if (need_test) e_ok = true;
                    logpost += logpr;
                    if (++nacc >= a->recompute_every) {
                        wsync();
                        rebuild_model();
                    }
                }
            }
            record();
            update_grids();
        }
        return kind;
    }

    __device__ __forceinline__ float total_photons_seq() {
        // sequential single-precision sum (the oracle's order) for step records
        float ph = 0.0f;
        const Em& c = cur();
        for (int k = 0; k < cN; k++) ph = ph + c.I[k];
        return ph;
    }

    // ---------------- segments ----------------
    __device__ __forceinline__ void save_state(SegState& g) {
        const Em& c = cur();
        for (int n = lane; n < cN; n += 64) {
            g.I[n] = c.I[n];
            g.X[n] = c.X[n];
            g.Y[n] = c.Y[n];
            g.S[n] = c.S[n];
        }
#pragma unroll
        for (int s = 0; s < PPL; s++)
            if (s < ppl) g.m[64 * s + lane] = m[s];
        if (lane < 32) g.cnt[lane] = L.cnt[lane];
        for (int k = lane; k < kNMax + 1; k += 64) g.hist[k] = L.hist[k];
        if (lane < 3) g.rec[lane] = L.rec[lane];
        if (lane == 0) {
            g.logpost = logpost;
            g.aflops = aflops;
            g.tcur = tcur;
            g.status = status;
            g.n = cN;
            g.e_ok = e_ok ? 1 : 0;
            g.nacc = nacc;
            g.bg = cBG;
            g.abg = cABG;
            g.bbg = cBBG;
        }
    }
    __device__ __forceinline__ void load_state(const SegState& g) {
        logpost = g.logpost;
        aflops = g.aflops;
        tcur = g.tcur;
        status = g.status;
        cN = g.n;
        e_ok = g.e_ok != 0;
        nacc = g.nacc;
        cBG = g.bg;
        cbg_pdf = findpdf(a->off, a->p.dx, cBG);
        cABG = g.abg;
        cBBG = g.bbg;
        Em& c = L.e[0];
        for (int n = lane; n < cN; n += 64) {
            c.I[n] = g.I[n];
            c.X[n] = g.X[n];
            c.Y[n] = g.Y[n];
            c.S[n] = g.S[n];
        }
#pragma unroll
        for (int s = 0; s < PPL; s++)
            if (s < ppl) m[s] = g.m[64 * s + lane];
        if (lane < 32) L.cnt[lane] = g.cnt[lane];
        for (int k = lane; k < kNMax + 1; k += 64) L.hist[k] = g.hist[k];
        if (lane < 3) L.rec[lane] = g.rec[lane];
        wsync();
    }

    __device__ __forceinline__ void run(int roi_) {
        // the ROI index and sizes in scalar registers (they arrive by vector
        // loads): the slot bookkeeping derived from them then runs on the
        // scalar unit
        roi = __builtin_amdgcn_readfirstlane(roi_);
        const bamf_mcmc_params& p = a->p;
        X = __builtin_amdgcn_readfirstlane(a->xs[roi]);
        Y = __builtin_amdgcn_readfirstlane(a->ys[roi]);
        if (X < 1 || Y < 1 || X > kMaxSide || Y > kMaxSide || slots_for(X, Y) > PPL) {
            if (a->status && lane == 0) a->status[roi] = ST_SIZE;
            return;
        }
        P = X * Y;
        lgY = 32 - __builtin_clz((unsigned)(Y - 1) | 1u);  // ceil(log2(Y)), 0 for Y = 1
        if (Y == 1) lgY = 0;
        cps = 64 >> lgY;
        ppl = slots_for(X, Y);
        xsub = lane >> lgY;
        ysub = lane & ((1 << lgY) - 1);
        C = 1 / (p.psf_sigma * sqrtf(2.0f));
        const float rho = a->rho ? a->rho[roi] : p.rho;
        meanN = rho * X * Y;  // cRJMCMC.cpp:116
        bnd = (int)p.bnd_out;
        xb = X + p.bnd_out;
        yb = Y + p.bnd_out;
        if (lane == 0) {
            const float Area = xb * yb;
            L.kc[0] = 1 / Area;
            const float sa = (float)sqrt(2.0f * kPiD), ss = p.split_std;
            L.kc[3] = 1.0f / (sa * ss);
            L.kc[4] = 2 * ss * ss;
        }
        // (read into scalar registers: the Philox key schedule then runs on the
        // scalar unit)
        key = __builtin_amdgcn_readfirstlane(a->key ? a->key[roi] : (uint32_t)roi);
        const int nb = p.n_burnin, total = p.n_burnin + p.n_trials;
        int s_begin = 0, s_end = total;
        SegState* ss = nullptr;
        if constexpr (SEG) {
            sd_lo = __builtin_amdgcn_readfirstlane(a->seed_dev ? a->seed_dev[0] : a->seed_lo);
            sd_hi = __builtin_amdgcn_readfirstlane(a->seed_dev ? a->seed_dev[1] : a->seed_hi);
            s_begin = a->seg_begin;
            s_end = min(a->seg_end, total);
            ss = a->seg_state ? a->seg_state + roi : nullptr;
        }
        const int64_t off = a->pix_off[roi];
#pragma unroll
        for (int s = 0; s < PPL; s++) {
            const bool ok = s < ppl && has(s);
            if (s < ppl) L.dpx[64 * s + lane] = ok ? a->data[off + pidx(s)] : 0.0f;
            m[s] = 1.0f;
            mt[s] = 1.0f;
        }
        ci = 0;
        t_bad = false;
        if (SEG && s_begin > 0) {
            load_state(*ss);
        } else {
            status = ST_OK;
            logpost = 0.0;
            aflops = 0;
            // counters start at N = 1, A = 0 (LibRJMCMC.cpp:672-677)
            if (lane < 32) L.cnt[lane] = (lane < 30 && ((lane / 5) % 2 == 0)) ? 1 : 0;
            for (int k = lane; k < kNMax + 1; k += 64) L.hist[k] = 0;
            // initial state (cRJMCMC.cpp:120-133)
            cN = a->n0 ? min(a->n0[roi], kNMax) : 0;
            if (a->n0) {
                const int64_t eo = a->em0_off[roi];
                for (int n = lane; n < cN; n += 64) {
                    L.e[0].I[n] = a->i0[eo + n];
                    L.e[0].X[n] = a->x0[eo + n];
                    L.e[0].Y[n] = a->y0[eo + n];
                    L.e[0].S[n] = 1;
                }
            }
            e_ok = false;  // initial emitters are checked in full by the first test
            const float bg0 = a->bg0 ? a->bg0[roi] : 0.0f;
            cBG = bg0 == 0 ? 1.0f : bg0;
            cbg_pdf = findpdf(a->off, a->p.dx, cBG);
            cABG = 0;
            cBBG = 0;
            wsync();
            rebuild_model();
            tcur = 0;
            if (lane < 3) L.rec[lane] = 0;
        }
        if (a->mode == BAMF_RJ_REPLAY) {
            tbase = a->tape_off[roi];
            tlen = a->tape_len[roi];
        } else {
            tbase = 0;
            tlen = 0;
        }
        pblk = -1;
        if constexpr (LRN) {
            if (a->mode == BAMF_RJ_LEARNED && a->prop_map != nullptr) map_sums();
        }
        const int64_t rstride = total;
        for (int s = s_begin; s < s_end && status == ST_OK; s++) {
            step = s;
            burn = s < nb;
            if (burn) {
                P1 = p.p_burnin[1]; P2 = p.p_burnin[2]; P3 = p.p_burnin[3]; P4 = p.p_burnin[4];
                P56 = p.p_burnin[5] / p.p_burnin[6];
            } else {
                P1 = a->pvt[1]; P2 = a->pvt[2]; P3 = a->pvt[3]; P4 = a->pvt[4];
                P56 = a->pvt[5] / a->pvt[6];
            }
            if (s == s_begin || s == nb) {  // the phase's fixed ratios
                if (lane == 0) {
                    L.kc[1] = P1 / P2;
                    L.kc[2] = P4 / P3;
                }
                wsync();
            }
            dj = 0;
            fill_window(0);
            r_jt = 0;
            r_flags = 0;
            r_pix = -1;
            r_llr = 0;
            r_pr = 0;
            r_logpr = __builtin_nan("");
            r_nprop = -1;
            if (a->mode == BAMF_RJ_REPLAY) {
                const bamf_step_record& rr = a->ref_steps[roi * rstride + s];
                ref_flags = rr.flags;
                ref_pix = rr.pix;
            }
            const int kind = one_step();
            wsync();
            if (a->steps) {
                const float ph = total_photons_seq();
                if (lane == 0) {
                    bamf_step_record& o = a->steps[roi * rstride + s];
                    o.kind = kind;
                    o.jump_type = r_jt;
                    o.flags = r_flags;
                    o.n = cN;
                    o.llr = r_llr;
                    o.pr = r_pr;
                    o.bg = cBG;
                    o.photons = ph;
                    o.log_pr = r_logpr;
                    o.draws = dj;
                    o.n_prop = r_nprop < 0 ? cN : r_nprop;
                    o.pix = r_pix;
                    o.aux = 0;
                }
            }
            if (!burn) {
                if (lane == 0) L.hist[min(cN, kNMax)]++;
                const int t = s - nb;
                if (a->trace_n && t % a->trace_thin == 0 && t / a->trace_thin < a->trace_len) {
                    const Em& c = cur();
                    float v = (lane < cN ? c.I[lane] : 0.0f) + (lane + 64 < cN ? c.I[lane + 64] : 0.0f);
                    v = bcastf(wave_scan(v), 63);
                    if (lane == 0) {
                        const int64_t ti = (int64_t)roi * a->trace_len + t / a->trace_thin;
                        a->trace_n[ti] = (int16_t)cN;
                        a->trace_ph[ti] = v;
                    }
                }
            }
            if (a->mode == BAMF_RJ_REPLAY && tcur + dj > tlen) status = ST_TAPE;
            tcur += dj;
        }
        wsync();
        if constexpr (SEG) {
            if (s_end < total) {  // a later segment continues this chain
                save_state(*ss);
                return;
            }
        }
        // ---- outputs ----
        if (a->status && lane == 0) a->status[roi] = status;
        if (a->counters && lane < 30) a->counters[(int64_t)roi * 30 + lane] = L.cnt[lane];
        if (a->logpost && lane == 0) a->logpost[roi] = logpost;
        if (a->alg_flops && lane == 0) a->alg_flops[roi] = aflops;
        if (a->acc_rec && lane == 0) {
            a->rec_count[3 * (int64_t)roi + 0] = L.rec[0];
            a->rec_count[3 * (int64_t)roi + 1] = L.rec[1];
            a->rec_count[3 * (int64_t)roi + 2] = L.rec[2];
        }
        if (a->hist) {
            const int hl = a->hist_len;
            for (int k = lane; k < hl; k += 64) {
                int v = 0;
                if (k < hl - 1) {
                    v = k <= kNMax ? L.hist[k] : 0;
                } else {
                    for (int q = k; q <= kNMax; q++) v += L.hist[q];
                }
                a->hist[(int64_t)roi * hl + k] = v;
            }
        }
        if (a->fin) {
            float* f = a->fin + (int64_t)roi * (4 + 4 * a->fin_cap);
            if (lane == 0) {
                f[0] = (float)cN;
                f[1] = cBG;
                f[2] = cABG;
                f[3] = cBBG;
            }
            const Em& c = cur();
            for (int n = lane; n < cN && n < a->fin_cap; n += 64) {
                f[4 + n] = c.I[n];
                f[4 + a->fin_cap + n] = c.X[n];
                f[4 + 2 * a->fin_cap + n] = c.Y[n];
                f[4 + 3 * a->fin_cap + n] = (float)c.S[n];
            }
        }
        wsync();
    }
};

// The argument block is the kernel's only parameter, so it sits at offset 0 of
// the kernarg segment; the chain reads it through that segment pointer.
// One instantiation per ROI size class (PPL pixel slots per lane), each
// draining its own class list through an atomic queue.
template <int PPL, int MINB, int F>
__global__ __launch_bounds__(64 * kWavesPerBlock, MINB) void k_chain(ChainArgs) {
    __shared__ WaveLds<PPL> lds[kWavesPerBlock];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int lane = threadIdx.x & 63;
    const ChainArgs* ap = (const ChainArgs*)(__builtin_amdgcn_kernarg_segment_ptr());
    Chain<PPL, F> ch(ap, lds[w], lane);
    const int n = *ap->list_count;
    while (true) {
        int qi = 0;
        if (lane == 0) qi = (int)atomicAdd(ap->queue, 1u);
        qi = bcasti(qi, 0);
        if (qi >= n) break;
        ch.run(ap->list[qi]);
    }
}

// Both size classes in one launch: the wavefronts drain the large-ROI list
// and then the small-ROI list through one work-queue counter, each ROI run by
// the chain of its class in the same LDS block (the small class's layout is
// smaller and starts at the same address; every chain's run begins and ends
// with a wavefront fence).  A wavefront takes the next ROI as soon as its
// chain ends, whatever the class, so no work-group holds its CU share while
// waiting for its last chain, and small ROIs fill the slots the large ones
// free.  The register budget is the large class's.
template <int MINB, int F>
__global__ __launch_bounds__(64 * kWavesPerBlock, MINB) void k_chain_all(ChainArgs) {
    static_assert(sizeof(WaveLds<kClassPPL[0]>) <= sizeof(WaveLds<kClassPPL[1]>), "the small layout fits");
    __shared__ WaveLds<kClassPPL[1]> lds[kWavesPerBlock];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const ChainArgs* ap = (const ChainArgs*)(__builtin_amdgcn_kernarg_segment_ptr());
    Chain<kClassPPL[1], F> large(ap, lds[w], lane);
    Chain<kClassPPL[0], F> small(ap, *reinterpret_cast<WaveLds<kClassPPL[0]>*>(&lds[w]), lane);
    const int n1 = *ap->list_count, n = n1 + *ap->list2_count;
    while (true) {
        int qi = 0;
        if (lane == 0) qi = (int)atomicAdd(ap->queue, 1u);
        qi = bcasti(qi, 0);
        if (qi >= n) break;
        if (qi < n1) large.run(ap->list[qi]);
        else small.run(ap->list2[qi - n1]);
    }
}

// Stable split of the ROI order into the size classes: ROIs of at most
// small_max pixel slots (<= kClassPPL[0]) to the small class, the rest to the
// large one; also resets the nq work-queue counters (two per segment).  One
// work-group walks the order in tiles of 1024 with a work-group scan.
__global__ __launch_bounds__(1024) void k_classify(int n_roi, const int32_t* order, const int32_t* xs,
                                                   const int32_t* ys, int32_t* list0, int32_t* list1,
                                                   int32_t* counts, uint32_t* queues, int nq, int small_max) {
    __shared__ int sc[2][1024];
    __shared__ int base[2];
    const int t = threadIdx.x;
    if (t < 2) base[t] = 0;
    __syncthreads();
    for (int i0 = 0; i0 < n_roi; i0 += 1024) {
        const int i = i0 + t;
        int roi = -1, c = -1;
        if (i < n_roi) {
            roi = order ? order[i] : i;
            c = slots_for(xs[roi], ys[roi]) <= small_max ? 0 : 1;
        }
        sc[0][t] = c == 0;
        sc[1][t] = c == 1;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int v0 = t >= o ? sc[0][t - o] : 0, v1 = t >= o ? sc[1][t - o] : 0;
            __syncthreads();
            sc[0][t] += v0;
            sc[1][t] += v1;
            __syncthreads();
        }
        if (c == 0) list0[base[0] + sc[0][t] - 1] = roi;
        if (c == 1) list1[base[1] + sc[1][t] - 1] = roi;
        __syncthreads();
        if (t < 2) base[t] += sc[t][1023];
        __syncthreads();
    }
    if (t < 2) counts[t] = base[t];
    for (int q = t; q < nq; q += 1024) queues[q] = 0u;
}

// The seed of a graph replay, written ahead of the graph on the launch stream.
__global__ void k_set_seed(uint32_t* d, uint32_t lo, uint32_t hi) {
    d[0] = lo;
    d[1] = hi;
}


__global__ void k_erf(const float* x, float* y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = erf_nr(x[i]);
}

}  // namespace

int launch_classify(int n_roi, const int32_t* order, const int32_t* xs, const int32_t* ys, int32_t* list0,
                    int32_t* list1, int32_t* counts, uint32_t* queues, int nq, int small_max, hipStream_t s) {
    BAMF_CHECK_ARG(small_max >= 0 && small_max <= kClassPPL[0], "small-class slot bound %d out of range", small_max);
    hipLaunchKernelGGL(k_classify, dim3(1), dim3(1024), 0, s, n_roi, order, xs, ys, list0, list1, counts, queues,
                       nq, small_max);
    BAMF_LAUNCH_CHECK("k_classify");
    return BAMF_OK;
}

int launch_set_seed(uint32_t* d, uint32_t lo, uint32_t hi, hipStream_t s) {
    hipLaunchKernelGGL(k_set_seed, dim3(1), dim3(1), 0, s, d, lo, hi);
    BAMF_LAUNCH_CHECK("k_set_seed");
    return BAMF_OK;
}

using ChainKernelFn = void (*)(ChainArgs);

// The instantiations per class: feature sets 0 (plain), kGJ, kSEG,
// kGJ | kSEG, kLRN and all three (learned mode with group jumps or
// segments), and one sampled-PSF kernel (kNUM with kGJ | kSEG, whose code
// for those features is inert in runs without them); the plain kernel also
// in each register budget of kTuneMinBlocks.
template <int C>
ChainKernelFn chain_kernel(int f, int tune) {
    constexpr int P = kClassPPL[C], M = kClassMinBlocks[C];
    static_assert(kTuneMinBlocks[kTuneDefault] == M, "the default budget is the class's");
    if (f & kNUM) return k_chain<P, M, kGJ | kSEG | kNUM>;
    if (f & kLRN) return f == kLRN ? k_chain<P, M, kLRN> : k_chain<P, M, kGJ | kSEG | kLRN>;
    if (f == kGJ) return k_chain<P, M, kGJ>;
    if (f == kSEG) return k_chain<P, M, kSEG>;
    if (f == (kGJ | kSEG)) return k_chain<P, M, kGJ | kSEG>;
    return tune == 0 ? k_chain<P, kTuneMinBlocks[0], 0> : tune == 2 ? k_chain<P, kTuneMinBlocks[2], 0> : k_chain<P, M, 0>;
}

// One launch over both classes (class index kClasses): plain, segmented and
// learned chains, at the large class's default register budget.
ChainKernelFn chain_kernel_all(int f) {
    constexpr int M = kClassMinBlocks[1];
    return f == kSEG ? k_chain_all<M, kSEG> : f == kLRN ? k_chain_all<M, kLRN> : k_chain_all<M, 0>;
}
ChainKernelFn select_chain(int cls, int features, int tune) {
    return cls == kClasses ? chain_kernel_all(features)
         : cls == 0        ? chain_kernel<0>(features, tune)
                           : chain_kernel<1>(features, tune);
}

int launch_chain(const ChainArgs& a, int cls, int features, int tune, int grid, hipStream_t s) {
    BAMF_CHECK_ARG(cls != kClasses || one_launch_features(features), "no one-launch kernel for features %d", features);
    const ChainKernelFn k = select_chain(cls, features, tune);
    hipLaunchKernelGGL(k, dim3(grid), dim3(64 * kWavesPerBlock), 0, s, a);
    BAMF_LAUNCH_CHECK("k_chain");
    return BAMF_OK;
}

int launch_erf(const float* x, float* y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_erf, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, x, y, n);
    BAMF_LAUNCH_CHECK("k_erf");
    return BAMF_OK;
}

int chain_occupancy_blocks(int cls, int features, int tune, int* blocks_per_cu) {
    int n = 0;
    BAMF_CHECK_ARG(cls != kClasses || one_launch_features(features), "no one-launch kernel for features %d", features);
    const ChainKernelFn k = select_chain(cls, features, tune);
    BAMF_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(k),
                                                               64 * kWavesPerBlock, 0));
    *blocks_per_cu = n;
    return BAMF_OK;
}

}  // namespace rj
}  // namespace bamf
