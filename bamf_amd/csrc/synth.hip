// Synthetic data from BAMF's forward model, generated on the device: the
// frames and ROIs the benchmarks and the proposal network's training use.
//
// Model (the convention of LibRJMCMC.cpp:15-59): pixel (y, x) covers
// [x, x+1) x [y, y+1); its expected count is
//   mu = bg + sum_e I_e * 0.5 (erf((x+1-X_e) c) - erf((x-X_e) c))
//                      * 0.5 (erf((y+1-Y_e) c) - erf((y-Y_e) c)),  c = 1/(sigma sqrt 2),
// and the observed count is Poisson(mu).  Emitters form a Poisson point
// process of the given density with Gamma(k, theta) intensities.
//
// Frames: the process is drawn per 16 x 16 pixel cell (disjoint cells of a
// Poisson process are independent, each with a Poisson count and uniform
// positions), so every output tile gathers the emitters of its 3 x 3
// neighbouring cells and renders without atomics: the frames are a
// deterministic function of the seed.  A PSF tail beyond 16 pixels (12
// sigma at sigma 1.3) is dropped.
//
// Random variates (all from Philox4x32-10, philox.hpp):
//   uniform   (r >> 8 + 1/2) / 2^24, in the open interval (0, 1);
//   normal    Box-Muller;
//   Gamma     Marsaglia & Tsang, "A simple method for generating gamma
//             variables", ACM TOMS 26 (2000) 363-372 (shape >= 1; shape < 1
//             by the boost U^(1/k));
//   Poisson   inversion by sequential search below mu = 10; above,
//             W. Hormann, "The transformed rejection method for generating
//             Poisson random variables", Insurance: Mathematics and
//             Economics 12 (1993) 39-45 (PTRS).
#include <cmath>

#include "common.hpp"
#include "philox.hpp"

namespace bamf {
namespace {

constexpr int kCell = 16;
constexpr int kCellCap = 128;  // emitters kept per cell (beyond: P < 1e-40 at 0.1 per px^2)

// A stream of uniforms: Philox with counter (n, a, b, tag), key = seed.
struct Stream {
    uint32_t a, b, tag, k0, k1, n;
    uint32_t buf[4];
    int used;
    __device__ Stream(uint64_t seed, uint32_t a_, uint32_t b_, uint32_t tag_)
        : a(a_), b(b_), tag(tag_), k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), n(0), used(4) {}
    __device__ float u() {
        if (used == 4) {
            philox(n++, a, b, tag, k0, k1, buf);
            used = 0;
        }
        return ((float)(buf[used++] >> 8) + 0.5f) * (1.0f / 16777216.0f);
    }
    __device__ double normal() {
        const double u1 = u(), u2 = u();
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

__device__ double gamma_ge1(Stream& g, double k) {
    const double d = k - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (int it = 0; it < 1000; it++) {
        const double x = g.normal();
        double v = 1.0 + c * x;
        if (v <= 0.0) continue;
        v = v * v * v;
        const double uu = g.u();
        if (uu < 1.0 - 0.0331 * x * x * x * x) return d * v;
        if (log(uu) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v;
    }
    return d;  // not reached in practice (acceptance rate > 0.95)
}

__device__ double gamma_variate(Stream& g, double k, double theta) {
    if (k < 1.0) return theta * gamma_ge1(g, k + 1.0) * pow((double)g.u(), 1.0 / k);  // Gamma(k+1) U^(1/k)
    return theta * gamma_ge1(g, k);
}

__device__ int poisson_variate(Stream& g, double mu) {
    if (!(mu > 0.0)) return 0;
    if (mu < 10.0) {
        double p = exp(-mu), F = p;
        const double uu = g.u();
        int k = 0;
        while (uu > F && k < 200) {
            k++;
            p *= mu / k;
            F += p;
        }
        return k;
    }
    const double slam = sqrt(mu), loglam = log(mu);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 10000; it++) {
        const double U = (double)g.u() - 0.5, V = g.u();
        const double us = 0.5 - fabs(U);
        const double kf = floor((2.0 * a / us + b) * U + mu + 0.43);
        if (us >= 0.07 && V <= vr) return (int)kf;
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -mu + kf * loglam - lgamma(kf + 1.0))
            return (int)kf;
    }
    return (int)floor(mu);  // not reached in practice
}

// Cell (cx, cy) of frame f: a Poisson count for its area inside the frame,
// positions uniform there, Gamma intensities.  truth[(f * ncell + c) * cap + k]
// = {x, y, I}; count[f * ncell + c] = emitters kept (at most cap).
__global__ void k_cells(int nframes, int ysize, int xsize, float density, float gk, float gtheta, uint64_t seed,
                        int cap, float* __restrict__ truth, int32_t* __restrict__ count) {
    const int ncx = (xsize + kCell - 1) / kCell, ncy = (ysize + kCell - 1) / kCell, ncell = ncx * ncy;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)ncell * nframes) return;
    const int f = (int)(t / ncell), c = (int)(t % ncell);
    const int cx = c % ncx, cy = c / ncx;
    const float x0 = (float)(cx * kCell), y0 = (float)(cy * kCell);
    const float w = fminf((float)kCell, (float)xsize - x0), h = fminf((float)kCell, (float)ysize - y0);
    Stream g(seed, (uint32_t)f, (uint32_t)c, 0x5eed0001u);
    int n = poisson_variate(g, (double)density * (double)w * (double)h);
    n = n < cap ? n : cap;
    float* out = truth + ((int64_t)f * ncell + c) * cap * 3;
    for (int k = 0; k < n; k++) {
        const float x = x0 + w * g.u(), y = y0 + h * g.u();
        const float I = (float)gamma_variate(g, gk, gtheta);
        out[3 * k + 0] = x;
        out[3 * k + 1] = y;
        out[3 * k + 2] = I;
    }
    count[(int64_t)f * ncell + c] = n;
}

// One 16 x 16 output tile (= cell) of one frame per work-group: the 3 x 3
// neighbouring cells' emitters in chunks of 64, their pixel-integrated PSF
// factors over the tile's columns and rows, then each pixel's sum and its
// Poisson draw (noise = 0: the expected count itself).
__global__ __launch_bounds__(256) void k_render(int ysize, int xsize, float bg, float sigma, uint64_t seed, int cap,
                                                const float* __restrict__ truth, const int32_t* __restrict__ count,
                                                int noise, float* __restrict__ frames) {
    __shared__ float gx[64][kCell], gy[64][kCell], gi[64];
    __shared__ int base[9], ncnt[9];
    const int ncx = (xsize + kCell - 1) / kCell, ncy = (ysize + kCell - 1) / kCell, ncell = ncx * ncy;
    const int f = blockIdx.y, c = blockIdx.x;
    const int cx = c % ncx, cy = c / ncx;
    const int tid = threadIdx.x, tx = tid % kCell, ty = tid / kCell;
    const int x = cx * kCell + tx, y = cy * kCell + ty;
    const float cc = 1.0f / (sigma * sqrtf(2.0f));
    if (tid < 9) {
        const int ncxi = cx + tid % 3 - 1, ncyi = cy + tid / 3 - 1;
        const bool ok = ncxi >= 0 && ncxi < ncx && ncyi >= 0 && ncyi < ncy;
        const int nc = ncyi * ncx + ncxi;
        base[tid] = ok ? nc : 0;
        ncnt[tid] = ok ? count[(int64_t)f * ncell + nc] : 0;
    }
    __syncthreads();
    int total = 0;
    for (int q = 0; q < 9; q++) total += ncnt[q];
    float mu = bg;
    for (int e0 = 0; e0 < total; e0 += 64) {
        // the chunk's emitters: column and row factors over this tile
        for (int j = tid; j < 64 * kCell; j += 256) {
            const int e = j / kCell, i = j % kCell, eg = e0 + e;
            float fx = 0.0f, fy = 0.0f, I = 0.0f;
            if (eg < total) {
                int q = 0, r = eg;
                while (r >= ncnt[q]) r -= ncnt[q++];
                const float* em = truth + (((int64_t)f * ncell + base[q]) * cap + r) * 3;
                const float X = em[0], Y = em[1];
                I = em[2];
                const float xi = (float)(cx * kCell + i), yi = (float)(cy * kCell + i);
                fx = 0.5f * (erff((xi + 1.0f - X) * cc) - erff((xi - X) * cc));
                fy = 0.5f * (erff((yi + 1.0f - Y) * cc) - erff((yi - Y) * cc));
            }
            gx[e][i] = fx;
            gy[e][i] = fy;
            if (i == 0) gi[e] = I;
        }
        __syncthreads();
        const int m = total - e0 < 64 ? total - e0 : 64;
        for (int e = 0; e < m; e++) mu = mu + gi[e] * gx[e][tx] * gy[e][ty];
        __syncthreads();
    }
    if (x < xsize && y < ysize) {
        float v = mu;
        if (noise) {
            Stream g(seed, (uint32_t)f, (uint32_t)(y * xsize + x), 0x5eed0002u);
            v = (float)poisson_variate(g, (double)mu);
        }
        frames[((int64_t)f * ysize + y) * xsize + x] = v;
    }
}

// Training ROIs: one work-group per ROI of side s (truth, when given: per
// ROI 1 + 3 * kCellCap floats, the background then {x, y, I} per emitter).  Density ~ U[d_lo, d_hi]
// per pixel^2, N ~ Poisson(density s^2) (at most kCellCap), positions uniform
// in the ROI, Gamma intensities, background ~ U[bg_lo, bg_hi].  Outputs in
// the ROI layout [roi][x][y]: the Poisson image, the occupancy target (1 at
// pixel (floor(x + 1/2), floor(y + 1/2)), clamped to the ROI, of each
// emitter: the pixel a birth proposes it from) and the log of the intensity
// summed per occupied pixel (0 elsewhere).
__global__ __launch_bounds__(256) void k_roi_batch(int n, int side, float d_lo, float d_hi, float bg_lo,
                                                   float bg_hi, float gk, float gtheta, float sigma, uint64_t seed,
                                                   int noise, float* __restrict__ img, float* __restrict__ occ,
                                                   float* __restrict__ logi, int32_t* __restrict__ nem,
                                                   float* __restrict__ truth) {
    __shared__ float ex[kCellCap], ey[kCellCap], ei[kCellCap];
    __shared__ float isum[32 * 32];
    __shared__ int s_n;
    __shared__ float s_bg;
    const int r = blockIdx.x;
    const int P = side * side;
    if (threadIdx.x == 0) {
        Stream g(seed, (uint32_t)r, 0u, 0x5eed0003u);
        const float dens = d_lo + (d_hi - d_lo) * g.u();
        int k = poisson_variate(g, (double)dens * (double)P);
        k = k < kCellCap ? k : kCellCap;
        for (int e = 0; e < k; e++) {
            ex[e] = side * g.u();
            ey[e] = side * g.u();
            ei[e] = (float)gamma_variate(g, gk, gtheta);
        }
        s_n = k;
        s_bg = bg_lo + (bg_hi - bg_lo) * g.u();
    }
    for (int p = threadIdx.x; p < P; p += blockDim.x) isum[p] = 0.0f;
    __syncthreads();
    if (threadIdx.x == 0) {  // occupancy in emitter order (deterministic sums)
        for (int e = 0; e < s_n; e++) {
            int px = (int)floorf(ex[e] + 0.5f), py = (int)floorf(ey[e] + 0.5f);
            px = px < side ? px : side - 1;
            py = py < side ? py : side - 1;
            isum[px * side + py] += ei[e];
        }
        if (nem) nem[r] = s_n;
        if (truth) {
            float* t = truth + (int64_t)r * (3 * kCellCap + 1);
            t[0] = s_bg;
            for (int e = 0; e < s_n; e++) {
                t[1 + 3 * e] = ex[e];
                t[2 + 3 * e] = ey[e];
                t[3 + 3 * e] = ei[e];
            }
        }
    }
    __syncthreads();
    const float cc = 1.0f / (sigma * sqrtf(2.0f));
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        const int x = p / side, y = p % side;
        float mu = s_bg;
        for (int e = 0; e < s_n; e++) {
            const float fx = 0.5f * (erff(((float)x + 1.0f - ex[e]) * cc) - erff(((float)x - ex[e]) * cc));
            const float fy = 0.5f * (erff(((float)y + 1.0f - ey[e]) * cc) - erff(((float)y - ey[e]) * cc));
            mu = mu + ei[e] * fx * fy;
        }
        Stream g(seed, (uint32_t)r, (uint32_t)(p + 1), 0x5eed0004u);
        const int64_t o = (int64_t)r * P + p;
        img[o] = noise ? (float)poisson_variate(g, (double)mu) : mu;
        if (occ) occ[o] = isum[p] > 0.0f ? 1.0f : 0.0f;
        if (logi) logi[o] = isum[p] > 0.0f ? logf(isum[p]) : 0.0f;
    }
}

}  // namespace
}  // namespace bamf

using namespace bamf;

extern "C" {

int64_t bamf_synth_cells(int ysize, int xsize) {
    if (ysize < 1 || xsize < 1) return 0;
    return (int64_t)((xsize + kCell - 1) / kCell) * ((ysize + kCell - 1) / kCell);
}

int bamf_synth_frames(int nframes, int ysize, int xsize, float density, float gamma_k, float gamma_theta, float bg,
                      float sigma, uint64_t seed, int cap, int noise, float* d_frames, float* d_truth,
                      int32_t* d_count, void* stream) {
    BAMF_CHECK_ARG(nframes >= 0 && nframes <= 65535 && ysize >= 1 && xsize >= 1, "bad sizes");
    BAMF_CHECK_ARG(cap >= 1 && cap <= kCellCap, "cap must be in 1..%d", kCellCap);
    BAMF_CHECK_ARG(density >= 0 && gamma_k > 0 && gamma_theta > 0 && sigma > 0, "bad model parameters");
    BAMF_CHECK_ARG(d_frames && d_truth && d_count, "null device pointer");
    if (nframes == 0) return BAMF_OK;
    hipStream_t s = as_stream(stream);
    const int64_t ncell = bamf_synth_cells(ysize, xsize);
    const int64_t nt = ncell * nframes;
    hipLaunchKernelGGL(k_cells, dim3((unsigned)ceil_div(nt, 256)), dim3(256), 0, s, nframes, ysize, xsize, density,
                       gamma_k, gamma_theta, seed, cap, d_truth, d_count);
    BAMF_LAUNCH_CHECK("k_cells");
    hipLaunchKernelGGL(k_render, dim3((unsigned)ncell, (unsigned)nframes), dim3(256), 0, s, ysize, xsize, bg, sigma,
                       seed, cap, d_truth, d_count, noise, d_frames);
    BAMF_LAUNCH_CHECK("k_render");
    return BAMF_OK;
}

int bamf_synth_roi_batch(int n, int side, float density_lo, float density_hi, float bg_lo, float bg_hi,
                         float gamma_k, float gamma_theta, float sigma, uint64_t seed, int noise, float* d_img,
                         float* d_occ, float* d_logi, int32_t* d_nem, float* d_truth, void* stream) {
    BAMF_CHECK_ARG(n >= 0 && side >= 1 && side <= 32, "side must be in 1..32");
    BAMF_CHECK_ARG(density_lo >= 0 && density_hi >= density_lo && bg_hi >= bg_lo && gamma_k > 0 && gamma_theta > 0 &&
                       sigma > 0,
                   "bad model parameters");
    BAMF_CHECK_ARG(d_img, "null device pointer");
    if (n == 0) return BAMF_OK;
    hipLaunchKernelGGL(k_roi_batch, dim3((unsigned)n), dim3(256), 0, as_stream(stream), n, side, density_lo,
                       density_hi, bg_lo, bg_hi, gamma_k, gamma_theta, sigma, seed, noise, d_img, d_occ, d_logi,
                       d_nem, d_truth);
    BAMF_LAUNCH_CHECK("k_roi_batch");
    return BAMF_OK;
}

}  // extern "C"
