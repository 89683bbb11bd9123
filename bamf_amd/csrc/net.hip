// Birth-proposal network on gfx950: four 3x3 convolutions (1->32, then
// 32->32 three times) with ReLU and a 1x1 head (32->2: occupancy logit,
// log-intensity), trained on the device.
//
// Layout: activations NHWC bfloat16 with 32 channels (64 B per pixel); the
// weights of the 3x3 layers as [cout][tap * cin + c] bfloat16 for the
// matrix cores, with fp32 master copies for training.  All products run on
// v_mfma_f32_16x16x32_bf16 with fp32 accumulation:
//   forward      D[pixel][cout]  = sum_k im2col(A)[pixel][k] W[cout][k]   (k = tap * 32 + c)
//   input grad   dA[pixel][c]    = the forward product on dZ with the
//                                  taps mirrored and W transposed
//   weight grad  dW[cout][k]     = sum_pixel dZ[pixel][cout] im2col(A)[pixel][k]
// Operand layout of the 16x16x32 bf16 product: lane l supplies A row l % 16
// and B column l % 16, each with the eight K entries 8 (l / 16) .. + 7, and
// receives D rows 4 (l / 16) .. + 3 of column l % 16.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace bamf {
namespace net {
namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kC = 32;        // hidden channels
constexpr int kK = 9 * kC;    // im2col width of a 32-channel layer
constexpr int kK1 = 32;       // first layer: 9 taps padded to one 32-wide K block

__host__ __device__ inline uint16_t f2bf(float f) {
    // round to nearest even (finite inputs)
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

__device__ inline v8bf as_v8bf(uint4 u) { return __builtin_bit_cast(v8bf, u); }

struct Geo {
    int H, W;             // canvas size
    const int32_t* vh;    // per-image valid rows (NULL: H)
    const int32_t* vw;    // per-image valid columns (NULL: W)
};
__device__ inline void valid_hw(const Geo& g, int b, int& hv, int& wv) {
    hv = g.vh ? g.vh[b] : g.H;
    wv = g.vw ? g.vw[b] : g.W;
}

// ---------------------------------------------------------------------------
// 3x3 convolution, 64 pixels x 32 output channels per wavefront.
//   CIN = 1:  input x0 [pixel] bf16, K = 9 taps padded to 32
//   CIN = 32: input [pixel][32] bf16, K = 9 x 32
// MODE 0: out = relu(D + bias); MODE 1: out = D * (mask_act > 0) (input
// gradient, weights already mirrored/transposed, no bias).  Pixels outside
// an image's valid region are written as 0.
template <int CIN, int MODE>
__global__ __launch_bounds__(256) void k_conv3(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wt,
                                               const float* __restrict__ bias, const uint16_t* __restrict__ mask_act,
                                               uint16_t* __restrict__ out, int64_t npix, Geo g) {
    const int lane = threadIdx.x & 63;
    const int64_t tile = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    if (tile >= npix) return;
    const int row = lane & 15, grp = lane >> 4;
    const int hw = g.H * g.W;
    // the four pixels this lane feeds as A rows (one per M tile)
    int pb[4], ph[4], pw[4], hv[4], wv[4];
    bool pin[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const int64_t p = tile + 16 * mt + row;
        pin[mt] = p < npix;
        const int64_t pp = pin[mt] ? p : 0;
        pb[mt] = (int)(pp / hw);
        const int r = (int)(pp - (int64_t)pb[mt] * hw);
        ph[mt] = r / g.W;
        pw[mt] = r - ph[mt] * g.W;
        valid_hw(g, pb[mt], hv[mt], wv[mt]);
    }
    v4f acc[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++) acc[mt][nt] = v4f{0.f, 0.f, 0.f, 0.f};

    if (CIN == 1) {
        // K entry k = 8 grp + j is tap k (k < 9)
        uint4 bw[2];
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
            bw[nt] = *reinterpret_cast<const uint4*>(wt + (size_t)(16 * nt + row) * kK1 + 8 * grp);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            uint16_t v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = 8 * grp + j;
                uint16_t x = 0;
                if (k < 9 && pin[mt]) {
                    const int hh = ph[mt] + k / 3 - 1, ww = pw[mt] + k % 3 - 1;
                    if (hh >= 0 && hh < hv[mt] && ww >= 0 && ww < wv[mt])
                        x = in[(int64_t)pb[mt] * hw + hh * g.W + ww];
                }
                v[j] = x;
            }
            uint4 a;
            a.x = v[0] | ((uint32_t)v[1] << 16);
            a.y = v[2] | ((uint32_t)v[3] << 16);
            a.z = v[4] | ((uint32_t)v[5] << 16);
            a.w = v[6] | ((uint32_t)v[7] << 16);
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_v8bf(a), as_v8bf(bw[nt]), acc[mt][nt], 0, 0, 0);
        }
    } else {
#pragma unroll 1
        for (int tap = 0; tap < 9; tap++) {
            const int dh = tap / 3 - 1, dw = tap % 3 - 1;
            uint4 bw[2];
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                bw[nt] = *reinterpret_cast<const uint4*>(wt + (size_t)(16 * nt + row) * kK + tap * kC + 8 * grp);
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                const int hh = ph[mt] + dh, ww = pw[mt] + dw;
                uint4 a = make_uint4(0, 0, 0, 0);
                if (pin[mt] && hh >= 0 && hh < hv[mt] && ww >= 0 && ww < wv[mt])
                    a = *reinterpret_cast<const uint4*>(in + (((int64_t)pb[mt] * hw + hh * g.W + ww) * kC + 8 * grp));
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
                    acc[mt][nt] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_v8bf(a), as_v8bf(bw[nt]), acc[mt][nt], 0, 0, 0);
            }
        }
    }
    // D: rows 4 grp + i of M tile mt, column row of N tile nt
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t p = tile + 16 * mt + 4 * grp + i;
            if (p >= npix) continue;
            const int b = (int)(p / hw);
            const int r = (int)(p - (int64_t)b * hw);
            const int h = r / g.W, w = r - (r / g.W) * g.W;
            int hvv, wvv;
            valid_hw(g, b, hvv, wvv);
            const bool ok = h < hvv && w < wvv;
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const int co = 16 * nt + row;
                float v = acc[mt][nt][i];
                if (MODE == 0) {
                    v = v + bias[co];
                    v = v > 0.f ? v : 0.f;
                } else {
                    v = bf2f(mask_act[p * kC + co]) > 0.f ? v : 0.f;
                }
                out[p * kC + co] = ok ? f2bf(v) : (uint16_t)0;
            }
        }
}

// Weight and bias gradients of a 3 x 3 convolution, dW_tap = dZ^T A_tap with
// the pixels as K: one work-group of four wavefronts per image (grid-strided
// over the batch, each wavefront accumulating its output tiles in registers
// across its images), the image's dZ transposed to [channel][pixel] and its
// input zero-padded by one pixel in LDS, so every MFMA operand is an LDS read;
// one atomic add per output element per work-group at the end.
//   CIN = 32: 36 tiles (9 taps x 2 x 2 of 16 x 16), nine per wavefront;
//   CIN = 1:  2 tiles (output channels x the 9 taps padded to 16).
template <int CIN>
__global__ __launch_bounds__(256) void k_conv3_wgrad(const uint16_t* __restrict__ dz, const uint16_t* __restrict__ in,
                                                    float* __restrict__ dw, float* __restrict__ db, int batch, Geo g) {
    constexpr int TPW = CIN == 1 ? 1 : 9;  // tiles per wavefront
    extern __shared__ uint16_t smem[];
    const int P = g.H * g.W, Pp = (P + 31) & ~31, W2 = g.W + 2, HW2 = (g.H + 2) * W2;
    uint16_t* dzT = smem;                                         // [32][Pp]
    uint16_t* halo = dzT + 32 * Pp;                               // [CIN][HW2]
    int32_t* hidx = reinterpret_cast<int32_t*>(halo + ((CIN * HW2 + 1) & ~1));  // [Pp]
    __shared__ float bsum[32];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const bool active = CIN == 32 || wv < 2;
    v4f acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; i++) acc[i] = v4f{0.f, 0.f, 0.f, 0.f};
    if (tid < 32) bsum[tid] = 0.f;
    __syncthreads();
    float bpart[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // this thread's eight dZ channels
    const int c8 = tid & 3;
    for (int b = blockIdx.x; b < batch; b += gridDim.x) {
        int hv, wvv;
        valid_hw(g, b, hv, wvv);
        const int64_t pb = (int64_t)b * P;
        // stage dZ transposed (zeros past P), the padded input and the pixel -> halo index table
        for (int idx = tid; idx < Pp * 4; idx += 256) {
            const int p = idx >> 2;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (p < P) v = *reinterpret_cast<const uint4*>(dz + (pb + p) * kC + 8 * c8);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint16_t h16 = (uint16_t)(w4[j >> 1] >> (16 * (j & 1)));
                dzT[(8 * c8 + j) * Pp + p] = h16;
                bpart[j] += bf2f(h16);
            }
        }
        for (int idx = tid; idx < CIN * HW2; idx += 256) halo[idx] = 0;
        __syncthreads();
        for (int idx = tid; idx < Pp; idx += 256) {
            const int h = idx / g.W, w = idx - h * g.W;
            // padding pixels (dZ = 0): an interior cell, so every tap offset stays inside the halo
            hidx[idx] = idx < P ? (h + 1) * W2 + (w + 1) : W2 + 1;
        }
        if constexpr (CIN == 1) {
            for (int p = tid; p < P; p += 256) {
                const int h = p / g.W, w = p - h * g.W;
                if (h < hv && w < wvv) halo[(h + 1) * W2 + (w + 1)] = in[pb + p];
            }
        } else {
            for (int idx = tid; idx < P * 4; idx += 256) {
                const int p = idx >> 2, h = p / g.W, w = p - h * g.W;
                if (h < hv && w < wvv) {
                    const uint4 v = *reinterpret_cast<const uint4*>(in + (pb + p) * kC + 8 * c8);
                    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        halo[(8 * c8 + j) * HW2 + (h + 1) * W2 + (w + 1)] = (uint16_t)(w4[j >> 1] >> (16 * (j & 1)));
                }
            }
        }
        __syncthreads();
        if (active) {
            for (int k0 = 0; k0 < Pp; k0 += 32) {
                const int p8 = k0 + 8 * grp;
                int hx[8];
#pragma unroll
                for (int j = 0; j < 8; j++) hx[j] = hidx[p8 + j];
#pragma unroll
                for (int i = 0; i < TPW; i++) {
                    const int t = wv * TPW + i;
                    // CIN = 32: tile t = (tap, mt, nt); CIN = 1: tile = mt, the N columns are the taps
                    const int tap = CIN == 1 ? row : t >> 2;
                    const int mt = CIN == 1 ? t : (t >> 1) & 1, nt = CIN == 1 ? 0 : t & 1;
                    const uint4 av = *reinterpret_cast<const uint4*>(dzT + (16 * mt + row) * Pp + p8);
                    uint16_t bv[8];
                    const bool tv = CIN == 32 || tap < 9;
                    const int off = tv ? (tap / 3 - 1) * W2 + (tap % 3 - 1) : 0;
                    const uint16_t* src = halo + (CIN == 1 ? 0 : (16 * nt + row) * HW2) + off;
#pragma unroll
                    for (int j = 0; j < 8; j++) bv[j] = tv ? src[hx[j]] : (uint16_t)0;
                    uint4 bb;
                    bb.x = bv[0] | ((uint32_t)bv[1] << 16);
                    bb.y = bv[2] | ((uint32_t)bv[3] << 16);
                    bb.z = bv[4] | ((uint32_t)bv[5] << 16);
                    bb.w = bv[6] | ((uint32_t)bv[7] << 16);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_v8bf(av), as_v8bf(bb), acc[i], 0, 0, 0);
                }
            }
        }
        __syncthreads();  // the buffers are restaged for the next image
    }
    // D rows 4 grp + r (output channels), column row (input channel, or tap for CIN = 1)
    if (active) {
#pragma unroll
        for (int i = 0; i < TPW; i++) {
            const int t = wv * TPW + i;
            const int tap = CIN == 1 ? row : t >> 2;
            const int mt = CIN == 1 ? t : (t >> 1) & 1, nt = CIN == 1 ? 0 : t & 1;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int o = 16 * mt + 4 * grp + r;
                if (CIN == 1) {
                    if (tap < 9) atomicAdd(&dw[(size_t)o * 9 + tap], acc[i][r]);
                } else {
                    atomicAdd(&dw[(size_t)o * kK + tap * kC + 16 * nt + row], acc[i][r]);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) atomicAdd(&bsum[8 * c8 + j], bpart[j]);
    __syncthreads();
    if (tid < 32) atomicAdd(&db[tid], bsum[tid]);
}

// Input normalisation: one work-group per image: x = (sqrt(D + 3/8) - mean) / 2
// over the valid region (bfloat16), zero elsewhere.
__global__ __launch_bounds__(256) void k_prep(const float* __restrict__ img, const int64_t* __restrict__ off,
                                              uint16_t* __restrict__ x0, Geo g) {
    __shared__ float red[256];
    const int b = blockIdx.x;
    int hv, wv;
    valid_hw(g, b, hv, wv);
    const int hw = g.H * g.W;
    const float* src = img + (off ? off[b] : (int64_t)b * hw);
    float s = 0.f;
    for (int i = threadIdx.x; i < hv * wv; i += 256) s += sqrtf(fmaxf(src[i], 0.f) + 0.375f);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const float mean = red[0] / (float)(hv * wv);
    for (int i = threadIdx.x; i < hw; i += 256) {
        const int h = i / g.W, w = i - (i / g.W) * g.W;
        float v = 0.f;
        if (h < hv && w < wv) v = (sqrtf(fmaxf(src[h * wv + w], 0.f) + 0.375f) - mean) * 0.5f;
        x0[(int64_t)b * hw + i] = f2bf(v);
    }
}

// 1x1 head and loss.  Per pixel: logit = wh0 . a + bh0, mu = wh1 . a + bh1.
// Training (tgt != NULL): L = mean over valid pixels of
//   BCE(logit, t) + t (mu - log I_t)^2,
// writes dZ4 = dL/d(pre-activation of layer 4) (bf16) and accumulates the
// head gradients and the loss sums.  Inference: writes (logit, mu) pairs.
__global__ __launch_bounds__(256) void k_head(const uint16_t* __restrict__ a4, const float* __restrict__ wh,
                                              const float* __restrict__ bh, const float* __restrict__ t_occ,
                                              const float* __restrict__ t_li, float inv_n, uint16_t* __restrict__ dz4,
                                              float* __restrict__ dwh, float* __restrict__ dbh,
                                              float* __restrict__ loss, float* __restrict__ outp, int64_t npix,
                                              Geo g) {
    __shared__ float red[4][2 * kC + 4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int hw = g.H * g.W;
    float gw0[kC], gw1[kC];
#pragma unroll
    for (int c = 0; c < kC; c++) gw0[c] = gw1[c] = 0.f;
    float s0 = 0.f, s1 = 0.f, q0 = 0.f, q1 = 0.f;
    // grid-strided over pixels; a pixel's 32 activations and 32 gradients as
    // four 16-byte loads and stores
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < npix; p += (int64_t)gridDim.x * 256) {
        const int b = (int)(p / hw);
        const int r = (int)(p - (int64_t)b * hw);
        int hv, wv2;
        valid_hw(g, b, hv, wv2);
        const bool ok = (r / g.W) < hv && (r - (r / g.W) * g.W) < wv2;
        float a[kC];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 v = ok ? *reinterpret_cast<const uint4*>(a4 + p * kC + 8 * q) : make_uint4(0u, 0u, 0u, 0u);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 8; j++) a[8 * q + j] = bf2f((uint16_t)(w4[j >> 1] >> (16 * (j & 1))));
        }
        float h0 = bh[0], h1 = bh[1];
#pragma unroll
        for (int c = 0; c < kC; c++) {
            h0 += wh[c] * a[c];
            h1 += wh[kC + c] * a[c];
        }
        if (!t_occ) {
            outp[2 * p] = ok ? h0 : -INFINITY;
            outp[2 * p + 1] = h1;
            continue;
        }
        float g0 = 0.f, g1 = 0.f;
        if (ok) {
            const float t = t_occ[p];
            const float sg = 1.f / (1.f + expf(-h0));
            q0 += (fmaxf(h0, 0.f) - h0 * t + log1pf(expf(-fabsf(h0)))) * inv_n;
            g0 = (sg - t) * inv_n;
            const float e = h1 - t_li[p];
            q1 += t * e * e * inv_n;
            g1 = 2.f * t * e * inv_n;
        }
        s0 += g0;
        s1 += g1;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t w4[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int c0 = 8 * q + 2 * j, c1 = c0 + 1;
                const float d0 = wh[c0] * g0 + wh[kC + c0] * g1, d1 = wh[c1] * g0 + wh[kC + c1] * g1;
                w4[j] = (uint32_t)f2bf(a[c0] > 0.f ? d0 : 0.f) | ((uint32_t)f2bf(a[c1] > 0.f ? d1 : 0.f) << 16);
            }
            *reinterpret_cast<uint4*>(dz4 + p * kC + 8 * q) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
        }
#pragma unroll
        for (int c = 0; c < kC; c++) {
            gw0[c] += a[c] * g0;
            gw1[c] += a[c] * g1;
        }
    }
    if (!t_occ) return;
    // block sums (wavefront shuffles, then the four wavefronts through LDS), one atomic per block and entry
#pragma unroll
    for (int c = 0; c < kC; c++) {
        float x = gw0[c], y = gw1[c];
        for (int o = 32; o > 0; o >>= 1) {
            x += __shfl_xor(x, o);
            y += __shfl_xor(y, o);
        }
        if (lane == 0) {
            red[wv][c] = x;
            red[wv][kC + c] = y;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
        q0 += __shfl_xor(q0, o);
        q1 += __shfl_xor(q1, o);
    }
    if (lane == 0) {
        red[wv][2 * kC + 0] = s0;
        red[wv][2 * kC + 1] = s1;
        red[wv][2 * kC + 2] = q0;
        red[wv][2 * kC + 3] = q1;
    }
    __syncthreads();
    if (threadIdx.x < 2 * kC + 4) {
        const int e = threadIdx.x;
        const float v = red[0][e] + red[1][e] + red[2][e] + red[3][e];
        if (e < 2 * kC) atomicAdd(&dwh[e], v);
        else if (e < 2 * kC + 2) atomicAdd(&dbh[e - 2 * kC], v);
        else atomicAdd(&loss[e - 2 * kC - 2], v);
    }
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Adam (Kingma & Ba) on the fp32 master parameters.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Packs the fp32 masters into the bfloat16 operand layouts: forward weights
// [cout][k] of each 3x3 layer (first layer padded to 32), and for layers 2..4
// the mirrored/transposed weights of the input-gradient product:
// WT[c][tap' * 32 + co] = W[co][(8 - tap') * 32 + c].
struct Offs {
    int64_t w[4], b[4], wh, bh, total;
};
__global__ void k_pack(const float* __restrict__ prm, Offs o, uint16_t* __restrict__ w1, uint16_t* __restrict__ wf,
                       uint16_t* __restrict__ wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kC * kK1) {
        const int co = i / kK1, k = i % kK1;
        w1[i] = k < 9 ? f2bf(prm[o.w[0] + co * 9 + k]) : (uint16_t)0;
    }
    if (i < 3 * kC * kK) {
        const int l = i / (kC * kK), r = i % (kC * kK);
        const int co = r / kK, k = r % kK;
        wf[i] = f2bf(prm[o.w[l + 1] + (int64_t)co * kK + k]);
        // transposed/mirrored: row c, column tap' * 32 + co'
        const int c = co, tapp = k / kC, cop = k % kC;
        wt[i] = f2bf(prm[o.w[l + 1] + (int64_t)cop * kK + (8 - tapp) * kC + c]);
    }
}

// Proposal maps for a ROI batch from the head outputs on the 32 x 32 canvas:
// q = (1 - eps) sigmoid(logit) / sum + eps / P over the ROI's pixels (ROI
// layout data[y + x*Ysize], canvas row = x, column = y), mu = log-intensity.
__global__ __launch_bounds__(256) void k_maps(const float* __restrict__ outp, const int64_t* __restrict__ off,
                                              const int32_t* __restrict__ xs, const int32_t* __restrict__ ys,
                                              float eps, float* __restrict__ q, float* __restrict__ mu, int H, int W) {
    __shared__ float red[256];
    const int b = blockIdx.x;
    const int X = xs[b], Y = ys[b], P = X * Y;
    const int64_t base = (int64_t)b * H * W;
    float s = 0.f;
    for (int i = threadIdx.x; i < P; i += 256) {
        const int x = i / Y, y = i - (i / Y) * Y;
        s += 1.f / (1.f + expf(-outp[2 * (base + x * W + y)]));
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const float tot = red[0];
    for (int i = threadIdx.x; i < P; i += 256) {
        const int x = i / Y, y = i - (i / Y) * Y;
        const float sg = 1.f / (1.f + expf(-outp[2 * (base + x * W + y)]));
        q[off[b] + i] = (tot > 0.f ? (1.f - eps) * sg / tot : 0.f) + (tot > 0.f ? eps : 1.f) / (float)P;
        mu[off[b] + i] = outp[2 * (base + x * W + y) + 1];
    }
}

// Copies each ROI (layout data[y + x*Ysize]) into its canvas slot, and its
// size into the per-image valid region (rows = x, columns = y).
__global__ void k_canvas(const float* __restrict__ data, const int64_t* __restrict__ off,
                         const int32_t* __restrict__ xs, const int32_t* __restrict__ ys, float* __restrict__ canvas,
                         int32_t* __restrict__ vh, int32_t* __restrict__ vw, int n, int H, int W) {
    const int b = blockIdx.x;
    if (b >= n) return;
    const int X = xs[b], Y = ys[b];
    if (threadIdx.x == 0) {
        vh[b] = X;
        vw[b] = Y;
    }
    // compact per image (row length Y); the normalisation reads the valid region in this layout
    for (int i = threadIdx.x; i < X * Y; i += blockDim.x) canvas[(int64_t)b * H * W + i] = data[off[b] + i];
}

}  // namespace
}  // namespace net
}  // namespace bamf

using namespace bamf;
using namespace bamf::net;

struct bamf_net {
    int max_batch = 0, H = 0, W = 0;
    Offs o{};
    float* prm = nullptr;   // fp32 masters
    float* grad = nullptr;  // gradients (same layout)
    float* m = nullptr;
    float* v = nullptr;
    float* loss = nullptr;  // 2 floats
    uint16_t* w1 = nullptr;  // [32][32]
    uint16_t* wf = nullptr;  // 3 x [32][288]
    uint16_t* wt = nullptr;  // 3 x [32][288]
    uint16_t* x0 = nullptr;  // [npix]
    uint16_t* act[4] = {};   // [npix][32]
    uint16_t* dz[2] = {};    // ping-pong gradients [npix][32]
    float* outp = nullptr;   // [npix][2]
    float* canvas = nullptr; // [batch][H][W] for proposal maps
    int32_t* vh = nullptr;
    int32_t* vw = nullptr;
    int64_t step = 0;
};

extern "C" int bamf_net_destroy(bamf_net* n);

namespace {
Offs make_offs() {
    Offs o;
    int64_t at = 0;
    o.w[0] = at;
    at += kC * 9;
    o.b[0] = at;
    at += kC;
    for (int l = 1; l < 4; l++) {
        o.w[l] = at;
        at += (int64_t)kC * kK;
        o.b[l] = at;
        at += kC;
    }
    o.wh = at;
    at += 2 * kC;
    o.bh = at;
    at += 2;
    o.total = at;
    return o;
}

int pack(bamf_net* n, hipStream_t s) {
    hipLaunchKernelGGL(k_pack, dim3((unsigned)ceil_div(3 * kC * kK, 256)), dim3(256), 0, s, n->prm, n->o, n->w1,
                       n->wf, n->wt);
    BAMF_LAUNCH_CHECK("k_pack");
    return BAMF_OK;
}

int fill(float* p, int64_t cnt, float v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)ceil_div(cnt, 256)), dim3(256), 0, s, p, cnt, v);
    BAMF_LAUNCH_CHECK("k_fill_f32");
    return BAMF_OK;
}

// forward through the four 3x3 layers (x0 must be prepared)
int forward_convs(bamf_net* n, int64_t npix, Geo g, hipStream_t s) {
    const dim3 grid((unsigned)ceil_div(npix, 256)), blk(256);
    hipLaunchKernelGGL((k_conv3<1, 0>), grid, blk, 0, s, n->x0, n->w1, n->prm + n->o.b[0], nullptr, n->act[0], npix, g);
    BAMF_LAUNCH_CHECK("k_conv3<1>");
    for (int l = 1; l < 4; l++) {
        hipLaunchKernelGGL((k_conv3<kC, 0>), grid, blk, 0, s, n->act[l - 1], n->wf + (size_t)(l - 1) * kC * kK,
                           n->prm + n->o.b[l], nullptr, n->act[l], npix, g);
        BAMF_LAUNCH_CHECK("k_conv3<32>");
    }
    return BAMF_OK;
}

// loss and gradients for one batch (grad must be zeroed by the caller)
int loss_and_grad(bamf_net* n, int batch, const float* d_occ, const float* d_li, Geo g, hipStream_t s) {
    const int64_t npix = (int64_t)batch * n->H * n->W;
    int rc = forward_convs(n, npix, g, s);
    if (rc) return rc;
    const float inv_n = 1.0f / (float)npix;
    // at most 256 work-groups: the head gradient's atomics are one per work-group and entry
    const unsigned hg = (unsigned)(ceil_div(npix, 256) < 256 ? ceil_div(npix, 256) : 256);
    hipLaunchKernelGGL(k_head, dim3(hg), dim3(256), 0, s, n->act[3], n->prm + n->o.wh, n->prm + n->o.bh, d_occ, d_li,
                       inv_n, n->dz[0], n->grad + n->o.wh, n->grad + n->o.bh, n->loss, n->outp, npix, g);
    BAMF_LAUNCH_CHECK("k_head");
    // back through layers 4..2 (dz[cur] = dL/dZ_l), then layer 1's weights
    const int P = g.H * g.W, Pp = (P + 31) & ~31, HW2 = (g.H + 2) * (g.W + 2);
    const size_t lds32 = (size_t)32 * Pp * 2 + (size_t)((kC * HW2 + 1) & ~1) * 2 + (size_t)Pp * 4;
    const size_t lds1 = (size_t)32 * Pp * 2 + (size_t)((HW2 + 1) & ~1) * 2 + (size_t)Pp * 4;
    BAMF_CHECK_ARG(lds32 <= 160 * 1024, "gradients need an image of at most about 1,500 pixels");
    if (lds32 > 64 * 1024)
        BAMF_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3_wgrad<kC>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds32));
    if (lds1 > 64 * 1024)
        BAMF_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3_wgrad<1>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1));
    const unsigned wg = (unsigned)(batch < 64 ? batch : 64);
    int cur = 0;
    for (int l = 3; l >= 1; l--) {
        hipLaunchKernelGGL((k_conv3_wgrad<kC>), dim3(wg), dim3(256), lds32, s, n->dz[cur], n->act[l - 1],
                           n->grad + n->o.w[l], n->grad + n->o.b[l], batch, g);
        BAMF_LAUNCH_CHECK("k_conv3_wgrad<32>");
        hipLaunchKernelGGL((k_conv3<kC, 1>), dim3((unsigned)ceil_div(npix, 256)), dim3(256), 0, s, n->dz[cur],
                           n->wt + (size_t)(l - 1) * kC * kK, nullptr, n->act[l - 1], n->dz[cur ^ 1], npix, g);
        BAMF_LAUNCH_CHECK("k_conv3<32,grad>");
        cur ^= 1;
    }
    hipLaunchKernelGGL((k_conv3_wgrad<1>), dim3(wg), dim3(256), lds1, s, n->dz[cur], n->x0, n->grad + n->o.w[0],
                       n->grad + n->o.b[0], batch, g);
    BAMF_LAUNCH_CHECK("k_conv3_wgrad<1>");
    return BAMF_OK;
}

int prep(bamf_net* n, const float* d_img, int batch, Geo g, hipStream_t s, const int64_t* off = nullptr) {
    hipLaunchKernelGGL(k_prep, dim3(batch), dim3(256), 0, s, d_img, off, n->x0, g);
    BAMF_LAUNCH_CHECK("k_prep");
    return BAMF_OK;
}
}  // namespace

extern "C" {

int bamf_net_param_count(void) { return (int)make_offs().total; }

int bamf_net_create(int max_batch, int h, int w, const float* h_params, bamf_net** out) {
    BAMF_CHECK_ARG(out && h_params, "null argument");
    BAMF_CHECK_ARG(max_batch >= 1 && h >= 1 && w >= 1 && h <= 64 && w <= 64, "bad network geometry");
    auto* n = new bamf_net();
    n->max_batch = max_batch;
    n->H = h;
    n->W = w;
    n->o = make_offs();
    const int64_t P = n->o.total, npix = (int64_t)max_batch * h * w;
    hipError_t e = hipSuccess;
    auto al = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    al((void**)&n->prm, P * 4);
    al((void**)&n->grad, P * 4);
    al((void**)&n->m, P * 4);
    al((void**)&n->v, P * 4);
    al((void**)&n->loss, 16);
    al((void**)&n->w1, kC * kK1 * 2);
    al((void**)&n->wf, 3 * kC * kK * 2);
    al((void**)&n->wt, 3 * kC * kK * 2);
    al((void**)&n->x0, npix * 2);
    for (int l = 0; l < 4; l++) al((void**)&n->act[l], npix * kC * 2);
    for (int l = 0; l < 2; l++) al((void**)&n->dz[l], npix * kC * 2);
    al((void**)&n->outp, npix * 8);
    al((void**)&n->canvas, npix * 4);
    al((void**)&n->vh, max_batch * 4);
    al((void**)&n->vw, max_batch * 4);
    if (e == hipSuccess) e = hipMemcpy(n->prm, h_params, P * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(n->m, 0, P * 4);
    if (e == hipSuccess) e = hipMemset(n->v, 0, P * 4);
    if (e == hipSuccess && pack(n, nullptr) != BAMF_OK) e = hipErrorUnknown;
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        set_error("network allocation failed: %s", hipGetErrorString(e));
        bamf_net_destroy(n);
        return BAMF_E_HIP;
    }
    *out = n;
    return BAMF_OK;
}

int bamf_net_destroy(bamf_net* n) {
    if (!n) return BAMF_OK;
    void* ps[] = {n->prm, n->grad, n->m, n->v, n->loss, n->w1, n->wf, n->wt, n->x0, n->act[0], n->act[1],
                  n->act[2], n->act[3], n->dz[0], n->dz[1], n->outp, n->canvas, n->vh, n->vw};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete n;
    return BAMF_OK;
}

int bamf_net_get_params(bamf_net* n, float* h_params) {
    BAMF_CHECK_ARG(n && h_params, "null argument");
    BAMF_HIP_TRY(hipMemcpy(h_params, n->prm, n->o.total * 4, hipMemcpyDeviceToHost));
    return BAMF_OK;
}

int bamf_net_forward(bamf_net* n, const float* d_images, int batch, float* d_out, void* stream) {
    BAMF_CHECK_ARG(n && d_images && d_out, "null argument");
    BAMF_CHECK_ARG(batch >= 1 && batch <= n->max_batch, "batch outside 1..max_batch");
    hipStream_t s = as_stream(stream);
    const Geo g{n->H, n->W, nullptr, nullptr};
    const int64_t npix = (int64_t)batch * n->H * n->W;
    int rc = prep(n, d_images, batch, g, s);
    if (!rc) rc = forward_convs(n, npix, g, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_head, dim3((unsigned)ceil_div(npix, 256)), dim3(256), 0, s, n->act[3], n->prm + n->o.wh,
                       n->prm + n->o.bh, nullptr, nullptr, 0.f, nullptr, nullptr, nullptr, nullptr, d_out, npix, g);
    BAMF_LAUNCH_CHECK("k_head");
    return BAMF_OK;
}

int bamf_net_gradients(bamf_net* n, const float* d_images, const float* d_occ, const float* d_logi, int batch,
                       float* d_grad, float* d_loss, void* stream) {
    BAMF_CHECK_ARG(n && d_images && d_occ && d_logi, "null argument");
    BAMF_CHECK_ARG(batch >= 1 && batch <= n->max_batch, "batch outside 1..max_batch");
    hipStream_t s = as_stream(stream);
    const Geo g{n->H, n->W, nullptr, nullptr};
    int rc = fill(n->grad, n->o.total, 0.f, s);
    if (!rc) rc = fill(n->loss, 2, 0.f, s);
    if (!rc) rc = prep(n, d_images, batch, g, s);
    if (!rc) rc = loss_and_grad(n, batch, d_occ, d_logi, g, s);
    if (rc) return rc;
    if (d_grad) BAMF_HIP_TRY(hipMemcpyAsync(d_grad, n->grad, n->o.total * 4, hipMemcpyDeviceToDevice, s));
    if (d_loss) BAMF_HIP_TRY(hipMemcpyAsync(d_loss, n->loss, 8, hipMemcpyDeviceToDevice, s));
    return BAMF_OK;
}

int bamf_net_train_step(bamf_net* n, const float* d_images, const float* d_occ, const float* d_logi, int batch,
                        float lr, float* d_loss, void* stream) {
    int rc = bamf_net_gradients(n, d_images, d_occ, d_logi, batch, nullptr, d_loss, stream);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)ceil_div(n->o.total, 256)), dim3(256), 0, s, n->prm, n->grad, n->m,
                       n->v, n->o.total, lr, b1, b2, 1e-8f, c1, c2);
    BAMF_LAUNCH_CHECK("k_adam");
    return pack(n, s);
}

int bamf_net_proposal_maps(bamf_net* n, const bamf_roi_batch* batch, float eps, float* d_q, float* d_mu,
                           void* stream) {
    BAMF_CHECK_ARG(n && batch && d_q && d_mu, "null argument");
    BAMF_CHECK_ARG(n->H >= 32 && n->W >= 32, "proposal maps need a network built for a 32 x 32 canvas");
    BAMF_CHECK_ARG(eps > 0.f && eps <= 1.f, "eps must be in (0, 1]");
    hipStream_t s = as_stream(stream);
    for (int b0 = 0; b0 < batch->n_roi; b0 += n->max_batch) {
        const int nb = batch->n_roi - b0 < n->max_batch ? batch->n_roi - b0 : n->max_batch;
        hipLaunchKernelGGL(k_canvas, dim3(nb), dim3(256), 0, s, batch->d_data, batch->d_pix_offset + b0,
                           batch->d_xsize + b0, batch->d_ysize + b0, n->canvas, n->vh, n->vw, nb, n->H, n->W);
        BAMF_LAUNCH_CHECK("k_canvas");
        const Geo g{n->H, n->W, n->vh, n->vw};
        const int64_t npix = (int64_t)nb * n->H * n->W;
        int rc = prep(n, n->canvas, nb, g, s);
        if (!rc) rc = forward_convs(n, npix, g, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_head, dim3((unsigned)ceil_div(npix, 256)), dim3(256), 0, s, n->act[3],
                           n->prm + n->o.wh, n->prm + n->o.bh, nullptr, nullptr, 0.f, nullptr, nullptr, nullptr,
                           nullptr, n->outp, npix, g);
        BAMF_LAUNCH_CHECK("k_head");
        hipLaunchKernelGGL(k_maps, dim3(nb), dim3(256), 0, s, n->outp, batch->d_pix_offset + b0,
                           batch->d_xsize + b0, batch->d_ysize + b0, eps, d_q, d_mu, n->H, n->W);
        BAMF_LAUNCH_CHECK("k_maps");
    }
    return BAMF_OK;
}

}  // extern "C"
