// Difference-of-Gaussian ROI finder (source/cuda/cuda_FindROI/cuda_FindROI.cu
// and the MATLAB driver @RJ/findROI.m:60-225), resident on the device.
//
// Bit-exactness: each output value is computed with the reference's
// expression and operation order, in single precision, with FMA contraction
// off and IEEE division (HIP's default), so results match the reference's
// semantics bit for bit regardless of how the work is tiled.
//
// gfx950 design
//   rows pass   k_rows_iir   : 64 rows per work-group staged through LDS in
//                              64-sample chunks (coalesced 256 B row segments),
//                              the recursion for one or two sigmas per row;
//   column pass k_cols_iir   : one work-item per column, adjacent lanes on
//                              adjacent columns (coalesced), forward pass in
//                              place, backward pass fused with the DoG subtract;
//   local max   k_localmax   : 64x32 output tile, D and the row-max image held
//                              in LDS with a kernelsz halo, bits set with sparse
//                              atomicOr into a Res bitmap;
//   compaction  k_bits_*     : ordered stream compaction of the bitmap
//                              (per-block popcount, scan, ordered emit) plus the
//                              box corners of findROI.m:130-148.
#include <cmath>

#include "common.hpp"

#pragma clang fp contract(off)

namespace bamf {
namespace {

constexpr int kRowsPerBlock = 64;
constexpr int kChunk = 32;

struct Coeffs {
    float b0, b1, b2, b3, B;
    float rb0;  // 1 / b0 correctly rounded (set by with_rcp on the host)
};

// The correctly rounded reciprocal of b: the float nearest 1/(double)b, or
// the neighbour whose product with b (exact in double) is closer to 1.
inline Coeffs with_rcp(Coeffs c) {
    const float y = (float)(1.0 / (double)c.b0);
    float best = y;
    double err = std::fabs(1.0 - (double)y * (double)c.b0);
    for (float z : {std::nextafter(y, 0.0f), std::nextafter(y, 2.0f * y)}) {
        const double e = std::fabs(1.0 - (double)z * (double)c.b0);
        if (e < err) {
            err = e;
            best = z;
        }
    }
    c.rb0 = best;
    return c;
}

__device__ inline float iir_step(float w0, float w1, float w2, float w3, const Coeffs& c) {
    // cuda_FindROI.cu:28 — temp = w0*B + (b1*w1 + b2*w2 + b3*w3) / b0.  The
    // quotient is the correctly rounded x / b0 formed from the correctly
    // rounded reciprocal with one residual correction (Markstein's theorem:
    // q0 = RN(x y), r = x - q0 b0 exact by fma, RN(q0 + r y) = RN(x / b0) when
    // y = RN(1 / b0)), three instructions instead of the division sequence.
    const float x = c.b1 * w1 + c.b2 * w2 + c.b3 * w3;
    const float q0 = x * c.rb0;
    const float q = fmaf(fmaf(-q0, c.b0, x), c.rb0, q0);
    return w0 * c.B + q;
}

// Row recursion (the reference's gaussX) for NS sigmas at once.  in may alias out0.
// The chunk is staged in LDS; sigma1's results overwrite the staged input in
// place (each input sample is read before it is overwritten).
template <int NS>
__global__ __launch_bounds__(kRowsPerBlock) void k_rows_iir(const float* in, float* out0, float* out1,
                                                            int xsize, int64_t nrows, Coeffs c0,
                                                            Coeffs c1) {
    __shared__ float tA[kRowsPerBlock][kChunk + 1];
    __shared__ float tB[NS == 2 ? kRowsPerBlock : 1][kChunk + 1];
    const int t = threadIdx.x;
    const int64_t row0 = (int64_t)blockIdx.x * kRowsPerBlock;
    const int nr = (int)min<int64_t>(kRowsPerBlock, nrows - row0);
    float a1 = 0.f, a2 = 0.f, a3 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
    const int nchunk = (xsize + kChunk - 1) / kChunk;

    // ---- forward (cuda_FindROI.cu:23-33) ----
    for (int ch = 0; ch < nchunk; ch++) {
        const int x0 = ch * kChunk;
        const int nc = min(kChunk, xsize - x0);
        for (int e = t; e < nr * nc; e += kRowsPerBlock) {
            const int r = e / nc, c = e - r * nc;
            tA[r][c] = in[(row0 + r) * xsize + x0 + c];
        }
        __syncthreads();
        if (t < nr) {
            if (ch == 0) {
                a1 = a2 = a3 = tA[t][0];
                b1 = b2 = b3 = a1;
            }
            for (int k = 0; k < nc; k++) {
                const float w0 = tA[t][k];
                const float va = iir_step(w0, a1, a2, a3, c0);
                tA[t][k] = va;
                a3 = a2; a2 = a1; a1 = va;
                if (NS == 2) {
                    const float vb = iir_step(w0, b1, b2, b3, c1);
                    tB[t][k] = vb;
                    b3 = b2; b2 = b1; b1 = vb;
                }
            }
        }
        __syncthreads();
        for (int e = t; e < nr * nc; e += kRowsPerBlock) {
            const int r = e / nc, c = e - r * nc;
            out0[(row0 + r) * xsize + x0 + c] = tA[r][c];
            if (NS == 2) out1[(row0 + r) * xsize + x0 + c] = tB[r][c];
        }
        __syncthreads();
    }
    // ---- backward (cuda_FindROI.cu:35-45), on the forward results ----
    for (int ch = nchunk - 1; ch >= 0; ch--) {
        const int x0 = ch * kChunk;
        const int nc = min(kChunk, xsize - x0);
        for (int e = t; e < nr * nc; e += kRowsPerBlock) {
            const int r = e / nc, c = e - r * nc;
            tA[r][c] = out0[(row0 + r) * xsize + x0 + c];
            if (NS == 2) tB[r][c] = out1[(row0 + r) * xsize + x0 + c];
        }
        __syncthreads();
        if (t < nr) {
            if (ch == nchunk - 1) {
                a1 = a2 = a3 = tA[t][nc - 1];
                if (NS == 2) b1 = b2 = b3 = tB[t][nc - 1];
            }
            for (int k = nc - 1; k >= 0; k--) {
                const float va = iir_step(tA[t][k], a1, a2, a3, c0);
                tA[t][k] = va;
                a3 = a2; a2 = a1; a1 = va;
                if (NS == 2) {
                    const float vb = iir_step(tB[t][k], b1, b2, b3, c1);
                    tB[t][k] = vb;
                    b3 = b2; b2 = b1; b1 = vb;
                }
            }
        }
        __syncthreads();
        for (int e = t; e < nr * nc; e += kRowsPerBlock) {
            const int r = e / nc, c = e - r * nc;
            out0[(row0 + r) * xsize + x0 + c] = tA[r][c];
            if (NS == 2) out1[(row0 + r) * xsize + x0 + c] = tB[r][c];
        }
        __syncthreads();
    }
}

// Column recursion (the reference's gaussY).  DOG=false: in place on a.
// DOG=true: a and b hold the row-filtered images for sigma1 and sigma2; both
// are column-filtered and a receives G(sigma1) - G(sigma2) (kernel_subtract).
template <bool DOG>
__global__ __launch_bounds__(64) void k_cols_iir(float* a, float* b, int xsize, int ysize, Coeffs c0,
                                                 Coeffs c1) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int64_t z = blockIdx.y;
    if (x >= xsize) return;
    const int64_t base = z * xsize * ysize + x;
    const int64_t st = xsize;
    float* pa = a + base;
    float* pb = b + base;
    // forward (cuda_FindROI.cu:60-70)
    float a1, a2, a3, b1 = 0.f, b2 = 0.f, b3 = 0.f;
    a1 = a2 = a3 = pa[0];
    if (DOG) b1 = b2 = b3 = pb[0];
    for (int ii = 0; ii < ysize; ii++) {
        const float va = iir_step(pa[ii * st], a1, a2, a3, c0);
        pa[ii * st] = va;
        a3 = a2; a2 = a1; a1 = va;
        if (DOG) {
            const float vb = iir_step(pb[ii * st], b1, b2, b3, c1);
            pb[ii * st] = vb;
            b3 = b2; b2 = b1; b1 = vb;
        }
    }
    // backward (cuda_FindROI.cu:72-82)
    a1 = a2 = a3 = pa[(int64_t)(ysize - 1) * st];
    if (DOG) b1 = b2 = b3 = pb[(int64_t)(ysize - 1) * st];
    for (int ii = ysize - 1; ii >= 0; ii--) {
        const float va = iir_step(pa[ii * st], a1, a2, a3, c0);
        a3 = a2; a2 = a1; a1 = va;
        if (DOG) {
            const float vb = iir_step(pb[ii * st], b1, b2, b3, c1);
            b3 = b2; b2 = b1; b1 = vb;
            pa[ii * st] = va - vb;  // cuda_FindROI.cu:91
        } else {
            pa[ii * st] = va;
        }
    }
}

__global__ void k_subtract(float* a, const float* b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = a[i] - b[i];
}

// maxX output value (cuda_FindROI.cu:115).
__device__ inline float maxx_value(float maxval, float inpixel) {
    const float t1 = (float)(maxval > inpixel) * -maxval;
    const float t2 = (float)(maxval == inpixel) * maxval;
    return t1 + t2;
}
// maxY flag test (cuda_FindROI.cu:144): |max - in| < 1e-6 compared in double.
// 1e-6f is the largest float below 1e-6, so for a float v, (double)v < 1e-6
// holds exactly when v <= 1e-6f.
__device__ inline bool maxy_flag(float maxval, float inpixel) { return fabsf(maxval - inpixel) <= 1e-6f; }

__global__ void k_max_x(const float* a, float* b, int xsize, int ysize, int k, float minval) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int64_t rowi = blockIdx.y;  // y + ysize*z
    if (x >= xsize) return;
    const float* row = a + rowi * xsize;
    const int start = max(0, x - k), end = min(xsize - 1, x + k);
    float maxval = minval;
    const float in = row[x];
    for (int ii = start; ii < end + 1; ii++) maxval = fmaxf(maxval, row[ii]);
    b[rowi * xsize + x] = maxx_value(maxval, in);
}

__global__ void k_max_y(float* a, const float* b, int xsize, int ysize, int k, float minval) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y;
    const int64_t z = blockIdx.z;
    if (x >= xsize) return;
    const int64_t base = z * xsize * ysize;
    const int64_t idx = x + (int64_t)xsize * y + base;
    const int start = max(0, y - k), end = min(ysize - 1, y + k);
    float maxval = minval;
    const float in = b[idx];
    for (int ii = start; ii <= end; ii++) maxval = fmaxf(maxval, fabsf(b[(int64_t)ii * xsize + x + base]));
    a[idx] = maxy_flag(maxval, in) ? 1.0f : 0.0f;
}

// Fused maxX + maxY on the DoG image, writing Res bits.  Window half-width K <= kMaxK.
constexpr int kTX = 64, kTY = 32, kMaxK = 16;
__global__ __launch_bounds__(256) void k_localmax(const float* __restrict__ dog, uint32_t* __restrict__ bits,
                                                  int xsize, int ysize, int k, float minval) {
    __shared__ float sD[kTY + 2 * kMaxK][kTX + 2 * kMaxK + 1];
    __shared__ float sM[kTY + 2 * kMaxK][kTX + 1];
    const int x0 = blockIdx.x * kTX, y0 = blockIdx.y * kTY;
    const int64_t z = blockIdx.z;
    const int64_t fbase = z * xsize * ysize;
    const int t = threadIdx.x;
    const int dw = kTX + 2 * k, dh = kTY + 2 * k;
    // D tile: rows y0-k .. y0+kTY+k-1, cols x0-k .. x0+kTX+k-1 (outside the frame: -inf,
    // which never changes fmaxf(maxval, .) so the clipped windows are reproduced).
    for (int e = t; e < dw * dh; e += 256) {
        const int r = e / dw, cc = e - r * dw;
        const int y = y0 - k + r, x = x0 - k + cc;
        float v = -INFINITY;
        if (y >= 0 && y < ysize && x >= 0 && x < xsize) v = dog[fbase + (int64_t)y * xsize + x];
        sD[r][cc] = v;
    }
    __syncthreads();
    // row-max image (maxX) for the tile columns and all halo rows
    for (int e = t; e < kTX * dh; e += 256) {
        const int r = e / kTX, cc = e - r * kTX;
        float maxval = minval;
        for (int q = 0; q <= 2 * k; q++) maxval = fmaxf(maxval, sD[r][cc + q]);
        sM[r][cc] = maxx_value(maxval, sD[r][cc + k]);
    }
    __syncthreads();
    // column max of |maxX| (maxY) and the flag
    for (int e = t; e < kTX * kTY; e += 256) {
        const int r = e / kTX, cc = e - r * kTX;
        const int y = y0 + r, x = x0 + cc;
        if (y >= ysize || x >= xsize) continue;
        const int ys = max(0, y - k), ye = min(ysize - 1, y + k);
        float maxval = minval;
        for (int yy = ys; yy <= ye; yy++) maxval = fmaxf(maxval, fabsf(sM[yy - y0 + k][cc]));
        if (maxy_flag(maxval, sM[r + k][cc])) {
            const int64_t lin = fbase + (int64_t)y * xsize + x;
            atomicOr(&bits[lin >> 5], 1u << (lin & 31));
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Float map (0/1) -> bits, for the unfused path.
__global__ void k_pack_bits(const float* res, uint32_t* bits, int64_t n) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 32 >= n) return;
    uint32_t v = 0;
    for (int b = 0; b < 32; b++) {
        const int64_t i = w * 32 + b;
        if (i < n && res[i] != 0.0f) v |= 1u << b;
    }
    bits[w] = v;
}

constexpr int kWordsPerThread = 8;
constexpr int kCompactThreads = 256;
constexpr int kWordsPerBlock = kWordsPerThread * kCompactThreads;

__global__ __launch_bounds__(kCompactThreads) void k_bits_count(const uint32_t* bits, int64_t nwords,
                                                                int64_t* block_counts) {
    __shared__ int64_t red[kCompactThreads];
    const int t = threadIdx.x;
    const int64_t w0 = (int64_t)blockIdx.x * kWordsPerBlock + (int64_t)t * kWordsPerThread;
    int64_t c = 0;
    for (int q = 0; q < kWordsPerThread; q++)
        if (w0 + q < nwords) c += __popc(bits[w0 + q]);
    red[t] = c;
    __syncthreads();
    for (int s = kCompactThreads / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) block_counts[blockIdx.x] = red[0];
}

// Exclusive scan of nb block counts in place (one work-group); total -> *count.
__global__ __launch_bounds__(1024) void k_bits_scan(int64_t* counts, int64_t nb, int64_t* count) {
    __shared__ int64_t s[1024];
    __shared__ int64_t carry;
    const int t = threadIdx.x;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + t;
        const int64_t v = i < nb ? counts[i] : 0;
        s[t] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int64_t add = t >= off ? s[t - off] : 0;
            __syncthreads();
            s[t] += add;
            __syncthreads();
        }
        if (i < nb) counts[i] = carry + s[t] - v;
        __syncthreads();
        if (t == 1023) carry += s[1023];
        __syncthreads();
    }
    if (t == 0) *count = carry;
}

__global__ __launch_bounds__(kCompactThreads) void k_bits_emit(const uint32_t* bits, int64_t nwords,
                                                               const int64_t* block_offsets, int xsize,
                                                               int ysize, int boxsize, int64_t capacity,
                                                               int64_t* lin_idx, int32_t* start_row,
                                                               int32_t* start_col, int32_t* frame) {
    __shared__ int64_t s[kCompactThreads];
    const int t = threadIdx.x;
    const int64_t w0 = (int64_t)blockIdx.x * kWordsPerBlock + (int64_t)t * kWordsPerThread;
    uint32_t wv[kWordsPerThread];
    int64_t c = 0;
    for (int q = 0; q < kWordsPerThread; q++) {
        wv[q] = (w0 + q < nwords) ? bits[w0 + q] : 0u;
        c += __popc(wv[q]);
    }
    s[t] = c;
    __syncthreads();
    for (int off = 1; off < kCompactThreads; off <<= 1) {
        const int64_t add = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    int64_t pos = block_offsets[blockIdx.x] + s[t] - c;
    const int halfbox = (int)floor(boxsize / 2.0 + 0.5);  // findROI.m:130
    const int64_t fr = (int64_t)xsize * ysize;
    for (int q = 0; q < kWordsPerThread; q++) {
        uint32_t v = wv[q];
        while (v) {
            const int b = __ffs(v) - 1;
            v &= v - 1;
            if (pos < capacity) {
                const int64_t lin = (w0 + q) * 32 + b;
                // ind2sub(size(Res) = [Ysize Xsize Z], lin+1) -> 1-based (row, col, frame)
                const int ii = (int)(lin % ysize) + 1;
                const int jj = (int)((lin / ysize) % xsize) + 1;
                const int kk = (int)(lin / fr) + 1;
                // findROI.m:131-148
                int sr = ii - halfbox + 1, sc = jj - halfbox + 1;
                if (sr < 1) sr = 1;
                if (sc < 1) sc = 1;
                if (sr + boxsize - 1 > ysize) sr = ysize - boxsize + 1;
                if (sc + boxsize - 1 > xsize) sc = xsize - boxsize + 1;
                if (lin_idx) lin_idx[pos] = lin;
                if (start_row) start_row[pos] = sr;
                if (start_col) start_col[pos] = sc;
                if (frame) frame[pos] = kk;
            }
            pos++;
        }
    }
}

__global__ void k_crop(const float* stack, int xsize, int ysize, const int32_t* sr, const int32_t* sc,
                       const int32_t* fr, int64_t n, int boxsize, float* out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)boxsize * boxsize;
    if (e >= n * per) return;
    const int64_t b = e / per;
    const int rc = (int)(e - b * per);
    const int c = rc / boxsize, r = rc - c * boxsize;
    // Image(StartRow:EndRow, StartCol:EndCol, Kk) in MATLAB column-major layout
    const int64_t src = (int64_t)(sr[b] - 1 + r) + (int64_t)ysize * (sc[b] - 1 + c) +
                        (int64_t)ysize * xsize * (fr[b] - 1);
    out[e] = stack[src];
}

Coeffs make_coeffs(float sigma) {
    float o[5];
    bamf_gauss_coeffs(sigma, o);
    return with_rcp(Coeffs{o[0], o[1], o[2], o[3], o[4], 0.0f});
}

int launch_rows(const float* in, float* o0, float* o1, int xsize, int64_t nrows, Coeffs c0, Coeffs c1,
                int ns, hipStream_t s) {
    dim3 grid((unsigned)ceil_div(nrows, kRowsPerBlock));
    if (ns == 1)
        hipLaunchKernelGGL(k_rows_iir<1>, grid, dim3(kRowsPerBlock), 0, s, in, o0, o1, xsize, nrows, c0, c1);
    else
        hipLaunchKernelGGL(k_rows_iir<2>, grid, dim3(kRowsPerBlock), 0, s, in, o0, o1, xsize, nrows, c0, c1);
    BAMF_LAUNCH_CHECK("k_rows_iir");
    return BAMF_OK;
}

}  // namespace
}  // namespace bamf

using namespace bamf;

extern "C" {

void bamf_gauss_coeffs(float sigma, float out[5]) {
    // findROI.m:169-182, double precision, handed to the single-precision kernels.
    double s = (double)sigma, q;
    if (s > 2.5) q = 0.98711 * s - 0.96330;
    else q = 3.97156 - 4.14554 * std::sqrt(1 - 0.26891 * s);
    const double qq = q * q, qqq = qq * q;
    const double b0 = 1.57825 + 2.44413 * q + 1.4281 * qq + 0.422205 * qqq;
    const double b1 = 2.44413 * q + 2.85619 * qq + 1.26661 * qqq;
    const double b2 = -1.4281 * qq - 1.26661 * qqq;
    const double b3 = 0.422205 * qqq;
    const double B = 1 - (b1 + b2 + b3) / b0;
    out[0] = (float)b0; out[1] = (float)b1; out[2] = (float)b2; out[3] = (float)b3; out[4] = (float)B;
}

int bamf_find_roi_gauss_x(float* d, int xsize, int ysize, int nframes, float b0, float b1, float b2,
                          float b3, float B, void* stream) {
    BAMF_CHECK_ARG(xsize > 0 && ysize > 0 && nframes >= 0, "bad sizes");
    if (nframes == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d, "null device pointer");
    const Coeffs c = with_rcp(Coeffs{b0, b1, b2, b3, B, 0.0f});
    return launch_rows(d, d, nullptr, xsize, (int64_t)ysize * nframes, c, c, 1, as_stream(stream));
}

int bamf_find_roi_gauss_y(float* d, int xsize, int ysize, int nframes, float b0, float b1, float b2,
                          float b3, float B, void* stream) {
    BAMF_CHECK_ARG(xsize > 0 && ysize > 0 && nframes >= 0, "bad sizes");
    if (nframes == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d, "null device pointer");
    BAMF_CHECK_ARG(nframes <= 65535, "nframes must be <= 65535 per call");
    const Coeffs c = with_rcp(Coeffs{b0, b1, b2, b3, B, 0.0f});
    dim3 grid((unsigned)ceil_div(xsize, 64), (unsigned)nframes);
    hipLaunchKernelGGL(k_cols_iir<false>, grid, dim3(64), 0, as_stream(stream), d, d, xsize, ysize, c, c);
    BAMF_LAUNCH_CHECK("k_cols_iir");
    return BAMF_OK;
}

int bamf_find_roi_subtract(float* d_a, const float* d_b, int64_t n, void* stream) {
    BAMF_CHECK_ARG(n >= 0, "bad n");
    if (n == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_a && d_b, "null device pointer");
    hipLaunchKernelGGL(k_subtract, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, as_stream(stream), d_a, d_b, n);
    BAMF_LAUNCH_CHECK("k_subtract");
    return BAMF_OK;
}

int bamf_find_roi_max_x(const float* d_a, float* d_b, int xsize, int ysize, int nframes, int kernelsz,
                        float minval, void* stream) {
    BAMF_CHECK_ARG(xsize > 0 && ysize > 0 && nframes >= 0 && kernelsz >= 0, "bad sizes");
    if (nframes == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_a && d_b, "null device pointer");
    const int64_t nrows = (int64_t)ysize * nframes;
    BAMF_CHECK_ARG(nrows <= 0x7fffffff, "too many rows");
    dim3 grid((unsigned)ceil_div(xsize, 64), (unsigned)nrows);
    hipLaunchKernelGGL(k_max_x, grid, dim3(64), 0, as_stream(stream), d_a, d_b, xsize, ysize, kernelsz, minval);
    BAMF_LAUNCH_CHECK("k_max_x");
    return BAMF_OK;
}

int bamf_find_roi_max_y(float* d_a, const float* d_b, int xsize, int ysize, int nframes, int kernelsz,
                        float minval, void* stream) {
    BAMF_CHECK_ARG(xsize > 0 && ysize > 0 && nframes >= 0 && kernelsz >= 0, "bad sizes");
    if (nframes == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_a && d_b, "null device pointer");
    BAMF_CHECK_ARG(ysize <= 65535 && nframes <= 65535, "ysize and nframes must be <= 65535");
    dim3 grid((unsigned)ceil_div(xsize, 64), (unsigned)ysize, (unsigned)nframes);
    hipLaunchKernelGGL(k_max_y, grid, dim3(64), 0, as_stream(stream), d_a, d_b, xsize, ysize, kernelsz, minval);
    BAMF_LAUNCH_CHECK("k_max_y");
    return BAMF_OK;
}

static int64_t nwords_of(int64_t n) { return (n + 31) / 32; }
static int64_t nblocks_of(int64_t nwords) { return (nwords + kWordsPerBlock - 1) / kWordsPerBlock; }
static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

size_t bamf_find_roi_workspace_bytes(int xsize, int ysize, int nframes) {
    const int64_t n = (int64_t)xsize * ysize * nframes;
    const int64_t nw = nwords_of(n);
    return align256((size_t)n * 4) * 2 + align256((size_t)nw * 4) + align256((size_t)nblocks_of(nw) * 8);
}

int bamf_find_roi(const float* d_stack, int xsize, int ysize, int nframes, float sigma1, float sigma2,
                  int boxsize, float minval, void* d_workspace, size_t workspace_bytes, uint32_t* d_res_bits,
                  int64_t* d_lin_idx, int32_t* d_start_row, int32_t* d_start_col, int32_t* d_frame,
                  int64_t capacity, int64_t* d_count, void* stream) {
    BAMF_CHECK_ARG(xsize > 0 && ysize > 0 && nframes >= 0, "bad sizes");
    BAMF_CHECK_ARG(boxsize >= 0, "boxsize must be >= 0");
    BAMF_CHECK_ARG(capacity >= 0, "capacity must be >= 0");
    BAMF_CHECK_ARG(d_count, "d_count is required");
    BAMF_CHECK_ARG(nframes <= 65535, "nframes must be <= 65535 per call");
    const size_t need = bamf_find_roi_workspace_bytes(xsize, ysize, nframes);
    BAMF_CHECK_ARG(d_workspace && workspace_bytes >= need, "workspace too small (%zu < %zu)", workspace_bytes, need);
    BAMF_CHECK_ARG(d_stack, "null stack");
    hipStream_t s = as_stream(stream);
    const int64_t n = (int64_t)xsize * ysize * nframes;
    const int64_t nw = nwords_of(n), nb = nblocks_of(nw);
    char* ws = static_cast<char*>(d_workspace);
    float* A = reinterpret_cast<float*>(ws);
    float* Bf = reinterpret_cast<float*>(ws + align256((size_t)n * 4));
    uint32_t* wbits = reinterpret_cast<uint32_t*>(ws + 2 * align256((size_t)n * 4));
    int64_t* bcounts = reinterpret_cast<int64_t*>(ws + 2 * align256((size_t)n * 4) + align256((size_t)nw * 4));
    uint32_t* bits = d_res_bits ? d_res_bits : wbits;
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(d_count), 2, 0u);
        BAMF_LAUNCH_CHECK("k_fill_u32");
        return BAMF_OK;
    }
    const Coeffs c1 = make_coeffs(sigma1), c2 = make_coeffs(sigma2);
    int rc = launch_rows(d_stack, A, Bf, xsize, (int64_t)ysize * nframes, c1, c2, 2, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cols_iir<true>, dim3((unsigned)ceil_div(xsize, 64), (unsigned)nframes), dim3(64), 0, s,
                       A, Bf, xsize, ysize, c1, c2);
    BAMF_LAUNCH_CHECK("k_cols_iir");
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<int64_t>(ceil_div(nw, 256), 4096)), dim3(256), 0, s,
                       bits, nw, 0u);
    BAMF_LAUNCH_CHECK("k_fill_u32");
    if (boxsize <= kMaxK) {
        dim3 grid((unsigned)ceil_div(xsize, kTX), (unsigned)ceil_div(ysize, kTY), (unsigned)nframes);
        hipLaunchKernelGGL(k_localmax, grid, dim3(256), 0, s, A, bits, xsize, ysize, boxsize, minval);
        BAMF_LAUNCH_CHECK("k_localmax");
    } else {
        // wide windows: the reference's two passes, then pack the 0/1 map
        if ((rc = bamf_find_roi_max_x(A, Bf, xsize, ysize, nframes, boxsize, minval, stream))) return rc;
        if ((rc = bamf_find_roi_max_y(A, Bf, xsize, ysize, nframes, boxsize, minval, stream))) return rc;
        hipLaunchKernelGGL(k_pack_bits, dim3((unsigned)ceil_div(nw, 256)), dim3(256), 0, s, A, bits, n);
        BAMF_LAUNCH_CHECK("k_pack_bits");
    }
    hipLaunchKernelGGL(k_bits_count, dim3((unsigned)nb), dim3(kCompactThreads), 0, s, bits, nw, bcounts);
    BAMF_LAUNCH_CHECK("k_bits_count");
    hipLaunchKernelGGL(k_bits_scan, dim3(1), dim3(1024), 0, s, bcounts, nb, d_count);
    BAMF_LAUNCH_CHECK("k_bits_scan");
    hipLaunchKernelGGL(k_bits_emit, dim3((unsigned)nb), dim3(kCompactThreads), 0, s, bits, nw, bcounts, xsize,
                       ysize, boxsize, capacity, d_lin_idx, d_start_row, d_start_col, d_frame);
    BAMF_LAUNCH_CHECK("k_bits_emit");
    return BAMF_OK;
}

int bamf_find_roi_crop(const float* d_stack, int xsize, int ysize, const int32_t* d_start_row,
                       const int32_t* d_start_col, const int32_t* d_frame, int64_t n, int boxsize,
                       float* d_roi_stack, void* stream) {
    BAMF_CHECK_ARG(n >= 0 && boxsize > 0 && boxsize <= xsize && boxsize <= ysize, "bad sizes");
    if (n == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_stack && d_start_row && d_start_col && d_frame && d_roi_stack, "null device pointer");
    const int64_t total = n * boxsize * boxsize;
    hipLaunchKernelGGL(k_crop, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, as_stream(stream), d_stack,
                       xsize, ysize, d_start_row, d_start_col, d_frame, n, boxsize, d_roi_stack);
    BAMF_LAUNCH_CHECK("k_crop");
    return BAMF_OK;
}

}  // extern "C"
