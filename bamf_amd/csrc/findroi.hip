// Difference-of-Gaussian ROI finder (source/cuda/cuda_FindROI/cuda_FindROI.cu
// and the MATLAB driver @RJ/findROI.m:60-225), resident on the device.
//
// Bit-exactness: each output value is computed with the reference's
// expression and operation order, in single precision, with FMA contraction
// off and IEEE division (HIP's default), so results match the reference's
// semantics bit for bit regardless of how the work is tiled.
//
// gfx950 design
//   rows pass   k_rows_dog   : 64 rows per work-group staged through LDS in
//                              32-sample chunks (coalesced row segments), the
//                              recursion for both sigmas per row; the forward
//                              pass keeps only chunk-start states and the
//                              backward pass recomputes each chunk's forward
//                              outputs in registers (no HBM round trip);
//   column pass k_cols_dog   : one work-item per column, adjacent lanes on
//                              adjacent columns (coalesced), chunk-start states
//                              in LDS, forward outputs recomputed per chunk in
//                              registers, backward pass fused with the DoG
//                              subtract;
//   local max   k_localmax<K>: 64x64 output tile, D staged in LDS with a K
//                              halo and overwritten by the row maxima; both
//                              window maxima by doubling in registers; bits set
//                              with sparse atomicOr into a Res bitmap;
//   compaction  k_bits_*     : ordered stream compaction of the bitmap
//                              (per-block popcount, scan, ordered emit) plus the
//                              box corners of findROI.m:130-148.
#include <cmath>
#include <mutex>
#include <type_traits>

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

// Row recursion (the reference's gaussX) for NS sigmas at once, writing the
// forward outputs and re-reading them.  in may alias out0 (the per-kernel
// entry point runs it in place with NS = 1).
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

// Column recursion (the reference's gaussY) in place on a, one work-item per
// column, adjacent lanes on adjacent columns.
__global__ __launch_bounds__(64) void k_cols_iir(float* a, int xsize, int ysize, Coeffs c0) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int64_t z = blockIdx.y;
    if (x >= xsize) return;
    const int64_t st = xsize;
    float* pa = a + z * xsize * ysize + x;
    // forward (cuda_FindROI.cu:60-70)
    float a1, a2, a3;
    a1 = a2 = a3 = pa[0];
    for (int ii = 0; ii < ysize; ii++) {
        const float va = iir_step(pa[ii * st], a1, a2, a3, c0);
        pa[ii * st] = va;
        a3 = a2; a2 = a1; a1 = va;
    }
    // backward (cuda_FindROI.cu:72-82)
    a1 = a2 = a3 = pa[(int64_t)(ysize - 1) * st];
    for (int ii = ysize - 1; ii >= 0; ii--) {
        const float va = iir_step(pa[ii * st], a1, a2, a3, c0);
        a3 = a2; a2 = a1; a1 = va;
        pa[ii * st] = va;
    }
}

// Row recursion for both sigmas of the DoG pipeline (in must not alias the
// outputs).  The forward pass keeps only the recursion state: the three last
// forward outputs of each chunk, which are the next chunk's starting state, are
// parked in the checkpoint array ck, laid out [chunk][value][row] so that a
// wave's 64 rows store and load each value as one 256-byte run (parking them
// inside the outputs would make partial-line stores).  The backward pass walks
// the chunks from the end, re-reads a
// chunk, recomputes its forward outputs in registers from the parked state (the
// same operations on the same inputs, hence the same values) and runs the
// backward recursion on them.  HBM traffic: the stack read twice and each
// output written once, where writing and re-reading the forward outputs costs
// four image passes more.
// FULL: every block has kRowsPerBlock rows and xsize is a multiple of kChunk
// (no guards, 16-byte row transfers); otherwise guarded scalar transfers.
// Rows row_begin + blockIdx.x * kRowsPerBlock ... of nrows.
template <bool FULL>
__global__ __launch_bounds__(kRowsPerBlock) void k_rows_dog(const float* __restrict__ in, float* __restrict__ out0,
                                                           float* __restrict__ out1, float* __restrict__ ck,
                                                           int xsize, int64_t nrows, int64_t row_begin, Coeffs c0,
                                                           Coeffs c1) {
    __shared__ float tA[kRowsPerBlock][kChunk + 1];
    const int t = threadIdx.x;
    const int64_t row0 = row_begin + (int64_t)blockIdx.x * kRowsPerBlock;
    const int nr = FULL ? kRowsPerBlock : (int)min<int64_t>(kRowsPerBlock, nrows - row0);
    const int nchunk = (xsize + kChunk - 1) / kChunk;
    // Chunk transfers.  FULL: element i of work-item t is the float4 at row
    // kRowStep4 * i + t / 8, columns 4 (t % 8) ..+3.  Otherwise element i is row
    // kRowStep * i + t / kChunk, column t % kChunk, guarded.  Both address by a
    // per-lane pointer stepped by whole rows, and the next chunk is fetched
    // into registers while the current one is processed.
    constexpr int kRowStep = kRowsPerBlock / kChunk, kPer = kChunk;
    constexpr int kQuads = kChunk / 4, kRowStep4 = kRowsPerBlock / kQuads, kPer4 = kChunk / 4;
    const int tr = t / kChunk, tc = t % kChunk;
    const int tr4 = t / kQuads, tc4 = (t % kQuads) * 4;
    const int64_t bstep = (int64_t)kRowStep * xsize, bstep4 = (int64_t)kRowStep4 * xsize;
    const int64_t loff = FULL ? (row0 + tr4) * xsize + tc4 : (row0 + tr) * xsize + tc;
    float pre[kPer];
    auto fetch_any = [&](int x0, int nc) {
        const float* p = in + loff + x0;
        if (FULL) {
#pragma unroll
            for (int i = 0; i < kPer4; i++) {
                const float4 v = *reinterpret_cast<const float4*>(p);
                pre[4 * i] = v.x; pre[4 * i + 1] = v.y; pre[4 * i + 2] = v.z; pre[4 * i + 3] = v.w;
                p += bstep4;
            }
        } else {
#pragma unroll
            for (int i = 0; i < kPer; i++) {
                if (kRowStep * i + tr < nr && tc < nc) pre[i] = *p;
                p += bstep;
            }
        }
    };
    auto put_any = [&](int nc) {
        if (FULL) {
#pragma unroll
            for (int i = 0; i < kPer4; i++)
#pragma unroll
                for (int q = 0; q < 4; q++) tA[kRowStep4 * i + tr4][tc4 + q] = pre[4 * i + q];
        } else {
#pragma unroll
            for (int i = 0; i < kPer; i++)
                if (kRowStep * i + tr < nr && tc < nc) tA[kRowStep * i + tr][tc] = pre[i];
        }
    };
    auto store_any = [&](float* out, int x0, int nc) {
        float* p = out + loff + x0;
        if (FULL) {
#pragma unroll
            for (int i = 0; i < kPer4; i++) {
                const int r = kRowStep4 * i + tr4;
                *reinterpret_cast<float4*>(p) = make_float4(tA[r][tc4], tA[r][tc4 + 1], tA[r][tc4 + 2], tA[r][tc4 + 3]);
                p += bstep4;
            }
        } else {
#pragma unroll
            for (int i = 0; i < kPer; i++) {
                if (kRowStep * i + tr < nr && tc < nc) *p = tA[kRowStep * i + tr][tc];
                p += bstep;
            }
        }
    };
    float a1 = 0.f, a2 = 0.f, a3 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
    auto forward = [&](auto full, int nc) {
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            if (decltype(full)::value || k < nc) {
                const float w0 = tA[t][k];
                const float va = iir_step(w0, a1, a2, a3, c0);
                a3 = a2; a2 = a1; a1 = va;
                const float vb = iir_step(w0, b1, b2, b3, c1);
                b3 = b2; b2 = b1; b1 = vb;
            }
        }
    };
    // backward state, and the chunk's forward outputs for sigma2 (then its results)
    float p1 = 0.f, p2 = 0.f, p3 = 0.f, q1 = 0.f, q2 = 0.f, q3 = 0.f;
    float fb[kChunk];
    auto backward = [&](auto full, int nc) {
        float fa[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            if (decltype(full)::value || k < nc) {
                const float w0 = tA[t][k];
                fa[k] = iir_step(w0, a1, a2, a3, c0);
                a3 = a2; a2 = a1; a1 = fa[k];
                fb[k] = iir_step(w0, b1, b2, b3, c1);
                b3 = b2; b2 = b1; b1 = fb[k];
            }
        }
#pragma unroll
        for (int k = kChunk - 1; k >= 0; k--) {
            if (decltype(full)::value || k < nc) {
                const float va = iir_step(fa[k], p1, p2, p3, c0);
                p3 = p2; p2 = p1; p1 = va;
                tA[t][k] = va;
                const float vb = iir_step(fb[k], q1, q2, q3, c1);
                q3 = q2; q2 = q1; q1 = vb;
                fb[k] = vb;
            }
        }
    };
    // starting state of the chunk the backward pass works on next
    float s1 = 0.f, s2 = 0.f, s3 = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f;
    // ---- forward (cuda_FindROI.cu:23-33), states only ----
    fetch_any(0, min(kChunk, xsize));
    for (int ch = 0; ch < nchunk; ch++) {
        const int x0 = ch * kChunk, nc = min(kChunk, xsize - x0);
        put_any(nc);
        __syncthreads();
        if (ch + 1 < nchunk) fetch_any(x0 + kChunk, min(kChunk, xsize - x0 - kChunk));
        if (t < nr) {
            if (ch == 0) {
                a1 = a2 = a3 = tA[t][0];
                b1 = b2 = b3 = a1;
            }
            s1 = a1; s2 = a2; s3 = a3; u1 = b1; u2 = b2; u3 = b3;
            if (FULL || nc == kChunk) forward(std::true_type{}, nc);
            else forward(std::false_type{}, nc);
            if (ch + 2 < nchunk) {  // park the state of chunk ch + 1: [chunk][value][row], coalesced
                float* q = ck + (int64_t)(ch + 1) * 6 * nrows + (row0 + t);
                q[0] = a3; q[nrows] = a2; q[2 * nrows] = a1;
                q[3 * nrows] = b3; q[4 * nrows] = b2; q[5 * nrows] = b1;
            }
        }
        __syncthreads();
    }
    // ---- backward (cuda_FindROI.cu:35-45), from the last forward outputs;
    // tA holds the last chunk's input and s/u its starting state ----
    p1 = p2 = p3 = a1;
    q1 = q2 = q3 = b1;
    for (int ch = nchunk - 1; ch >= 0; ch--) {
        const int x0 = ch * kChunk, nc = min(kChunk, xsize - x0);
        if (ch != nchunk - 1) {
            put_any(nc);
            __syncthreads();
        }
        a1 = s1; a2 = s2; a3 = s3; b1 = u1; b2 = u2; b3 = u3;
        // the next chunk's input and starting state travel together, so waiting
        // for one never waits on the other's later loads
        if (ch > 0) {
            fetch_any(x0 - kChunk, kChunk);
            if (t < nr && ch > 1) {
                const float* q = ck + (int64_t)(ch - 1) * 6 * nrows + (row0 + t);
                s3 = q[0]; s2 = q[nrows]; s1 = q[2 * nrows];
                u3 = q[3 * nrows]; u2 = q[4 * nrows]; u1 = q[5 * nrows];
            }
        }
        if (t < nr) {
            if (ch == 0) {
                a1 = a2 = a3 = tA[t][0];
                b1 = b2 = b3 = a1;
            }
            if (FULL || nc == kChunk) backward(std::true_type{}, nc);
            else backward(std::false_type{}, nc);
        }
        __syncthreads();
        store_any(out0, x0, nc);
        __syncthreads();
        if (t < nr) {
#pragma unroll
            for (int k = 0; k < kChunk; k++)
                if (FULL || k < nc) tA[t][k] = fb[k];
        }
        __syncthreads();
        store_any(out1, x0, nc);
        __syncthreads();
    }
}

// Column recursion for both sigmas fused with the DoG subtract (cuda_FindROI.cu:
// 56-91 and kernel_subtract): a receives G(sigma1) - G(sigma2), b is only read.
// The forward pass keeps only the recursion state at the start of each chunk
// of kColChunk rows, in ck (6 floats per column per chunk, lanes on adjacent
// columns); the backward pass walks the chunks from the end, re-reads a chunk
// of both images, recomputes its forward outputs in registers and runs the
// backward recursion on them.  A work-group spans 256 adjacent columns, so each
// row access covers 1 KiB of contiguous bytes.
constexpr int kColChunk = 32, kColThreads = 256;
__global__ __launch_bounds__(kColThreads) void k_cols_dog(float* a, const float* b, float* __restrict__ ck,
                                                          int xsize, int ysize, Coeffs c0, Coeffs c1) {
    const int x = blockIdx.x * kColThreads + threadIdx.x;
    if (x >= xsize) return;
    const int nchunk = (ysize + kColChunk - 1) / kColChunk;
    const int64_t st = xsize;
    float* const ac = a + (int64_t)blockIdx.y * xsize * ysize + x;  // this column
    const float* const bc = b + (int64_t)blockIdx.y * xsize * ysize + x;
    float* const cc = ck + (int64_t)blockIdx.y * nchunk * 6 * xsize + x;
    // Every helper has a full form (no guards, kColChunk rows) and a guarded
    // one for the ragged last chunk.
    float wa[kColChunk], wb[kColChunk];
    auto load = [&](auto full, int y0, int n) {  // every load is issued before the first use
        const float* pa = ac + y0 * st;
        const float* pb = bc + y0 * st;
#pragma unroll
        for (int k = 0; k < kColChunk; k++) {
            if (decltype(full)::value || k < n) {
                wa[k] = *pa;
                wb[k] = *pb;
            }
            pa += st;
            pb += st;
        }
    };
    float a1, a2, a3, b1, b2, b3;
    auto forward = [&](auto full, int n) {
#pragma unroll
        for (int k = 0; k < kColChunk; k++)
            if (decltype(full)::value || k < n) {
                const float va = iir_step(wa[k], a1, a2, a3, c0);
                a3 = a2; a2 = a1; a1 = va;
                const float vb = iir_step(wb[k], b1, b2, b3, c1);
                b3 = b2; b2 = b1; b1 = vb;
            }
    };
    float p1, p2, p3, q1, q2, q3;
    auto backward = [&](auto full, int y0, int n) {
#pragma unroll
        for (int k = 0; k < kColChunk; k++)
            if (decltype(full)::value || k < n) {
                wa[k] = iir_step(wa[k], a1, a2, a3, c0);
                a3 = a2; a2 = a1; a1 = wa[k];
                wb[k] = iir_step(wb[k], b1, b2, b3, c1);
                b3 = b2; b2 = b1; b1 = wb[k];
            }
        float* pa = ac + (y0 + kColChunk - 1) * st;
#pragma unroll
        for (int k = kColChunk - 1; k >= 0; k--) {
            if (decltype(full)::value || k < n) {
                const float va = iir_step(wa[k], p1, p2, p3, c0);
                p3 = p2; p2 = p1; p1 = va;
                const float vb = iir_step(wb[k], q1, q2, q3, c1);
                q3 = q2; q2 = q1; q1 = vb;
                *pa = va - vb;  // cuda_FindROI.cu:91
            }
            pa -= st;
        }
    };
    a1 = a2 = a3 = ac[0];
    b1 = b2 = b3 = bc[0];
    for (int ch = 0; ch < nchunk; ch++) {
        const int y0 = ch * kColChunk, n = min(kColChunk, ysize - y0);
        if (ch > 0) {  // chunk 0 starts from its first row
            float* c = cc + (int64_t)ch * 6 * st;
            c[0] = a1; c[st] = a2; c[2 * st] = a3; c[3 * st] = b1; c[4 * st] = b2; c[5 * st] = b3;
        }
        if (n == kColChunk) {
            load(std::true_type{}, y0, n);
            forward(std::true_type{}, n);
        } else {
            load(std::false_type{}, y0, n);
            forward(std::false_type{}, n);
        }
    }
    // wa/wb hold the last chunk; the backward pass starts at the last forward outputs
    p1 = p2 = p3 = a1;
    q1 = q2 = q3 = b1;
    for (int ch = nchunk - 1; ch >= 0; ch--) {
        const int y0 = ch * kColChunk, n = min(kColChunk, ysize - y0);
        if (ch > 0) {
            const float* c = cc + (int64_t)ch * 6 * st;
            a1 = c[0]; a2 = c[st]; a3 = c[2 * st]; b1 = c[3 * st]; b2 = c[4 * st]; b3 = c[5 * st];
        } else {
            a1 = a2 = a3 = ac[0];
            b1 = b2 = b3 = bc[0];
        }
        if (n == kColChunk) {
            if (ch != nchunk - 1) load(std::true_type{}, y0, n);
            backward(std::true_type{}, y0, n);
        } else {
            if (ch != nchunk - 1) load(std::false_type{}, y0, n);
            backward(std::false_type{}, y0, n);
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

// Fused maxX + maxY on the DoG image, writing Res bits, for window half-width
// K <= kMaxK (one instantiation per K).  A 64 x 64 output tile: the D tile with
// its K halo is staged in LDS (-inf outside the frame, which never changes
// fmaxf(maxval, .) and so reproduces the reference's clipped windows), the row
// maxima (maxX) overwrite it in place, and both window maxima are taken in
// registers by doubling.  fmaxf is exact, so any grouping of a window's
// maximum gives the reference's value bit for bit.
constexpr int kLmT = 64, kMaxK = 16;

// out[j] = max(v[j .. j+2K]) for j < N.  After the pass of width w, v[i] holds
// the maximum over [i, i+2w); the last step joins two overlapping windows of
// length P (the largest power of two <= 2K+1).
template <int N, int K>
__device__ __forceinline__ void window_max(float (&v)[N + 2 * K], float (&out)[N]) {
    constexpr int L = 2 * K + 1;
    constexpr int P = L >= 32 ? 32 : L >= 16 ? 16 : L >= 8 ? 8 : L >= 4 ? 4 : L >= 2 ? 2 : 1;
#pragma unroll
    for (int w = 1; w < P; w *= 2) {
#pragma unroll
        for (int i = 0; i + 2 * w <= N + 2 * K; i++) v[i] = fmaxf(v[i], v[i + w]);
    }
#pragma unroll
    for (int j = 0; j < N; j++) out[j] = fmaxf(v[j], v[j + L - P]);
}

template <int K>
__global__ __launch_bounds__(256) void k_localmax(const float* __restrict__ dog, uint32_t* __restrict__ bits,
                                                  int xsize, int ysize, float minval) {
    constexpr int R = kLmT + 2 * K;      // staged rows (tile rows y0-K .. y0+63+K)
    constexpr int S = kLmT + 2 * K + 1;  // row stride, odd: lanes on consecutive rows use distinct banks
    constexpr int RS = 32 + 2 * K;       // row strip: 32 outputs and the halo
    constexpr int CS = 16 + 2 * K;       // column strip: 16 outputs and the halo
    __shared__ float sD[R * S];
    const int x0 = blockIdx.x * kLmT, y0 = blockIdx.y * kLmT;
    const int64_t fbase = (int64_t)blockIdx.z * xsize * ysize;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    {
        // all of this wave's loads are issued before the first LDS store
        constexpr int RPW = (R + 3) / 4;
        float va[RPW], vb[RPW];
#pragma unroll
        for (int i = 0; i < RPW; i++) {
            const int r = wv + 4 * i, y = y0 - K + r;
            const bool yin = r < R && y >= 0 && y < ysize;
            const int64_t rb = fbase + (int64_t)y * xsize;
            const int xa = x0 - K + lane, xb = xa + kLmT;
            va[i] = (yin && xa >= 0 && xa < xsize) ? dog[rb + xa] : -INFINITY;
            vb[i] = (yin && lane < 2 * K && xb < xsize) ? dog[rb + xb] : -INFINITY;
        }
#pragma unroll
        for (int i = 0; i < RPW; i++) {
            const int r = wv + 4 * i;
            if (r < R) {
                sD[r * S + lane] = va[i];
                if (lane < 2 * K) sD[r * S + kLmT + lane] = vb[i];
            }
        }
    }
    __syncthreads();
    // maxX (cuda_FindROI.cu:105-116): work-item t < 2R takes row t % R, outputs
    // 32 (t / R) .. +31; the results replace columns 0..63 of the row.
    float mx[32];
    const bool act = t < 2 * R;
    const int seg = t / R, rr = t - seg * R, cs = seg * 32;
    if (act) {
        float v[RS], w[32], cv[32];
#pragma unroll
        for (int q = 0; q < RS; q++) v[q] = sD[rr * S + cs + q];
#pragma unroll
        for (int j = 0; j < 32; j++) cv[j] = v[j + K];
        window_max<32, K>(v, w);
#pragma unroll
        for (int j = 0; j < 32; j++) mx[j] = maxx_value(fmaxf(minval, w[j]), cv[j]);
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int j = 0; j < 32; j++) sD[rr * S + cs + j] = mx[j];
    }
    __syncthreads();
    // maxY (cuda_FindROI.cu:134-146): column lane, rows 16 wv .. +15; rows
    // outside the frame leave the window (-inf).
    const int x = x0 + lane, r0 = wv * 16;
    if (x >= xsize) return;
    float v[CS], w[16], cv[16];
#pragma unroll
    for (int q = 0; q < CS; q++) {
        const int y = y0 - K + r0 + q;
        const float m = sD[(r0 + q) * S + lane];
        v[q] = (y >= 0 && y < ysize) ? fabsf(m) : -INFINITY;
    }
#pragma unroll
    for (int j = 0; j < 16; j++) cv[j] = sD[(r0 + K + j) * S + lane];
    window_max<16, K>(v, w);
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int y = y0 + r0 + j;
        if (y < ysize && maxy_flag(fmaxf(minval, w[j]), cv[j])) {
            const int64_t lin = fbase + (int64_t)y * xsize + x;
            atomicOr(&bits[lin >> 5], 1u << (lin & 31));
        }
    }
}

using LocalMaxKernel = void (*)(const float*, uint32_t*, int, int, float);
constexpr LocalMaxKernel kLocalMax[kMaxK + 1] = {
    k_localmax<0>,  k_localmax<1>,  k_localmax<2>,  k_localmax<3>,  k_localmax<4>,  k_localmax<5>,
    k_localmax<6>,  k_localmax<7>,  k_localmax<8>,  k_localmax<9>,  k_localmax<10>, k_localmax<11>,
    k_localmax<12>, k_localmax<13>, k_localmax<14>, k_localmax<15>, k_localmax<16>};

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

// The row recursion in place on d for one sigma (the reference's gaussX).
int launch_rows_inplace(float* d, int xsize, int64_t nrows, Coeffs c, hipStream_t s) {
    dim3 grid((unsigned)ceil_div(nrows, kRowsPerBlock));
    hipLaunchKernelGGL(k_rows_iir<1>, grid, dim3(kRowsPerBlock), 0, s, d, d, nullptr, xsize, nrows, c, c);
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
    return launch_rows_inplace(d, xsize, (int64_t)ysize * nframes, c, as_stream(stream));
}

int bamf_find_roi_gauss_y(float* d, int xsize, int ysize, int nframes, float b0, float b1, float b2,
                          float b3, float B, void* stream) {
    BAMF_CHECK_ARG(xsize > 0 && ysize > 0 && nframes >= 0, "bad sizes");
    if (nframes == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d, "null device pointer");
    BAMF_CHECK_ARG(nframes <= 65535, "nframes must be <= 65535 per call");
    const Coeffs c = with_rcp(Coeffs{b0, b1, b2, b3, B, 0.0f});
    dim3 grid((unsigned)ceil_div(xsize, 64), (unsigned)nframes);
    hipLaunchKernelGGL(k_cols_iir, grid, dim3(64), 0, as_stream(stream), d, xsize, ysize, c);
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

// Recursion checkpoints, 6 floats each: the column pass's per column per
// kColChunk rows, the row pass's per row per kChunk columns (the passes run
// one after the other and share the region).
static size_t col_ck_bytes(int xsize, int ysize, int nframes) {
    const size_t cols = (size_t)xsize * nframes * ceil_div(ysize, kColChunk) * 6 * sizeof(float);
    const size_t rows = (size_t)ysize * nframes * ceil_div(xsize, kChunk) * 6 * sizeof(float);
    return cols > rows ? cols : rows;
}

size_t bamf_find_roi_workspace_bytes(int xsize, int ysize, int nframes) {
    const int64_t n = (int64_t)xsize * ysize * nframes;
    const int64_t nw = nwords_of(n);
    return align256((size_t)n * 4) * 2 + align256((size_t)nw * 4) + align256((size_t)nblocks_of(nw) * 8) +
           align256(col_ck_bytes(xsize, ysize, nframes));
}

}  // extern "C"

namespace {

// Argument checks shared by both entry points (nothing is launched on failure).
int check_find_roi_args(const float* d_stack, int xsize, int ysize, int nframes, int boxsize, int64_t capacity,
                        const int64_t* d_count) {
    BAMF_CHECK_ARG(xsize > 0 && ysize > 0 && nframes >= 0, "bad sizes");
    BAMF_CHECK_ARG(boxsize >= 0, "boxsize must be >= 0");
    BAMF_CHECK_ARG(capacity >= 0, "capacity must be >= 0");
    BAMF_CHECK_ARG(d_count, "d_count is required");
    BAMF_CHECK_ARG(nframes <= 65535, "nframes must be <= 65535 per call");
    BAMF_CHECK_ARG(d_stack, "null stack");
    return BAMF_OK;
}

// The pipeline of bamf_find_roi on a workspace of at least
// bamf_find_roi_workspace_bytes(xsize, ysize, nframes) bytes.  Every
// workspace byte that is read is written earlier in the same call (the two
// filtered stacks and the checkpoints by the row and column passes, the bitmap
// by k_fill_u32, the block counts by k_bits_count), so the workspace's prior
// contents never matter.
int find_roi_on(const float* d_stack, int xsize, int ysize, int nframes, float sigma1, float sigma2, int boxsize,
                float minval, void* d_workspace, uint32_t* d_res_bits, int64_t* d_lin_idx, int32_t* d_start_row,
                int32_t* d_start_col, int32_t* d_frame, int64_t capacity, int64_t* d_count, void* stream) {
    hipStream_t s = as_stream(stream);
    const int64_t n = (int64_t)xsize * ysize * nframes;
    const int64_t nw = nwords_of(n), nb = nblocks_of(nw);
    char* ws = static_cast<char*>(d_workspace);
    float* A = reinterpret_cast<float*>(ws);
    float* Bf = reinterpret_cast<float*>(ws + align256((size_t)n * 4));
    uint32_t* wbits = reinterpret_cast<uint32_t*>(ws + 2 * align256((size_t)n * 4));
    int64_t* bcounts = reinterpret_cast<int64_t*>(ws + 2 * align256((size_t)n * 4) + align256((size_t)nw * 4));
    float* colck = reinterpret_cast<float*>(ws + 2 * align256((size_t)n * 4) + align256((size_t)nw * 4) +
                                            align256((size_t)nb * 8));
    uint32_t* bits = d_res_bits ? d_res_bits : wbits;
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(d_count), 2, 0u);
        BAMF_LAUNCH_CHECK("k_fill_u32");
        return BAMF_OK;
    }
    const Coeffs c1 = make_coeffs(sigma1), c2 = make_coeffs(sigma2);
    const int64_t nrows = (int64_t)ysize * nframes;
    // whole blocks of a chunk-multiple row length unguarded, the rest guarded
    const int64_t nfull = xsize % kChunk == 0 ? nrows / kRowsPerBlock : 0;
    if (nfull > 0) {
        hipLaunchKernelGGL(k_rows_dog<true>, dim3((unsigned)nfull), dim3(kRowsPerBlock), 0, s, d_stack, A, Bf, colck,
                           xsize, nrows, (int64_t)0, c1, c2);
        BAMF_LAUNCH_CHECK("k_rows_dog");
    }
    if (nrows > nfull * kRowsPerBlock) {
        hipLaunchKernelGGL(k_rows_dog<false>, dim3((unsigned)ceil_div(nrows - nfull * kRowsPerBlock, kRowsPerBlock)),
                           dim3(kRowsPerBlock), 0, s, d_stack, A, Bf, colck, xsize, nrows, nfull * kRowsPerBlock, c1,
                           c2);
        BAMF_LAUNCH_CHECK("k_rows_dog");
    }
    hipLaunchKernelGGL(k_cols_dog, dim3((unsigned)ceil_div(xsize, kColThreads), (unsigned)nframes),
                       dim3(kColThreads), 0, s, A, Bf, colck, xsize, ysize, c1, c2);
    BAMF_LAUNCH_CHECK("k_cols_dog");
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<int64_t>(ceil_div(nw, 256), 4096)), dim3(256), 0, s,
                       bits, nw, 0u);
    BAMF_LAUNCH_CHECK("k_fill_u32");
    if (boxsize <= kMaxK) {
        dim3 grid((unsigned)ceil_div(xsize, kLmT), (unsigned)ceil_div(ysize, kLmT), (unsigned)nframes);
        hipLaunchKernelGGL(kLocalMax[boxsize], grid, dim3(256), 0, s, A, bits, xsize, ysize, minval);
        BAMF_LAUNCH_CHECK("k_localmax");
    } else {
        // wide windows: the reference's two passes, then pack the 0/1 map
        int rc;
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

// Workspace of bamf_find_roi_noalloc: one copy per device, held from the
// library's first use on that device until the process ends.  Host code never
// passes this name to a kernel: on the host it is only a handle for
// hipGetSymbolAddress, which gives each device's own address of its copy.
constexpr size_t kNoAllocScratchBytes = (size_t)BAMF_FIND_ROI_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's scratch copy, looked up once
// per device.
int scratch_on_current_device(void** ptr) {
    int dev = 0;
    BAMF_HIP_TRY(hipGetDevice(&dev));
    BAMF_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d this library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        BAMF_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        BAMF_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return BAMF_OK;
}

}  // namespace

extern "C" {

int bamf_find_roi(const float* d_stack, int xsize, int ysize, int nframes, float sigma1, float sigma2,
                  int boxsize, float minval, void* d_workspace, size_t workspace_bytes, uint32_t* d_res_bits,
                  int64_t* d_lin_idx, int32_t* d_start_row, int32_t* d_start_col, int32_t* d_frame,
                  int64_t capacity, int64_t* d_count, void* stream) {
    int rc = check_find_roi_args(d_stack, xsize, ysize, nframes, boxsize, capacity, d_count);
    if (rc) return rc;
    const size_t need = bamf_find_roi_workspace_bytes(xsize, ysize, nframes);
    BAMF_CHECK_ARG(d_workspace && workspace_bytes >= need, "workspace too small (%zu < %zu)", workspace_bytes, need);
    return find_roi_on(d_stack, xsize, ysize, nframes, sigma1, sigma2, boxsize, minval, d_workspace, d_res_bits,
                       d_lin_idx, d_start_row, d_start_col, d_frame, capacity, d_count, stream);
}

int bamf_find_roi_noalloc(const float* d_stack, int xsize, int ysize, int nframes, float sigma1, float sigma2,
                          int boxsize, float minval, uint32_t* d_res_bits, int64_t* d_lin_idx,
                          int32_t* d_start_row, int32_t* d_start_col, int32_t* d_frame, int64_t capacity,
                          int64_t* d_count, void* stream) {
    int rc = check_find_roi_args(d_stack, xsize, ysize, nframes, boxsize, capacity, d_count);
    if (rc) return rc;
    const size_t need = bamf_find_roi_workspace_bytes(xsize, ysize, nframes);
    BAMF_CHECK_ARG(need <= kNoAllocScratchBytes,
                   "stack too large for bamf_find_roi_noalloc: it needs %zu workspace bytes, the scratch holds %zu "
                   "(use bamf_find_roi with a caller workspace, or fewer frames per call)",
                   need, kNoAllocScratchBytes);
    void* ws = nullptr;
    if ((rc = scratch_on_current_device(&ws))) return rc;
    return find_roi_on(d_stack, xsize, ysize, nframes, sigma1, sigma2, boxsize, minval, ws, d_res_bits, d_lin_idx,
                       d_start_row, d_start_col, d_frame, capacity, d_count, stream);
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
