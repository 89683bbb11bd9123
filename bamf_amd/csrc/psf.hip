// Sampled-PSF construction of @RJ/findPSF.m:38-68 on the device: each ROI's
// centred spectrum fftshift(fft2(fftshift(R / 10))) times the phase ramp that
// moves its fitted emitter to the common position, summed over the first
// min(N, 1000) ROIs, then |fftshift(ifft2(.))| of that spectrum zero-padded by
// 12 per side.
//
// gfx950 design: the transforms are B x B (B = 8) and (B + 24)^2, far too
// small for an FFT library call per ROI to pay: one work-group per ROI
// evaluates its B^2 frequencies as direct sums in double, the twiddles an
// exact integer phase (p r + q c) mod B into an LDS table, and writes them;
// one work-group then sums the spectra in ROI order (deterministic) and
// evaluates the padded inverse, whose only nonzero inputs are the B^2
// spectrum entries (B^2 terms per output pixel).
//
// Index conventions (MATLAB column-major, R(r, c) at r + B c, rows = y):
// with p = u - B/2, q = v - B/2 the centred frequency of spectrum entry
// (u, v), fftshift-fft2-fftshift of an even-size R is
//   F(u, v) = (-1)^(p + q) sum_{r,c} R(r, c) exp(-2 pi i (p r + q c) / B),
// and findPSF.m:45 multiplies it by exp(2 pi i q (X - B/2 - 1/2) / B)
// exp(2 pi i p (Y - B/2 - 1/2) / B) (its XGm = q varies along columns).
#include <cmath>

#include "common.hpp"

namespace bamf {
namespace {

constexpr int kPsfBox = 8, kPsfPad = 12, kPsfSide = kPsfBox + 2 * kPsfPad, kPsfMaxRois = 1000;
constexpr double kTwoPi = 6.283185307179586;

__global__ __launch_bounds__(kPsfBox * kPsfBox) void k_psf_spectra(const float* __restrict__ rois,
                                                                   const float* __restrict__ x,
                                                                   const float* __restrict__ y, int n,
                                                                   double2* __restrict__ spec) {
    constexpr int B = kPsfBox, BB = B * B;
    __shared__ double sr[BB];
    __shared__ double tc[B], ts[B];
    const int t = threadIdx.x, roi = blockIdx.x;
    sr[t] = (double)rois[(int64_t)roi * BB + t] / 10.0;
    if (t < B) {
        double s, c;
        sincos(kTwoPi * t / B, &s, &c);
        tc[t] = c;
        ts[t] = s;
    }
    __syncthreads();
    const int u = t % B, v = t / B;  // column-major spectrum entry (u, v)
    const int p = u - B / 2, q = v - B / 2;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < B; c++)
        for (int r = 0; r < B; r++) {
            const int k = ((p * r + q * c) % B + B) % B;  // exp(-2 pi i k / B)
            const double a = sr[r + B * c];
            re += a * tc[k];
            im -= a * ts[k];
        }
    if ((p + q) & 1) {
        re = -re;
        im = -im;
    }
    const double dx = (double)x[roi] - B / 2.0 - 0.5, dy = (double)y[roi] - B / 2.0 - 0.5;
    double s, c;
    sincos(kTwoPi * (q * dx + p * dy) / B, &s, &c);
    spec[(int64_t)roi * BB + t] = make_double2(re * c - im * s, re * s + im * c);
}

// One work-group: the ordered sum of the spectra into LDS, then pixel (a, b)
// of |fftshift(ifft2(pad(F, 12)))| for every pixel of the padded square.
__global__ __launch_bounds__(1024) void k_psf_image(const double2* __restrict__ spec, int n,
                                                    float* __restrict__ image) {
    constexpr int B = kPsfBox, BB = B * B, M = kPsfSide;
    __shared__ double2 fm[BB];
    __shared__ double tc[M], ts[M];
    const int t = threadIdx.x;
    if (t < BB) {
        double re = 0.0, im = 0.0;
        for (int k = 0; k < n; k++) {
            const double2 v = spec[(int64_t)k * BB + t];
            re += v.x;
            im += v.y;
        }
        fm[t] = make_double2(re, im);
    }
    if (t < M) {
        double s, c;
        sincos(kTwoPi * t / M, &s, &c);
        tc[t] = c;
        ts[t] = s;
    }
    __syncthreads();
    for (int o = t; o < M * M; o += blockDim.x) {
        const int a2 = o % M, b2 = o / M;  // output (row, column) after fftshift
        const int a = (a2 + M / 2) % M, b = (b2 + M / 2) % M;
        double re = 0.0, im = 0.0;
        for (int v = 0; v < B; v++)
            for (int u = 0; u < B; u++) {
                const int k = ((u + kPsfPad) * a + (v + kPsfPad) * b) % M;  // exp(+2 pi i k / M)
                const double2 f = fm[u + B * v];
                re += f.x * tc[k] - f.y * ts[k];
                im += f.x * ts[k] + f.y * tc[k];
            }
        image[o] = (float)(sqrt(re * re + im * im) / (double)(M * M));
    }
}

}  // namespace
}  // namespace bamf

using namespace bamf;

extern "C" {

int bamf_psf_shift_average(const float* d_rois, const float* d_x, const float* d_y, int64_t n, int box,
                           void* d_workspace, float* d_image, void* stream) {
    BAMF_CHECK_ARG(box == kPsfBox, "findPSF's fixed 64 x 64 x 4 x 4 split needs BoxSize 8 (got %d)", box);
    BAMF_CHECK_ARG(n >= 1, "need at least one ROI");
    BAMF_CHECK_ARG(d_rois && d_x && d_y && d_workspace && d_image, "null device pointer");
    const int used = (int)(n < kPsfMaxRois ? n : kPsfMaxRois);  // findPSF.m:41
    double2* spec = static_cast<double2*>(d_workspace);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_psf_spectra, dim3(used), dim3(kPsfBox * kPsfBox), 0, s, d_rois, d_x, d_y, used, spec);
    BAMF_LAUNCH_CHECK("k_psf_spectra");
    hipLaunchKernelGGL(k_psf_image, dim3(1), dim3(1024), 0, s, spec, used, d_image);
    BAMF_LAUNCH_CHECK("k_psf_image");
    return BAMF_OK;
}

int64_t bamf_psf_workspace_bytes(int64_t n) {
    return (int64_t)sizeof(double2) * kPsfBox * kPsfBox * (n < kPsfMaxRois ? n : kPsfMaxRois);
}

}  // extern "C"
