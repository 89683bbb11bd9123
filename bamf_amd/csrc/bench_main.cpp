// Standalone C++ bench harness over the C ABI (no Python): synthetic ROIs of
// the batched workload from the forward model (pixel-integrated Gaussian
// plus background, Poisson counts), every chain run on the device, timed with
// HIP events around bamf_rjmcmc_run with the inputs resident.
//
//   bamf_bench [--rois N] [--jumps J] [--steps K] [--seed S]
//
// Prints one JSON line: chains/s, jumps/s, ms per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "bamf_gpu.h"

namespace {

#define CHECK_HIP(x)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));           \
            std::exit(1);                                                                \
        }                                                                                \
    } while (0)
#define CHECK_BAMF(x)                                                                    \
    do {                                                                                 \
        int r_ = (x);                                                                    \
        if (r_ != BAMF_OK) {                                                             \
            std::fprintf(stderr, "%s failed: %s\n", #x, bamf_last_error());              \
            std::exit(1);                                                                \
        }                                                                                \
    } while (0)

template <class T>
T* upload(const std::vector<T>& h) {
    T* d = nullptr;
    CHECK_HIP(hipMalloc(&d, h.size() * sizeof(T) + 1));
    if (!h.empty()) CHECK_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// Gamma(k, theta) pdf on 0, dx, 2 dx, ...
std::vector<float> gamma_table(double k, double theta, double dx, int n) {
    std::vector<float> t(n);
    for (int i = 0; i < n; i++) {
        const double x = i * dx;
        t[i] = x <= 0 ? 0.f : (float)std::exp((k - 1) * std::log(x) - x / theta - std::lgamma(k) - k * std::log(theta));
    }
    return t;
}

}  // namespace

int main(int argc, char** argv) {
    int n_roi = 20000, jumps = 10000, steps = 3;
    uint64_t seed = 20261203;
    for (int i = 1; i + 1 < argc; i += 2) {
        const std::string a = argv[i];
        if (a == "--rois") n_roi = std::atoi(argv[i + 1]);
        else if (a == "--jumps") jumps = std::atoi(argv[i + 1]);
        else if (a == "--steps") steps = std::atoi(argv[i + 1]);
        else if (a == "--seed") seed = std::strtoull(argv[i + 1], nullptr, 10);
    }
    // synthetic ROIs: side from {8,12,16,24,32}, density U[0.5,5] /um^2 at 100 nm
    // pixels, I ~ Gamma(5, 200), bg ~ U[5,30], sigma 1.3; prior density 2.75 /um^2
    std::mt19937_64 rng(seed);
    const int sides[5] = {8, 12, 16, 24, 32};
    std::vector<int32_t> xs(n_roi), ys(n_roi);
    std::vector<int64_t> off(n_roi + 1, 0);
    std::vector<float> data;
    const double c = 1.0 / (1.3 * std::sqrt(2.0));
    for (int r = 0; r < n_roi; r++) {
        const int s = sides[std::uniform_int_distribution<int>(0, 4)(rng)];
        xs[r] = ys[r] = s;
        const double dens = std::uniform_real_distribution<double>(0.5, 5.0)(rng);
        const int k = std::poisson_distribution<int>(dens * s * s * 0.01)(rng);
        const double bg = std::uniform_real_distribution<double>(5.0, 30.0)(rng);
        std::vector<double> mu((size_t)s * s, bg);
        for (int e = 0; e < k; e++) {
            const double X = std::uniform_real_distribution<double>(0, s)(rng);
            const double Y = std::uniform_real_distribution<double>(0, s)(rng);
            const double I = std::gamma_distribution<double>(5.0, 200.0)(rng);
            for (int x = 0; x < s; x++) {
                const double gx = 0.5 * (std::erf((x + 1 - X) * c) - std::erf((x - X) * c));
                for (int y = 0; y < s; y++)
                    mu[(size_t)x * s + y] += I * gx * 0.5 * (std::erf((y + 1 - Y) * c) - std::erf((y - Y) * c));
            }
        }
        for (double m : mu) data.push_back((float)std::poisson_distribution<int>(m)(rng));
        off[r + 1] = off[r] + (int64_t)s * s;
    }
    const std::vector<float> sig = gamma_table(5, 200, 10.0, 1001), offp = gamma_table(4, 5, 10.0, 1001);
    const std::vector<float> back(1001, 0.0f);
    bamf_mcmc_params p;
    std::memset(&p, 0, sizeof(p));
    p.psf_sigma = 1.3f;
    p.n_burnin = jumps / 2;
    p.n_trials = jumps - jumps / 2;
    p.i_std_fg = 10.f;
    p.i_std_bg = 5.f;
    p.x_std_fg = 0.1f;
    p.x_std_bg = 0.05f;
    const float pb[8] = {0.4f, 0.15f, 0.15f, 0.15f, 0.15f, 0, 0, 0}, pt[8] = {0.6f, 0.1f, 0.1f, 0.1f, 0.1f, 0, 0, 0};
    std::memcpy(p.p_burnin, pb, sizeof(pb));
    std::memcpy(p.p_trials, pt, sizeof(pt));
    p.rho = 0.0275f;
    p.bnd_out = 2.f;
    p.grid_zoom = 20;
    p.split_std = 0.5f;
    p.bg_std = 1.f;
    p.abbg_std = 0.02f;
    p.dx = 10.f;
    bamf_rjmcmc_plan* plan = nullptr;
    CHECK_BAMF(bamf_rjmcmc_plan_create(&p, offp.data(), (int)offp.size(), sig.data(), (int)sig.size(), back.data(),
                                       (int)back.size(), &plan));
    bamf_roi_batch b;
    std::memset(&b, 0, sizeof(b));
    b.n_roi = n_roi;
    b.d_pix_offset = upload(off);
    b.d_xsize = upload(xs);
    b.d_ysize = upload(ys);
    b.d_data = upload(data);
    // initial BG at the offset prior's mode (rjmcmc.m:47)
    int mode = 0;
    for (int i = 1; i < (int)offp.size(); i++)
        if (offp[i] > offp[mode]) mode = i;
    const std::vector<float> bg0(n_roi, p.dx * mode);
    b.d_bg0 = upload(bg0);
    // longest-first queue order
    std::vector<int32_t> order(n_roi);
    for (int r = 0; r < n_roi; r++) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return xs[a] * ys[a] > xs[c] * ys[c]; });
    b.d_order = upload(order);
    bamf_rjmcmc_out o;
    std::memset(&o, 0, sizeof(o));
    std::vector<int32_t> st(n_roi);
    o.d_status = upload(st);
    std::vector<int32_t> hist((size_t)n_roi * 64);
    o.d_n_hist = upload(hist);
    o.hist_len = 64;
    hipStream_t s;
    CHECK_HIP(hipStreamCreate(&s));
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    CHECK_BAMF(bamf_rjmcmc_run(plan, &b, seed + 99, BAMF_RJ_PLAIN, nullptr, &o, s));  // warm-up
    CHECK_HIP(hipStreamSynchronize(s));
    CHECK_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < steps; k++) CHECK_BAMF(bamf_rjmcmc_run(plan, &b, seed + k, BAMF_RJ_PLAIN, nullptr, &o, s));
    CHECK_HIP(hipEventRecord(e1, s));
    CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    ms /= steps;
    CHECK_HIP(hipMemcpy(st.data(), o.d_status, st.size() * 4, hipMemcpyDeviceToHost));
    int bad = 0;
    for (int v : st) bad += v != 0;
    std::printf("{\"harness\": \"bamf_bench\", \"n_roi\": %d, \"jumps_per_chain\": %d, \"steps\": %d, "
                "\"ms_per_step\": %.3f, \"chains_per_s\": %.1f, \"jumps_per_s\": %.4g, \"status_nonzero\": %d}\n",
                n_roi, jumps, steps, ms, n_roi / (ms / 1e3), (double)n_roi * jumps / (ms / 1e3), bad);
    bamf_rjmcmc_plan_destroy(plan);
    return bad ? 1 : 0;
}
