// Shared helpers for the HIP sources of libbamf_gpu.so (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/bamf_gpu.h"

namespace bamf {

void set_error(const char* fmt, ...);

#define BAMF_CHECK_ARG(cond, ...)                 \
    do {                                          \
        if (!(cond)) {                            \
            ::bamf::set_error(__VA_ARGS__);       \
            return BAMF_E_INVALID_ARG;            \
        }                                         \
    } while (0)

#define BAMF_HIP_TRY(expr)                                                          \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::bamf::set_error("%s failed: %s", #expr, hipGetErrorString(e_));       \
            return BAMF_E_HIP;                                                      \
        }                                                                           \
    } while (0)

#define BAMF_LAUNCH_CHECK(name)                                                     \
    do {                                                                            \
        hipError_t e_ = hipGetLastError();                                          \
        if (e_ != hipSuccess) {                                                     \
            ::bamf::set_error("launch of %s failed: %s", name, hipGetErrorString(e_)); \
            return BAMF_E_HIP;                                                      \
        }                                                                           \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace bamf
