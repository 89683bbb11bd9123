// Library-wide runtime pieces: error strings, version, device queries.
#include <cstdarg>
#include <cstring>

#include "common.hpp"

namespace bamf {
static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace bamf

extern "C" {

const char* bamf_last_error(void) { return bamf::g_err; }

int bamf_version(void) { return BAMF_ABI_VERSION; }

int bamf_device_info(int device, int* cu_count, int* clock_khz, char* name, int name_len) {
    hipDeviceProp_t prop;
    BAMF_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (name && name_len > 0) {
        std::strncpy(name, prop.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    return BAMF_OK;
}

}  // extern "C"
