// Error reporting, version and device queries of libhmx.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "hmx_internal.h"

namespace {
thread_local char g_last_error[1024] = "";

using hmx::kMaxDevices;
int g_cu_count[kMaxDevices] = {0};
std::mutex g_mutex;
}  // namespace

namespace hmx {

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_last_error[0] = '\0'; }

int device_cu_count() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(HMX_ERR_NODEV, "no HIP device: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(HMX_ERR_NODEV, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_cu_count[dev] == 0) {
        int n = 0;
        e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess || n <= 0)
            return fail(HMX_ERR_NODEV, "cannot query device %d: %s", dev, hipGetErrorString(e));
        g_cu_count[dev] = n;
    }
    return g_cu_count[dev];
}

}  // namespace hmx

extern "C" {

const char *hmx_last_error(void) { return g_last_error; }

int hmx_version(void) { return 10000; }

int hmx_device_count(int *count) {
    if (!count) return hmx::fail(HMX_ERR_ARG, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    int usable = 0;
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) ++usable;
    }
    *count = usable;
    return HMX_OK;
}

int hmx_forward_workspace_bytes(int64_t n_obs, int64_t n_sources, uint32_t field_mask,
                                size_t *bytes) {
    if (!bytes) return hmx::fail(HMX_ERR_ARG, "bytes is NULL");
    if (n_obs < 0 || n_sources < 0) return hmx::fail(HMX_ERR_ARG, "negative size");
    if (n_obs > INT32_MAX || n_sources > INT32_MAX)
        return hmx::fail(HMX_ERR_ARG, "more than 2^31-1 observation points or sources");
    int nf = hmx::popcount10(field_mask);
    if (nf == 0 || (field_mask & ~HMX_ALL_FIELDS))
        return hmx::fail(HMX_ERR_ARG, "invalid field mask 0x%x", field_mask);
    int cu = hmx::device_cu_count();
    if (cu < 0) return cu;
    int64_t chunks = hmx::forward_chunks(n_obs, n_sources, cu);
    *bytes = (size_t)chunks * (size_t)nf * (size_t)n_obs * sizeof(double);
    // the prism paths' point order, after the partial sums
    if (hmx::use_point_order(n_obs, n_sources))
        *bytes = hmx::point_order_offset(*bytes) + hmx::point_order_bytes(n_obs);
    return HMX_OK;
}

}  // extern "C"
