// Evaluation hook of the FP64 device math routines (devmath.h), used by the
// accuracy tests: out[i] = f(a[i], b[i]) computed on the GPU.
#include "devmath.h"
#include "hmx_internal.h"

namespace hmx {
namespace {

__global__ __launch_bounds__(256) void devmath_eval_kernel(int fn, const double *__restrict__ a,
                                                           const double *__restrict__ b,
                                                           double *__restrict__ out, int64_t n) {
    __shared__ dm::Tables tables;
    dm::load_tables(&tables);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = 0.0;
    switch (fn) {
        case HMX_DM_SQRT: r = dm::sqrt_nn(a[i]); break;
        case HMX_DM_DIV: r = dm::div(a[i], b[i]); break;
        case HMX_DM_LOG: r = dm::log_pos(a[i], tables); break;
        case HMX_DM_ATAN_RATIO: r = dm::atan_ratio(a[i], b[i], tables); break;
        case HMX_DM_ATAN2: r = dm::atan2_full(a[i], b[i], tables); break;
        case HMX_DM_RSQRT: r = dm::rsqrt_pos(a[i]); break;
    }
    out[i] = r;
}

}  // namespace
}  // namespace hmx

extern "C" int hmx_devmath_eval(int fn, const double *a, const double *b, double *out, int64_t n,
                                void *stream) {
    using namespace hmx;
    clear_error();
    if (fn < HMX_DM_SQRT || fn > HMX_DM_RSQRT) return fail(HMX_ERR_ARG, "unknown function %d", fn);
    if (n < 0) return fail(HMX_ERR_ARG, "negative size");
    if (n == 0) return HMX_OK;
    if (!a || !out || ((fn == HMX_DM_DIV || fn == HMX_DM_ATAN_RATIO || fn == HMX_DM_ATAN2) && !b))
        return fail(HMX_ERR_ARG, "NULL array");
    hipLaunchKernelGGL(devmath_eval_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, fn, a, b, out, n);
    HMX_LAUNCHED("devmath_eval_kernel");
    return HMX_OK;
}
