// Equivalent-source jacobian and predict kernels (Green's function 1 / r).
//
// jacobian: one workgroup row-block x column-tile; consecutive lanes write
// consecutive columns of the row-major (n_data, ld) matrix (coalesced
// stores; the kernel is HBM-write bound).
// predict: 256 observation points per workgroup, sources and coefficients
// read with scalar loads, FP64 accumulation (also for FP32 data).
#include "devmath.h"
#include "hmx_internal.h"

namespace hmx {
namespace {

// 1 / sqrt(s): FP64 through the library's reciprocal square root (about one
// ulp), FP32 in single precision (the reference's float32 jacobian).
__device__ __forceinline__ double inv_sqrt(double s) { return dm::rsqrt_pos(s); }
__device__ __forceinline__ float inv_sqrt(float s) { return 1.0f / sqrtf(s); }

template <class T>
__global__ __launch_bounds__(256) void eqs_jacobian_kernel(
    const T *__restrict__ e, const T *__restrict__ n, const T *__restrict__ u, int64_t n_data,
    const T *__restrict__ pe, const T *__restrict__ pn, const T *__restrict__ pu, int64_t n_src,
    T *__restrict__ jac, int64_t ld) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n_src) return;
    const T sj = pe[j], nj = pn[j], uj = pu[j];
    for (int64_t i = blockIdx.y; i < n_data; i += gridDim.y) {
        const T de = e[i] - sj, dn = n[i] - nj, du = u[i] - uj;
        jac[i * ld + j] = inv_sqrt(de * de + dn * dn + du * du);
    }
}

// Sources are the same for every lane of a wavefront, so they are read
// with scalar loads straight from global memory (the compiler turns the
// wavefront-uniform reads of an unrolled loop into 64-byte scalar loads):
// no LDS staging and no barriers.  Each lane adds its sources in index
// order.
template <class T>
__global__ __launch_bounds__(kObsPerBlock) void eqs_predict_kernel(
    const T *__restrict__ e, const T *__restrict__ n, const T *__restrict__ u, int n_obs,
    const T *__restrict__ pe, const T *__restrict__ pn, const T *__restrict__ pu,
    const T *__restrict__ coefs, int64_t n_src, int64_t chunk_len, double *__restrict__ partial) {
    const int i = blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = i < n_obs;
    const int ii = valid ? i : n_obs - 1;
    const T oe = e[ii], on = n[ii], ou = u[ii];
    double acc = 0.0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_src ? c0 + chunk_len : n_src;
#pragma unroll 8
    for (int64_t j = c0; j < c1; ++j) {
        const T de = oe - pe[j], dn = on - pn[j], du = ou - pu[j];
        acc += (double)(coefs[j] * inv_sqrt(de * de + dn * dn + du * du));
    }
    if (valid) partial[(int64_t)blockIdx.y * n_obs + i] = acc;
}

template <class T>
__global__ void eqs_predict_reduce(const double *__restrict__ partial, int64_t chunks,
                                   int64_t n_obs, T *__restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_obs) return;
    double sum = 0.0;
    for (int64_t c = 0; c < chunks; ++c) sum += partial[c * n_obs + i];
    out[i] = (T)sum;
}

template <class T>
int jacobian_typed(const void *e, const void *n, const void *u, int64_t n_data, const void *pe,
                   const void *pn, const void *pu, int64_t n_src, void *jac, int64_t ld,
                   hipStream_t s) {
    unsigned gy = (unsigned)(n_data < 8192 ? n_data : 8192);
    dim3 grid((unsigned)((n_src + 255) / 256), gy);
    hipLaunchKernelGGL(eqs_jacobian_kernel<T>, grid, dim3(256), 0, s, (const T *)e, (const T *)n,
                       (const T *)u, n_data, (const T *)pe, (const T *)pn, (const T *)pu, n_src,
                       (T *)jac, ld);
    HMX_LAUNCHED("eqs_jacobian_kernel");
    return HMX_OK;
}

template <class T>
int predict_typed(const void *e, const void *n, const void *u, int64_t n_obs, const void *pe,
                  const void *pn, const void *pu, const void *coefs, int64_t n_src, void *out,
                  double *partial, int64_t chunks, hipStream_t s) {
    const int64_t chunk_len = (n_src + chunks - 1) / chunks;
    if (n_src > 0) {
        dim3 grid((unsigned)((n_obs + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
        hipLaunchKernelGGL(eqs_predict_kernel<T>, grid, dim3(kObsPerBlock), 0, s, (const T *)e,
                           (const T *)n, (const T *)u, (int)n_obs, (const T *)pe, (const T *)pn,
                           (const T *)pu, (const T *)coefs, n_src, chunk_len, partial);
        HMX_LAUNCHED("eqs_predict_kernel");
    }
    hipLaunchKernelGGL(eqs_predict_reduce<T>, dim3((unsigned)((n_obs + 255) / 256)), dim3(256), 0,
                       s, partial, n_src > 0 ? chunks : 0, n_obs, (T *)out);
    HMX_LAUNCHED("eqs_predict_reduce");
    return HMX_OK;
}

}  // namespace
}  // namespace hmx

namespace hmx {
namespace {

// (longitude [deg], latitude [deg], radius [m]) -> geocentric Cartesian
// (x, y, z) [m].  The spherical equivalent sources' Green's function
// 1 / l, l^2 = (r - r')^2 + 2 r r' (1 - cos psi) (_forward/utils.py:163-201),
// is 1 / |X - X'| of these points, evaluated here without the cancellation
// of 1 - cos psi for nearby points.
__global__ void geocentric_kernel(const double *__restrict__ lon, const double *__restrict__ lat,
                                  const double *__restrict__ radius, int64_t n,
                                  double *__restrict__ x, double *__restrict__ y,
                                  double *__restrict__ z) {
    constexpr double kDeg = 3.14159265358979323846 / 180.0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sl, cl, sp, cp;
    sincos(lon[i] * kDeg, &sl, &cl);
    sincos(lat[i] * kDeg, &sp, &cp);
    const double r = radius[i];
    x[i] = r * cp * cl;
    y[i] = r * cp * sl;
    z[i] = r * sp;
}

}  // namespace
}  // namespace hmx

extern "C" {

int hmx_eqs_jacobian(const void *easting, const void *northing, const void *upward,
                     int64_t n_data, const void *p_easting, const void *p_northing,
                     const void *p_upward, int64_t n_src, int dtype, void *jac, int64_t ld,
                     void *stream) {
    using namespace hmx;
    clear_error();
    if (n_data < 0 || n_src < 0) return fail(HMX_ERR_ARG, "negative size");
    if (ld < n_src) return fail(HMX_ERR_ARG, "ld (%lld) < n_src (%lld)", (long long)ld, (long long)n_src);
    if (n_data == 0 || n_src == 0) return HMX_OK;
    if (!easting || !northing || !upward || !p_easting || !p_northing || !p_upward || !jac)
        return fail(HMX_ERR_ARG, "NULL array");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == HMX_F64)
        return jacobian_typed<double>(easting, northing, upward, n_data, p_easting, p_northing,
                                      p_upward, n_src, jac, ld, s);
    if (dtype == HMX_F32)
        return jacobian_typed<float>(easting, northing, upward, n_data, p_easting, p_northing,
                                     p_upward, n_src, jac, ld, s);
    return fail(HMX_ERR_ARG, "invalid dtype %d", dtype);
}

int hmx_eqs_predict_workspace_bytes(int64_t n_obs, int64_t n_src, size_t *bytes) {
    using namespace hmx;
    if (!bytes) return fail(HMX_ERR_ARG, "bytes is NULL");
    if (n_obs < 0 || n_src < 0) return fail(HMX_ERR_ARG, "negative size");
    int cu = device_cu_count();
    if (cu < 0) return cu;
    *bytes = (size_t)forward_chunks(n_obs, n_src, cu) * (size_t)n_obs * sizeof(double);
    return HMX_OK;
}

int hmx_eqs_predict(const void *easting, const void *northing, const void *upward,
                    int64_t n_obs, const void *p_easting, const void *p_northing,
                    const void *p_upward, const void *coefs, int64_t n_src, int dtype,
                    void *out, void *workspace, size_t workspace_bytes, void *stream) {
    using namespace hmx;
    clear_error();
    if (n_obs < 0 || n_src < 0) return fail(HMX_ERR_ARG, "negative size");
    if (n_obs > INT32_MAX) return fail(HMX_ERR_ARG, "more than 2^31-1 observation points");
    if (n_obs == 0) return HMX_OK;
    if (!easting || !northing || !upward || !out) return fail(HMX_ERR_ARG, "NULL array");
    if (n_src > 0 && (!p_easting || !p_northing || !p_upward || !coefs))
        return fail(HMX_ERR_ARG, "NULL source array");
    int cu = device_cu_count();
    if (cu < 0) return cu;
    const int64_t chunks = forward_chunks(n_obs, n_src, cu);
    const size_t need = (size_t)chunks * (size_t)n_obs * sizeof(double);
    if (n_src > 0 && (!workspace || workspace_bytes < need))
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    double *partial = static_cast<double *>(workspace);
    if (dtype == HMX_F64)
        return predict_typed<double>(easting, northing, upward, n_obs, p_easting, p_northing,
                                     p_upward, coefs, n_src, out, partial, chunks, s);
    if (dtype == HMX_F32)
        return predict_typed<float>(easting, northing, upward, n_obs, p_easting, p_northing,
                                    p_upward, coefs, n_src, out, partial, chunks, s);
    return fail(HMX_ERR_ARG, "invalid dtype %d", dtype);
}

int hmx_geocentric(const double *longitude, const double *latitude, const double *radius,
                   int64_t n, double *x, double *y, double *z, void *stream) {
    using namespace hmx;
    clear_error();
    if (n < 0) return fail(HMX_ERR_ARG, "negative size");
    if (n == 0) return HMX_OK;
    if (!longitude || !latitude || !radius || !x || !y || !z) return fail(HMX_ERR_ARG, "NULL array");
    hipLaunchKernelGGL(geocentric_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, longitude, latitude, radius, n, x, y, z);
    HMX_LAUNCHED("geocentric_kernel");
    return HMX_OK;
}

}  // extern "C"
