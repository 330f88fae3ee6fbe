// Gravitational potential and g_z of tesseroids (spherical prisms) by
// adaptive discretisation and Gauss-Legendre quadrature, as
// src/harmonica/_forward/tesseroid_gravity.py:236-336 and
// _tesseroid_utils.py do on the CPU:
//  * per (point, tesseroid), a stack of tesseroids starts with the
//    tesseroid; a popped tesseroid whose distance to the point (from its
//    centre) is below `ratio` times its longitudinal, latitudinal (or, with
//    radial discretisation, radial) size is bisected along those
//    directions and its parts pushed, else it is integrated;
//  * integration: 2 x 2 x 2 Gauss-Legendre nodes as point masses
//    rho(node) * a * r'^2 cos(lat') with the spherical point-mass kernels
//    (point.py:324-354) and the distance of _forward/utils.py:163-201.
// Launch shape as the prism kernels: 256 points per workgroup, one per lane
// (its stack in the lane's private memory), tesseroids staged in LDS tiles
// and read as wave-uniform broadcasts, one chunk of the tesseroid list per
// workgroup column, chunk partials reduced in order by a second kernel.
// Overflows of the stack (100 entries) or of the number of small
// tesseroids per pair (100000) set a flag the host turns into the
// reference's OverflowError.
#include "hmx_internal.h"

namespace hmx {
namespace {

constexpr int kTessStack = 100;
constexpr long kTessMaxSmall = 100000;
constexpr double kDegT = 3.14159265358979323846 / 180.0;
constexpr double kGT = 6.6743e-11;  // harmonica/constants.py:12
// numpy.polynomial.legendre.leggauss(2): nodes -+this, weights 1.
constexpr double kGlqNode = 0x1.279a74590331cp-1;

struct alignas(16) StagedTess {
    double t[6];
    double rho[2];  // density at the lower and upper radial node
};

struct TessPoint {
    double lon, lat, r;            // degrees, degrees, m
    double lon_rad, cosphi, sinphi;
};

// distance_spherical (_forward/utils.py:121-201) from precomputed point
// quantities to (lon_p [deg], lat_p [deg], r_p).
__device__ __forceinline__ double tess_distance(const TessPoint &P, double lon_p, double lat_p,
                                                double r_p) {
    double sp, cp;
    sincos(lat_p * kDegT, &sp, &cp);
    const double coslambda = cos(lon_p * kDegT - P.lon_rad);
    const double cospsi = sp * P.sinphi + cp * P.cosphi * coslambda;
    const double dr = P.r - r_p;
    return sqrt(dr * dr + 2.0 * P.r * r_p * (1.0 - cospsi));
}

// 2 x 2 x 2 Gauss-Legendre quadrature of one tesseroid (_tesseroid_utils.py
// gauss_legendre_quadrature; the density at the radial nodes as
// _tesseroid_variable_density.py).
template <bool kPot>
__device__ __forceinline__ double tess_glq(const TessPoint &P, const double (&t)[6],
                                           const double (&rho)[2]) {
    const double w = t[0], e = t[1], s = t[2], n = t[3], bottom = t[4], top = t[5];
    const double a_factor = 1.0 / 8.0 * ((e - w) * kDegT) * ((n - s) * kDegT) * (top - bottom);
    const double node[2] = {-kGlqNode, kGlqNode};
    double result = 0.0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        double sinphi_p, cosphi_p;
        sincos((0.5 * (n - s) * node[j] + 0.5 * (n + s)) * kDegT, &sinphi_p, &cosphi_p);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double radius_p = 0.5 * (top - bottom) * node[k] + 0.5 * (top + bottom);
            const double kappa = radius_p * radius_p * cosphi_p;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const double lon_p = (0.5 * (e - w) * node[i] + 0.5 * (e + w)) * kDegT;
                const double mass = rho[k] * a_factor * kappa;
                const double coslambda = cos(lon_p - P.lon_rad);
                const double cospsi = sinphi_p * P.sinphi + cosphi_p * P.cosphi * coslambda;
                const double dr = P.r - radius_p;
                const double dist = sqrt(dr * dr + 2.0 * P.r * radius_p * (1.0 - cospsi));
                double kv;
                if constexpr (kPot) {
                    kv = 1.0 / dist * kGT;
                } else {
                    const double dz = P.r - radius_p * cospsi;
                    kv = -kGT * dz / (dist * dist * dist);
                }
                result += mass * kv;
            }
        }
    }
    return result;
}

// One (point, tesseroid) pair: adaptive discretisation with the lane's
// stack, small tesseroids integrated in the order they are produced.
template <bool kPot, bool kRadial>
__device__ double tess_pair(const TessPoint &P, const StagedTess &T, double ratio, int *ovf) {
    double stk[kTessStack][6];
    int top = 0;
    long n_small = 0;
    double result = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) stk[0][c] = T.t[c];
    while (top >= 0) {
        double t[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) t[c] = stk[top][c];
        --top;
        // _tesseroid_dimensions
        const double w = t[0] * kDegT, e = t[1] * kDegT, s = t[2] * kDegT, n = t[3] * kDegT;
        double sn, cn, ss, cs, sc, cc;
        sincos(n, &sn, &cn);
        sincos(s, &ss, &cs);
        sincos((n + s) / 2, &sc, &cc);
        const double l_lat = t[5] * acos(sn * ss + cn * cs);
        const double l_lon = t[5] * acos(sc * sc + cc * cc * cos(e - w));
        const double l_rad = t[5] - t[4];
        // _distance_tesseroid_point
        const double d = tess_distance(P, (t[0] + t[1]) / 2, (t[2] + t[3]) / 2, (t[4] + t[5]) / 2);
        const int n_lon = d / l_lon < ratio ? 2 : 1;
        const int n_lat = d / l_lat < ratio ? 2 : 1;
        const int n_rad = (kRadial && d / l_rad < ratio) ? 2 : 1;
        if (n_lon * n_lat * n_rad > 1) {
            if ((top + 1) + n_lon * n_lat * n_rad > kTessStack) {
                *ovf = 1;
                return 0.0;
            }
            const double d_lon = (t[1] - t[0]) / n_lon, d_lat = (t[3] - t[2]) / n_lat;
            const double d_rad = (t[5] - t[4]) / n_rad;
            for (int i = 0; i < n_lon; ++i)
                for (int j = 0; j < n_lat; ++j)
                    for (int k = 0; k < n_rad; ++k) {
                        ++top;
                        stk[top][0] = t[0] + d_lon * i;
                        stk[top][1] = t[0] + d_lon * (i + 1);
                        stk[top][2] = t[2] + d_lat * j;
                        stk[top][3] = t[2] + d_lat * (j + 1);
                        stk[top][4] = t[4] + d_rad * k;
                        stk[top][5] = t[4] + d_rad * (k + 1);
                    }
        } else {
            if (n_small + 1 > kTessMaxSmall) {
                *ovf = 2;
                return 0.0;
            }
            ++n_small;
            result += tess_glq<kPot>(P, t, T.rho);
        }
    }
    return result;
}

template <bool kPot, bool kRadial>
__global__ __launch_bounds__(kObsPerBlock) void tesseroid_kernel(
    const double *__restrict__ lon, const double *__restrict__ lat,
    const double *__restrict__ rad, int n_obs, const double *__restrict__ tess,
    const double *__restrict__ rho, int64_t n_src, int64_t chunk_len, double ratio,
    double *__restrict__ partial, int *__restrict__ overflow) {
    __shared__ StagedTess tile[kTile];
    const int i = blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = i < n_obs;
    const int ii = valid ? i : n_obs - 1;
    TessPoint P;
    P.lon = lon[ii];
    P.lat = lat[ii];
    P.r = rad[ii];
    P.lon_rad = P.lon * kDegT;
    sincos(P.lat * kDegT, &P.sinphi, &P.cosphi);
    double acc = 0.0;
    int ovf = 0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_src ? c0 + chunk_len : n_src;
    for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
        const int cnt = (int)(c1 - t0 < kTile ? c1 - t0 : kTile);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            StagedTess st;
#pragma unroll
            for (int c = 0; c < 6; ++c) st.t[c] = tess[6 * (t0 + q) + c];
            st.rho[0] = rho[2 * (t0 + q)];
            st.rho[1] = rho[2 * (t0 + q) + 1];
            tile[q] = st;
        }
        __syncthreads();
        for (int p = 0; p < cnt && !ovf; ++p) acc += tess_pair<kPot, kRadial>(P, tile[p], ratio, &ovf);
    }
    if (valid) {
        partial[(int64_t)blockIdx.y * n_obs + i] = acc;
        if (ovf) atomicMax(overflow, ovf);
    }
}

// out[i] = scale * sum over chunks in order of partial[c][i].
__global__ void reduce_tess_kernel(const double *__restrict__ partial, int64_t chunks,
                                   int64_t n_obs, double scale, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_obs) return;
    double sum = 0.0;
    for (int64_t c = 0; c < chunks; ++c) sum += partial[c * n_obs + i];
    out[i] = scale * sum;
}

__global__ void fill_zero_tess_kernel(double *__restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 0.0;
}

}  // namespace
}  // namespace hmx

extern "C" {

int hmx_tesseroid_workspace_bytes(int64_t n_obs, int64_t n_tesseroids, size_t *bytes) {
    hmx::clear_error();
    if (!bytes) return hmx::fail(HMX_ERR_ARG, "bytes is NULL");
    if (n_obs < 0 || n_tesseroids < 0) return hmx::fail(HMX_ERR_ARG, "negative size");
    const int cu = hmx::device_cu_count();
    if (cu < 0) return cu;
    *bytes = (size_t)hmx::forward_chunks(n_obs, n_tesseroids, cu) * (size_t)n_obs * sizeof(double);
    return HMX_OK;
}

int hmx_tesseroid_gravity(const double *longitude, const double *latitude, const double *radius,
                          int64_t n_obs, const double *tesseroids, const double *density_nodes,
                          int64_t n_tesseroids, uint32_t field, int radial_discretization,
                          double *out, int *overflow, void *workspace, size_t workspace_bytes,
                          void *stream) {
    using namespace hmx;
    clear_error();
    if (n_obs < 0 || n_tesseroids < 0) return fail(HMX_ERR_ARG, "negative size");
    if (n_obs > INT32_MAX || n_tesseroids > INT32_MAX)
        return fail(HMX_ERR_ARG, "more than 2^31-1 points or tesseroids");
    if (field != HMX_POTENTIAL && field != HMX_G_Z)
        return fail(HMX_ERR_ARG, "tesseroids support only potential and g_z (mask 0x%x)", field);
    if (n_obs == 0) return HMX_OK;
    if (!longitude || !latitude || !radius || !out || !overflow)
        return fail(HMX_ERR_ARG, "NULL array");
    hipStream_t s = (hipStream_t)stream;
    if (n_tesseroids == 0) {
        hipLaunchKernelGGL(fill_zero_tess_kernel, dim3((unsigned)((n_obs + 255) / 256)),
                           dim3(256), 0, s, out, n_obs);
        HMX_LAUNCHED("fill_zero_tess_kernel");
        return HMX_OK;
    }
    if (!tesseroids || !density_nodes) return fail(HMX_ERR_ARG, "NULL tesseroid array");
    const int cu = device_cu_count();
    if (cu < 0) return cu;
    const int64_t chunks = forward_chunks(n_obs, n_tesseroids, cu);
    const size_t need = (size_t)chunks * (size_t)n_obs * sizeof(double);
    if (!workspace || workspace_bytes < need)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    double *partial = static_cast<double *>(workspace);
    const int64_t chunk_len = (n_tesseroids + chunks - 1) / chunks;
    const bool pot = field == HMX_POTENTIAL;
    // distance-size ratios of tesseroid_gravity.py:32
    const double ratio = pot ? 1.0 : 2.5;
    dim3 grid((unsigned)((n_obs + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
#define HMX_TESS(P, R)                                                                           \
    hipLaunchKernelGGL((tesseroid_kernel<P, R>), grid, dim3(kObsPerBlock), 0, s, longitude,     \
                       latitude, radius, (int)n_obs, tesseroids, density_nodes, n_tesseroids,  \
                       chunk_len, ratio, partial, overflow)
    if (pot && radial_discretization) HMX_TESS(true, true);
    else if (pot) HMX_TESS(true, false);
    else if (radial_discretization) HMX_TESS(false, true);
    else HMX_TESS(false, false);
#undef HMX_TESS
    HMX_LAUNCHED("tesseroid_kernel");
    // g_z points down: the reference flips the sign of the upward kernel
    // and converts to mGal (tesseroid_gravity.py:207-211).
    hipLaunchKernelGGL(reduce_tess_kernel, dim3((unsigned)((n_obs + 255) / 256)), dim3(256), 0,
                       s, partial, chunks, n_obs, pot ? 1.0 : -1e5, out);
    HMX_LAUNCHED("reduce_tess_kernel");
    return HMX_OK;
}

}  // extern "C"
