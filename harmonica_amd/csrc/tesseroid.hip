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

// Constants of one undivided tesseroid, computed once per tesseroid by
// tess_root_kernel and staged in LDS: the box and the densities at the
// radial nodes, the dimensions and centre of the discretisation test, and
// the trigonometry, radii and masses of the 2 x 2 x 2 quadrature nodes.
// Most pairs never divide the tesseroid: for them these leave three cosines
// of longitude differences per pair (the centre's and the two longitude
// nodes'), taken as the reference does, cos(lambda - lambda_p): where the
// point is close, 1 - cos(psi) is small and any other rounding of it would
// show in the result.
struct alignas(16) TessRoot {
    double t[6];    // w, e, s, n (degrees), bottom, top (m)
    double rho[2];  // density at the lower and upper radial node
    double l_lon, l_lat, l_rad;
    double sin_lat_c, cos_lat_c, lon_c, r_c;   // centre (lon_c in radians)
    double sin_lat_n[2], cos_lat_n[2], lon_n[2], radius_n[2];  // nodes (lon_n in radians)
    double mass[2][2];  // [latitude node][radial node]: rho a r^2 cos(lat)
};
constexpr int kRootDoubles = sizeof(TessRoot) / sizeof(double);
static_assert(sizeof(TessRoot) == 28 * sizeof(double), "TessRoot is 28 doubles");

struct TessPoint {
    double lon, lat, r;            // degrees, degrees, m
    double lon_rad, cosphi, sinphi;
};

// _tesseroid_dimensions (_tesseroid_utils.py): the longitudinal,
// latitudinal and radial sizes of a tesseroid at its top.
__device__ __forceinline__ void tess_dimensions(const double (&t)[6], double &l_lon,
                                                double &l_lat, double &l_rad) {
    const double w = t[0] * kDegT, e = t[1] * kDegT, s = t[2] * kDegT, n = t[3] * kDegT;
    double sn, cn, ss, cs, sc, cc;
    sincos(n, &sn, &cn);
    sincos(s, &ss, &cs);
    sincos((n + s) / 2, &sc, &cc);
    l_lat = t[5] * acos(sn * ss + cn * cs);
    l_lon = t[5] * acos(sc * sc + cc * cc * cos(e - w));
    l_rad = t[5] - t[4];
}

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

// tess_glq of an undivided tesseroid from its precomputed constants (the
// same nodes, masses and order of summation).
template <bool kPot>
__device__ __forceinline__ double tess_glq_root(const TessPoint &P, const TessRoot &R) {
    double coslambda[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) coslambda[i] = cos(R.lon_n[i] - P.lon_rad);
    double result = 0.0;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double radius_p = R.radius_n[k];
            const double dr = P.r - radius_p;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const double cospsi = R.sin_lat_n[j] * P.sinphi + R.cos_lat_n[j] * P.cosphi * coslambda[i];
                const double dist = sqrt(dr * dr + 2.0 * P.r * radius_p * (1.0 - cospsi));
                double kv;
                if constexpr (kPot) {
                    kv = 1.0 / dist * kGT;
                } else {
                    const double dz = P.r - radius_p * cospsi;
                    kv = -kGT * dz / (dist * dist * dist);
                }
                result += R.mass[j][k] * kv;
            }
        }
    return result;
}

// The parts of tesseroid t for a split into n_lon x n_lat x n_rad, pushed
// on the stack (false if the stack would overflow).
__device__ __forceinline__ bool tess_push_parts(double (&stk)[kTessStack][6], int &top,
                                                const double (&t)[6], int n_lon, int n_lat,
                                                int n_rad) {
    if ((top + 1) + n_lon * n_lat * n_rad > kTessStack) return false;
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
    return true;
}

// Density of a tesseroid's radial profile from its Chebyshev coefficients
// c[0..kChebN) over [B, T] (the undivided tesseroid's radial extent):
// sum_k c_k T_k(x), x = (2 r - (B + T)) / (T - B), by Clenshaw's recurrence.
constexpr int kChebN = 32;
__device__ __forceinline__ double cheb_density(const double *__restrict__ c, double r, double B,
                                               double T) {
    double x = (2.0 * r - (B + T)) / (T - B);
    x = fmin(1.0, fmax(-1.0, x));
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = kChebN - 1; k >= 1; --k) {
        const double b0 = fma(2.0 * x, b1, c[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    return fma(x, b1, c[0] - b2);
}

// One (point, tesseroid) pair: adaptive discretisation with the lane's
// stack, small tesseroids integrated in the order they are produced.  The
// undivided tesseroid is tested and, if small enough, integrated from its
// precomputed constants; its parts go through the generic path.
// With radial discretisation and a variable density (cheb: the tesseroid's
// Chebyshev coefficients, else NULL), the parts' radial nodes take their
// densities from the coefficients; otherwise they share the undivided
// tesseroid's nodes (radially undivided parts) or its constant density.
template <bool kPot, bool kRadial>
__device__ double tess_pair(const TessPoint &P, const TessRoot &R, double ratio, int *ovf,
                            const double *__restrict__ cheb) {
    {
        // _distance_tesseroid_point from the centre
        const double coslambda = cos(R.lon_c - P.lon_rad);
        const double cospsi = R.sin_lat_c * P.sinphi + R.cos_lat_c * P.cosphi * coslambda;
        const double dr = P.r - R.r_c;
        const double d = sqrt(dr * dr + 2.0 * P.r * R.r_c * (1.0 - cospsi));
        const int n_lon = d / R.l_lon < ratio ? 2 : 1;
        const int n_lat = d / R.l_lat < ratio ? 2 : 1;
        const int n_rad = (kRadial && d / R.l_rad < ratio) ? 2 : 1;
        if (n_lon * n_lat * n_rad == 1) return tess_glq_root<kPot>(P, R);
        // divided: continue below with the parts
        double stk[kTessStack][6];
        int top = -1;
        (void)tess_push_parts(stk, top, R.t, n_lon, n_lat, n_rad);  // at most 8 < kTessStack
        long n_small = 0;
        double result = 0.0;
        while (top >= 0) {
            double t[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) t[c] = stk[top][c];
            --top;
            double l_lon, l_lat, l_rad;
            tess_dimensions(t, l_lon, l_lat, l_rad);
            // _distance_tesseroid_point
            const double d2 =
                tess_distance(P, (t[0] + t[1]) / 2, (t[2] + t[3]) / 2, (t[4] + t[5]) / 2);
            const int m_lon = d2 / l_lon < ratio ? 2 : 1;
            const int m_lat = d2 / l_lat < ratio ? 2 : 1;
            const int m_rad = (kRadial && d2 / l_rad < ratio) ? 2 : 1;
            if (m_lon * m_lat * m_rad > 1) {
                if (!tess_push_parts(stk, top, t, m_lon, m_lat, m_rad)) {
                    *ovf = 1;
                    return 0.0;
                }
            } else {
                if (n_small + 1 > kTessMaxSmall) {
                    *ovf = 2;
                    return 0.0;
                }
                ++n_small;
                if (kRadial && cheb) {
                    double rho_p[2];
                    const double node[2] = {-kGlqNode, kGlqNode};
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        rho_p[k] = cheb_density(
                            cheb, 0.5 * (t[5] - t[4]) * node[k] + 0.5 * (t[5] + t[4]), R.t[4],
                            R.t[5]);
                    result += tess_glq<kPot>(P, t, rho_p);
                } else {
                    result += tess_glq<kPot>(P, t, R.rho);
                }
            }
        }
        return result;
    }
}

// The TessRoot constants of each tesseroid (one per thread).
__global__ __launch_bounds__(256) void tess_root_kernel(const double *__restrict__ tess,
                                                        const double *__restrict__ rho,
                                                        int64_t n, TessRoot *__restrict__ roots) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    TessRoot R;
#pragma unroll
    for (int c = 0; c < 6; ++c) R.t[c] = tess[6 * q + c];
    R.rho[0] = rho[2 * q];
    R.rho[1] = rho[2 * q + 1];
    tess_dimensions(R.t, R.l_lon, R.l_lat, R.l_rad);
    const double w = R.t[0], e = R.t[1], s = R.t[2], n_ = R.t[3], bottom = R.t[4], top = R.t[5];
    sincos((s + n_) / 2 * kDegT, &R.sin_lat_c, &R.cos_lat_c);
    R.lon_c = (w + e) / 2 * kDegT;
    R.r_c = (bottom + top) / 2;
    const double a_factor = 1.0 / 8.0 * ((e - w) * kDegT) * ((n_ - s) * kDegT) * (top - bottom);
    const double node[2] = {-kGlqNode, kGlqNode};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        sincos((0.5 * (n_ - s) * node[j] + 0.5 * (n_ + s)) * kDegT, &R.sin_lat_n[j], &R.cos_lat_n[j]);
        R.lon_n[j] = (0.5 * (e - w) * node[j] + 0.5 * (e + w)) * kDegT;
        R.radius_n[j] = 0.5 * (top - bottom) * node[j] + 0.5 * (top + bottom);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k)
            R.mass[j][k] = R.rho[k] * a_factor * (R.radius_n[k] * R.radius_n[k] * R.cos_lat_n[j]);
    roots[q] = R;
}

template <bool kPot, bool kRadial>
__global__ __launch_bounds__(kObsPerBlock) void tesseroid_kernel(
    const double *__restrict__ lon, const double *__restrict__ lat,
    const double *__restrict__ rad, int n_obs, const TessRoot *__restrict__ roots,
    int64_t n_src, int64_t chunk_len, double ratio, const double *__restrict__ cheb,
    double *__restrict__ partial, int *__restrict__ overflow) {
    __shared__ TessRoot tile[kTile];
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
        // cnt x kRootDoubles doubles, one per lane in turn
        const double *src = reinterpret_cast<const double *>(roots + t0);
        double *dst = reinterpret_cast<double *>(tile);
        for (int q = threadIdx.x; q < cnt * kRootDoubles; q += blockDim.x) dst[q] = src[q];
        __syncthreads();
        for (int p = 0; p < cnt && !ovf; ++p)
            acc += tess_pair<kPot, kRadial>(P, tile[p], ratio, &ovf,
                                            cheb ? cheb + (t0 + p) * kChebN : nullptr);
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

// Workspace: the chunk partials, then (256-byte aligned) one TessRoot per
// tesseroid.
inline size_t tess_partial_bytes(int64_t chunks, int64_t n_obs) {
    return ((size_t)chunks * (size_t)n_obs * sizeof(double) + 255) / 256 * 256;
}
inline size_t tess_workspace_bytes(int64_t chunks, int64_t n_obs, int64_t n_tess) {
    return tess_partial_bytes(chunks, n_obs) + (size_t)n_tess * sizeof(TessRoot);
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
    *bytes = hmx::tess_workspace_bytes(hmx::forward_chunks(n_obs, n_tesseroids, cu), n_obs,
                                       n_tesseroids);
    return HMX_OK;
}

int hmx_tesseroid_gravity(const double *longitude, const double *latitude, const double *radius,
                          int64_t n_obs, const double *tesseroids, const double *density_nodes,
                          const double *density_cheb, int64_t n_tesseroids, uint32_t field,
                          int radial_discretization, double *out, int *overflow, void *workspace,
                          size_t workspace_bytes, void *stream) {
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
    const size_t need = tess_workspace_bytes(chunks, n_obs, n_tesseroids);
    if (!workspace || workspace_bytes < need)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    double *partial = static_cast<double *>(workspace);
    TessRoot *roots = reinterpret_cast<TessRoot *>(static_cast<char *>(workspace) +
                                                   tess_partial_bytes(chunks, n_obs));
    hipLaunchKernelGGL(tess_root_kernel, dim3((unsigned)((n_tesseroids + 255) / 256)), dim3(256),
                       0, s, tesseroids, density_nodes, n_tesseroids, roots);
    HMX_LAUNCHED("tess_root_kernel");
    const int64_t chunk_len = (n_tesseroids + chunks - 1) / chunks;
    const bool pot = field == HMX_POTENTIAL;
    // distance-size ratios of tesseroid_gravity.py:32
    const double ratio = pot ? 1.0 : 2.5;
    dim3 grid((unsigned)((n_obs + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
#define HMX_TESS(P, R)                                                                           \
    hipLaunchKernelGGL((tesseroid_kernel<P, R>), grid, dim3(kObsPerBlock), 0, s, longitude,     \
                       latitude, radius, (int)n_obs, roots, n_tesseroids, chunk_len, ratio,     \
                       density_cheb, partial, overflow)
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
