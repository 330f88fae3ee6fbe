// Point-mass forward models (Cartesian ten fields, spherical potential and
// g_z) and the equivalent-source jacobian/predict kernels.
//
// Same launch shape as the prism kernels: 256 observation points per
// workgroup, one per lane, sources staged through LDS and read as
// wave-uniform broadcasts, partial sums per source chunk reduced in order.
#include "hmx_internal.h"
#include "prism_device.h"

namespace hmx {
namespace {

using dev::FieldSet;

// Cartesian point-mass kernels (point.py:281-302 table; the potential
// G m / l and its first and second derivatives, point.py:124-203), with
// gm = G * mass; before harmonica's sign flip and units.
template <unsigned M>
__device__ __forceinline__ void point_pair(double e, double n, double u, double pe, double pn,
                                           double pu, double gm,
                                           double (&acc)[FieldSet<M>::count]) {
    using F = FieldSet<M>;
    const double de = e - pe, dn = n - pn, du = u - pu;
    const double d2 = de * de + dn * dn + du * du;
    const double inv = dm::rsqrt_pos(d2);
    const double inv3 = inv * inv * inv;
    const double inv5 = inv3 * inv * inv;
    if constexpr (F::pot) acc[F::slot(0)] += gm * inv;
    if constexpr (F::ge) acc[F::slot(1)] += -gm * de * inv3;
    if constexpr (F::gn) acc[F::slot(2)] += -gm * dn * inv3;
    if constexpr (F::gu) acc[F::slot(3)] += -gm * du * inv3;
    if constexpr (F::ee) acc[F::slot(4)] += gm * (3.0 * de * de - d2) * inv5;
    if constexpr (F::nn) acc[F::slot(5)] += gm * (3.0 * dn * dn - d2) * inv5;
    if constexpr (F::uu) acc[F::slot(6)] += gm * (3.0 * du * du - d2) * inv5;
    if constexpr (F::en) acc[F::slot(7)] += 3.0 * gm * de * dn * inv5;
    if constexpr (F::eu) acc[F::slot(8)] += 3.0 * gm * de * du * inv5;
    if constexpr (F::nu) acc[F::slot(9)] += 3.0 * gm * dn * du * inv5;
}

struct alignas(16) StagedPoint {
    double a, b, c, m;  // Cartesian: e, n, u, G*m
};
struct alignas(16) StagedSphPoint {
    double lon, cosphi, sinphi, radius, mass, pad0, pad1, pad2;
};

template <unsigned M>
__global__ __launch_bounds__(kObsPerBlock) void point_cartesian_kernel(
    const double *__restrict__ oe, const double *__restrict__ on, const double *__restrict__ ou,
    int n_obs, const double *__restrict__ pe, const double *__restrict__ pn,
    const double *__restrict__ pu, const double *__restrict__ mass, int64_t n_src,
    int64_t chunk_len, double *__restrict__ partial) {
    using F = FieldSet<M>;
    __shared__ StagedPoint tile[kTile];
    const int i = blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = i < n_obs;
    const int ii = valid ? i : n_obs - 1;
    const double e = oe[ii], n = on[ii], u = ou[ii];
    double acc[F::count];
#pragma unroll
    for (int s = 0; s < F::count; ++s) acc[s] = 0.0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_src ? c0 + chunk_len : n_src;
    for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
        const int cnt = (int)(c1 - t0 < kTile ? c1 - t0 : kTile);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            StagedPoint sp;
            sp.a = pe[t0 + q];
            sp.b = pn[t0 + q];
            sp.c = pu[t0 + q];
            sp.m = dev::kG * mass[t0 + q];
            tile[q] = sp;
        }
        __syncthreads();
        for (int p = 0; p < cnt; ++p) {
            const StagedPoint sp = tile[p];
            point_pair<M>(e, n, u, sp.a, sp.b, sp.c, sp.m, acc);
        }
    }
    if (valid) {
#pragma unroll
        for (int s = 0; s < F::count; ++s)
            partial[((int64_t)blockIdx.y * F::count + s) * n_obs + i] = acc[s];
    }
}

// Spherical point masses (point.py:401-454): potential (field bit 0) or g_u
// (bit 3), distance of _forward/utils.py:163-201.
template <bool POT>
__global__ __launch_bounds__(kObsPerBlock) void point_spherical_kernel(
    const double *__restrict__ olon, const double *__restrict__ olat,
    const double *__restrict__ orad, int n_obs, const double *__restrict__ plon,
    const double *__restrict__ plat, const double *__restrict__ prad,
    const double *__restrict__ mass, int64_t n_src, int64_t chunk_len,
    double *__restrict__ partial) {
    constexpr double kDeg = 3.14159265358979323846 / 180.0;
    __shared__ StagedSphPoint tile[kTile];
    const int i = blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = i < n_obs;
    const int ii = valid ? i : n_obs - 1;
    const double lon = olon[ii] * kDeg, lat = olat[ii] * kDeg, r = orad[ii];
    const double cosphi = cos(lat), sinphi = sin(lat);
    double acc = 0.0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_src ? c0 + chunk_len : n_src;
    for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
        const int cnt = (int)(c1 - t0 < kTile ? c1 - t0 : kTile);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            StagedSphPoint sp;
            const double la = plat[t0 + q] * kDeg;
            sp.lon = plon[t0 + q] * kDeg;
            sp.cosphi = cos(la);
            sp.sinphi = sin(la);
            sp.radius = prad[t0 + q];
            sp.mass = mass[t0 + q];
            sp.pad0 = sp.pad1 = sp.pad2 = 0.0;
            tile[q] = sp;
        }
        __syncthreads();
        for (int p = 0; p < cnt; ++p) {
            const StagedSphPoint sp = tile[p];
            const double coslambda = cos(sp.lon - lon);
            const double cospsi = sp.sinphi * sinphi + sp.cosphi * cosphi * coslambda;
            const double dr = r - sp.radius;
            const double dist = sqrt(dr * dr + 2.0 * r * sp.radius * (1.0 - cospsi));
            double k;
            if (POT) {
                k = 1.0 / dist * dev::kG;
            } else {
                const double dz = r - sp.radius * cospsi;
                k = -dev::kG * dz / (dist * dist * dist);
            }
            acc += sp.mass * k;
        }
    }
    if (valid) partial[(int64_t)blockIdx.y * n_obs + i] = acc;
}

__global__ void reduce_point_kernel(const double *__restrict__ partial, int64_t chunks, int nf,
                                    int64_t n_obs, uint32_t mask, double *__restrict__ out) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nf * n_obs) return;
    int s = (int)(idx / n_obs);
    int64_t i = idx - (int64_t)s * n_obs;
    double sum = 0.0;
    for (int64_t c = 0; c < chunks; ++c) sum += partial[(c * nf + s) * n_obs + i];
    int f = 0, seen = -1;
    for (; f < 10; ++f)
        if ((mask >> f) & 1u)
            if (++seen == s) break;
    out[idx] = dev::to_harmonica_units(f, sum);
}

template <unsigned M>
int launch_point_cart(const double *c0, const double *c1, const double *c2, int64_t n_obs,
                      const double *p0, const double *p1, const double *p2, const double *m,
                      int64_t n_src, int64_t chunks, double *partial, hipStream_t stream) {
    const int64_t chunk_len = (n_src + chunks - 1) / chunks;
    dim3 grid((unsigned)((n_obs + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
    hipLaunchKernelGGL(point_cartesian_kernel<M>, grid, dim3(kObsPerBlock), 0, stream, c0, c1, c2,
                       (int)n_obs, p0, p1, p2, m, n_src, chunk_len, partial);
    HMX_LAUNCHED("point_cartesian_kernel");
    return HMX_OK;
}

int dispatch_point_cart(uint32_t mask, const double *c0, const double *c1, const double *c2,
                        int64_t n_obs, const double *p0, const double *p1, const double *p2,
                        const double *m, int64_t n_src, int64_t chunks, double *partial,
                        hipStream_t stream) {
#define HMX_CASE(MASK)                                                                        \
    case MASK:                                                                                \
        return launch_point_cart<MASK>(c0, c1, c2, n_obs, p0, p1, p2, m, n_src, chunks,       \
                                       partial, stream);
    switch (mask) {
        HMX_CASE(0x001u)
        HMX_CASE(0x002u)
        HMX_CASE(0x004u)
        HMX_CASE(0x008u)
        HMX_CASE(0x010u)
        HMX_CASE(0x020u)
        HMX_CASE(0x040u)
        HMX_CASE(0x080u)
        HMX_CASE(0x100u)
        HMX_CASE(0x200u)
        HMX_CASE(0x3ffu)
    }
#undef HMX_CASE
    return fail(HMX_ERR_ARG, "no point kernel for field mask 0x%x", mask);
}

}  // namespace
}  // namespace hmx

extern "C" int hmx_point_gravity(const double *c0, const double *c1, const double *c2,
                                 int64_t n_obs, const double *p0, const double *p1,
                                 const double *p2, const double *masses, int64_t n_points,
                                 int coordinate_system, uint32_t field_mask, double *out,
                                 void *workspace, size_t workspace_bytes, void *stream) {
    using namespace hmx;
    clear_error();
    hipStream_t s = (hipStream_t)stream;
    if (n_obs < 0 || n_points < 0) return fail(HMX_ERR_ARG, "negative size");
    if (n_obs > INT32_MAX || n_points > INT32_MAX)
        return fail(HMX_ERR_ARG, "more than 2^31-1 observation points or sources");
    if (field_mask == 0 || (field_mask & ~HMX_ALL_FIELDS))
        return fail(HMX_ERR_ARG, "invalid field mask 0x%x", field_mask);
    if (coordinate_system != HMX_CARTESIAN && coordinate_system != HMX_SPHERICAL)
        return fail(HMX_ERR_ARG, "invalid coordinate system %d", coordinate_system);
    if (coordinate_system == HMX_SPHERICAL && (field_mask & ~(HMX_POTENTIAL | HMX_G_Z)))
        return fail(HMX_ERR_ARG, "spherical point masses support only potential and g_z");
    if (n_obs == 0) return HMX_OK;
    if (!c0 || !c1 || !c2 || !out) return fail(HMX_ERR_ARG, "NULL array");
    if (n_points > 0 && (!p0 || !p1 || !p2 || !masses)) return fail(HMX_ERR_ARG, "NULL array");
    int cu = device_cu_count();
    if (cu < 0) return cu;
    const int nf = popcount10(field_mask);
    const int64_t chunks = forward_chunks(n_obs, n_points, cu);
    if (n_points == 0) {
        // No sources: zeros, written by the reduce kernel over a zero chunk count.
        hipLaunchKernelGGL(reduce_point_kernel, dim3((unsigned)((nf * n_obs + 255) / 256)),
                           dim3(256), 0, s, (const double *)nullptr, (int64_t)0, nf, n_obs,
                           field_mask, out);
        HMX_LAUNCHED("reduce_point_kernel");
        return HMX_OK;
    }
    const size_t need_one = (size_t)chunks * (size_t)n_obs * sizeof(double);
    double *partial = static_cast<double *>(workspace);
    const bool fused = coordinate_system == HMX_CARTESIAN &&
                       (nf == 1 || field_mask == HMX_ALL_FIELDS);
    const size_t need = fused ? need_one * nf : need_one;
    if (!workspace || workspace_bytes < need)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    const int64_t chunk_len = (n_points + chunks - 1) / chunks;
    dim3 grid((unsigned)((n_obs + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
    if (fused) {
        int st = dispatch_point_cart(field_mask, c0, c1, c2, n_obs, p0, p1, p2, masses, n_points,
                                     chunks, partial, s);
        if (st) return st;
        hipLaunchKernelGGL(reduce_point_kernel, dim3((unsigned)((nf * n_obs + 255) / 256)),
                           dim3(256), 0, s, partial, chunks, nf, n_obs, field_mask, out);
        HMX_LAUNCHED("reduce_point_kernel");
        return HMX_OK;
    }
    int slot = 0;
    for (int f = 0; f < 10; ++f) {
        if (!((field_mask >> f) & 1u)) continue;
        const uint32_t m = 1u << f;
        if (coordinate_system == HMX_CARTESIAN) {
            int st = dispatch_point_cart(m, c0, c1, c2, n_obs, p0, p1, p2, masses, n_points, chunks,
                                         partial, s);
            if (st) return st;
        } else if (f == 0) {
            hipLaunchKernelGGL(point_spherical_kernel<true>, grid, dim3(kObsPerBlock), 0, s, c0, c1,
                               c2, (int)n_obs, p0, p1, p2, masses, n_points, chunk_len, partial);
            HMX_LAUNCHED("point_spherical_kernel");
        } else {
            hipLaunchKernelGGL(point_spherical_kernel<false>, grid, dim3(kObsPerBlock), 0, s, c0,
                               c1, c2, (int)n_obs, p0, p1, p2, masses, n_points, chunk_len,
                               partial);
            HMX_LAUNCHED("point_spherical_kernel");
        }
        hipLaunchKernelGGL(reduce_point_kernel, dim3((unsigned)((n_obs + 255) / 256)), dim3(256), 0,
                           s, partial, chunks, 1, n_obs, m, out + (int64_t)slot * n_obs);
        HMX_LAUNCHED("reduce_point_kernel");
        ++slot;
    }
    return HMX_OK;
}
