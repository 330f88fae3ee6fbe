// Magnetic field of uniformly magnetised prisms (explicit prisms).
//
// For a prism with magnetisation M (A/m) the field at a point is
//   b_i = c_m sum_j M_j K_ij,   c_m = mu_0 / (4 pi),  mu_0 = 4 pi 1e-7 H/m,
// with K_ij the vertex sums of the second-derivative kernels of the gravity
// tensor in the east-north-up frame (the same grouped closed forms as the
// gravity kernels, prism_device.h): b_e uses K_ee, K_en, K_eu; b_n uses
// K_en, K_nn, K_nu; b_u uses K_eu, K_nu, K_uu.  Every component is NaN on any
// edge or vertex of a prism.  This replaces the per-pair calls of
// src/harmonica/_forward/prisms/magnetic.py:319-331 (field "b") and
// :384-397 (one component) and the nT conversion of :195-197 and :271.
//
// Launch shape as the gravity forward kernel: 256 points per workgroup (one
// per lane), prisms staged in LDS tiles of 128 and read as wave-uniform
// broadcasts, each tile in two passes (lean closed form for the prisms no
// lane of the wave needs the general form for, then the general form), one
// chunk of the prism list per workgroup column, chunk partial sums reduced
// in order by a second kernel (deterministic results).
#include "hmx_internal.h"
#include "prism_device.h"

namespace hmx {
namespace {

using dev::FieldSet;

constexpr double kPi = 3.141592653589793;
// mu_0 / 4 / pi with mu_0 = 4 pi 1e-7, evaluated in that order.
constexpr double kCm = 4.0 * kPi * 1e-7 / 4.0 / kPi;

// One staged prism (80 bytes): bounds and c_m times the magnetisation.
struct alignas(16) StagedMagPrism {
    double w, e, s, n, b, t, me, mn, mu;
    int pad0, pad1;
};
static_assert(sizeof(StagedMagPrism) == 80, "staged prism layout");
static_assert(kTile <= 128, "the two-pass tile loop marks at most 128 prisms");

// Components selected by bit c of C (0 = b_e, 1 = b_n, 2 = b_u) and the
// tensor kernels (gravity field bits 4..9) they need.
template <unsigned C>
struct MagSet {
    static constexpr unsigned tensor =
        ((C & 1u) ? 0x190u : 0u) | ((C & 2u) ? 0x2a0u : 0u) | ((C & 4u) ? 0x340u : 0u);
    static constexpr int count = __builtin_popcount(C & 7u);
};

__device__ __forceinline__ void stage_mag(StagedMagPrism *tile, const double *__restrict__ prisms,
                                          const double *__restrict__ me,
                                          const double *__restrict__ mn,
                                          const double *__restrict__ mu, int64_t t0, int cnt) {
    double *flat = reinterpret_cast<double *>(tile);
    const double *src = prisms + 6 * t0;
    for (int idx = threadIdx.x; idx < 6 * cnt; idx += blockDim.x) {
        const int p = idx / 6, c = idx - 6 * p;
        flat[10 * p + c] = src[idx];
    }
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        tile[p].me = kCm * me[t0 + p];
        tile[p].mn = kCm * mn[t0 + p];
        tile[p].mu = kCm * mu[t0 + p];
    }
}

// Add the selected components of one prism at (e, n, u) to acc[].  Without
// kGeneral, dev::needs_general<MagSet<C>::tensor> must be false for the
// point (it is then on no face plane, so on no edge), or with kXOnly
// dev::needs_full_general (the point may lie between the x-planes).
template <unsigned C, bool kGeneral, bool kXOnly = false>
__device__ __forceinline__ void magnetic_pair(double e, double n, double u,
                                              const StagedMagPrism &sp,
                                              double (&acc)[MagSet<C>::count],
                                              const dm::Tables &tab) {
    constexpr unsigned M = MagSet<C>::tensor;
    using F = FieldSet<M>;
    const double x0 = sp.e - e, x1 = sp.w - e, y0 = sp.n - n, y1 = sp.s - n;
    const double z0 = sp.t - u, z1 = sp.b - u;
    // K[f - 4] for gravity field bit f: ee 4, nn 5, uu 6, en 7, eu 8, nu 9.
    double K[6];
    if constexpr (kGeneral) {
        double vs[F::count];
        uint32_t sflags = 0;
        dev::prism_fields<M, true>(x0, x1, y0, y1, z0, z1, vs, sflags, tab);
#pragma unroll
        for (int f = 4; f < 10; ++f)
            if ((M >> f) & 1u) K[f - 4] = vs[F::slot(f)];
    } else if constexpr (kXOnly) {
        dev::tensor_xonly<M>(x0, x1, y0, y1, z0, z1, K, tab);
    } else {
        dev::tensor_lean<M>(x0, x1, y0, y1, z0, z1, K, tab);
    }
    int s = 0;
    if constexpr ((C & 1u) != 0) acc[s++] += sp.me * K[0] + sp.mn * K[3] + sp.mu * K[4];
    if constexpr ((C & 2u) != 0) acc[s++] += sp.me * K[3] + sp.mn * K[1] + sp.mu * K[5];
    if constexpr ((C & 4u) != 0) acc[s++] += sp.me * K[4] + sp.mn * K[5] + sp.mu * K[2];
    if constexpr (kGeneral) {
        // On an edge: two coordinates on a bound of the prism, the third
        // between its bounds (vertices included).
        const bool in_e = x1 <= 0.0 && 0.0 <= x0, in_n = y1 <= 0.0 && 0.0 <= y0;
        const bool in_u = z1 <= 0.0 && 0.0 <= z0;
        const bool at_e = x0 == 0.0 || x1 == 0.0, at_n = y0 == 0.0 || y1 == 0.0;
        const bool at_u = z0 == 0.0 || z1 == 0.0;
        const bool edge = (in_e && at_n && at_u) || (in_n && at_e && at_u) || (in_u && at_e && at_n);
        if (edge) {
#pragma unroll
            for (int c = 0; c < MagSet<C>::count; ++c) acc[c] = __builtin_nan("");
        }
    }
}

// The prisms tile[0, cnt) at (e, n, u), in three passes as dev::tile_pairs
// (lean, x-only, general).
template <unsigned C>
__device__ __forceinline__ void magnetic_tile(const StagedMagPrism *tile, int cnt, double e,
                                              double n, double u,
                                              double (&acc)[MagSet<C>::count],
                                              const dm::Tables &tab) {
    constexpr unsigned M = MagSet<C>::tensor;
    uint64_t general_lo = 0, general_hi = 0, xonly_lo = 0, xonly_hi = 0;
    for (int p = 0; p < cnt; ++p) {
        const StagedMagPrism sp = tile[p];
        if (dev::any_lane(dev::needs_general<M>(e, n, u, sp.w, sp.e, sp.s, sp.n, sp.b, sp.t))) {
            const bool full = dev::any_lane(
                dev::needs_full_general<M>(e, n, u, sp.w, sp.e, sp.s, sp.n, sp.b, sp.t));
            uint64_t &lo = full ? general_lo : xonly_lo, &hi = full ? general_hi : xonly_hi;
            if (p < 64)
                lo |= 1ull << p;
            else
                hi |= 1ull << (p - 64);
            continue;
        }
        magnetic_pair<C, false>(e, n, u, sp, acc, tab);
    }
    for (int half = 0; half < 2; ++half) {
        uint64_t m = half ? xonly_hi : xonly_lo;
        while (m) {
            const int p = 64 * half + __builtin_ctzll(m);
            m &= m - 1;
            magnetic_pair<C, false, true>(e, n, u, tile[p], acc, tab);
        }
    }
    for (int half = 0; half < 2; ++half) {
        uint64_t m = half ? general_hi : general_lo;
        while (m) {
            const int p = 64 * half + __builtin_ctzll(m);
            m &= m - 1;
            magnetic_pair<C, true>(e, n, u, tile[p], acc, tab);
        }
    }
}

template <unsigned C>
__global__ __launch_bounds__(kObsPerBlock) void prism_magnetic_kernel(
    const double *__restrict__ oe, const double *__restrict__ on, const double *__restrict__ ou,
    int n_obs, const double *__restrict__ prisms, const double *__restrict__ me,
    const double *__restrict__ mn, const double *__restrict__ mu, int64_t n_src,
    int64_t chunk_len, double *__restrict__ partial) {
    constexpr int NC = MagSet<C>::count;
    __shared__ StagedMagPrism tile[kTile];
    __shared__ dm::Tables tables;
    dm::load_tables(&tables);
    const int i = blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = i < n_obs;
    const int ii = valid ? i : n_obs - 1;
    const double e = oe[ii], n = on[ii], u = ou[ii];
    double acc[NC];
#pragma unroll
    for (int s = 0; s < NC; ++s) acc[s] = 0.0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_src ? c0 + chunk_len : n_src;
    for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
        const int cnt = (int)(c1 - t0 < kTile ? c1 - t0 : kTile);
        __syncthreads();
        stage_mag(tile, prisms, me, mn, mu, t0, cnt);
        __syncthreads();
        magnetic_tile<C>(tile, cnt, e, n, u, acc, tables);
    }
    if (valid) {
#pragma unroll
        for (int s = 0; s < NC; ++s) partial[((int64_t)blockIdx.y * NC + s) * n_obs + i] = acc[s];
    }
}

// out[s][perm[i]] = 1e9 * (sum over chunks in order of partial[c][s][i])
// (T -> nT; perm NULL: the identity).
__global__ void reduce_magnetic_kernel(const double *__restrict__ partial, int64_t chunks, int nc,
                                       int64_t n_obs, const int *__restrict__ perm,
                                       double *__restrict__ out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nc * n_obs) return;
    const int s = (int)(idx / n_obs);
    const int64_t i = idx - (int64_t)s * n_obs;
    double sum = 0.0;
    for (int64_t c = 0; c < chunks; ++c) sum += partial[(c * nc + s) * n_obs + i];
    out[(int64_t)s * n_obs + (perm ? perm[i] : i)] = 1e9 * sum;
}

__global__ void fill_zero_mag_kernel(double *__restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 0.0;
}

template <unsigned C>
int launch_magnetic(const double *oe, const double *on, const double *ou, int64_t n_obs,
                    const double *prisms, const double *me, const double *mn, const double *mu,
                    int64_t n_src, int64_t chunks, double *partial, hipStream_t stream) {
    const int64_t chunk_len = (n_src + chunks - 1) / chunks;
    dim3 grid((unsigned)((n_obs + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
    hipLaunchKernelGGL((prism_magnetic_kernel<C>), grid, dim3(kObsPerBlock), 0, stream, oe, on, ou,
                       (int)n_obs, prisms, me, mn, mu, n_src, chunk_len, partial);
    HMX_LAUNCHED("prism_magnetic_kernel");
    return HMX_OK;
}

bool valid_components(uint32_t mask) { return mask != 0 && (mask & ~7u) == 0; }

}  // namespace
}  // namespace hmx

extern "C" {

int hmx_prism_magnetic_workspace_bytes(int64_t n_obs, int64_t n_prisms, uint32_t component_mask,
                                       size_t *bytes) {
    hmx::clear_error();
    if (!bytes) return hmx::fail(HMX_ERR_ARG, "bytes is NULL");
    if (n_obs < 0 || n_prisms < 0) return hmx::fail(HMX_ERR_ARG, "negative size");
    if (n_obs > INT32_MAX || n_prisms > INT32_MAX)
        return hmx::fail(HMX_ERR_ARG, "more than 2^31-1 observation points or prisms");
    if (!hmx::valid_components(component_mask))
        return hmx::fail(HMX_ERR_ARG, "invalid component mask 0x%x", component_mask);
    const int cu = hmx::device_cu_count();
    if (cu < 0) return cu;
    const int64_t chunks = hmx::forward_chunks(n_obs, n_prisms, cu);
    *bytes = (size_t)chunks * (size_t)__builtin_popcount(component_mask) * (size_t)n_obs *
             sizeof(double);
    // the point order, after the partial sums
    if (hmx::use_point_order(n_obs, n_prisms))
        *bytes = hmx::point_order_offset(*bytes) + hmx::point_order_bytes(n_obs);
    return HMX_OK;
}

int hmx_prism_magnetic(const double *easting, const double *northing, const double *upward,
                       int64_t n_obs, const double *prisms, const double *magnetization_e,
                       const double *magnetization_n, const double *magnetization_u,
                       int64_t n_prisms, uint32_t component_mask, double *out, void *workspace,
                       size_t workspace_bytes, void *stream) {
    hmx::clear_error();
    if (n_obs < 0 || n_prisms < 0) return hmx::fail(HMX_ERR_ARG, "negative size");
    if (n_obs > INT32_MAX || n_prisms > INT32_MAX)
        return hmx::fail(HMX_ERR_ARG, "more than 2^31-1 observation points or prisms");
    if (!hmx::valid_components(component_mask))
        return hmx::fail(HMX_ERR_ARG, "invalid component mask 0x%x", component_mask);
    if (n_obs == 0) return HMX_OK;
    if (!out) return hmx::fail(HMX_ERR_ARG, "out is NULL");
    if (!easting || !northing || !upward) return hmx::fail(HMX_ERR_ARG, "NULL coordinate array");
    hipStream_t st = (hipStream_t)stream;
    const int nc = __builtin_popcount(component_mask);
    if (n_prisms == 0) {  // no prisms: zeros, as the reference's accumulators
        const int64_t total = (int64_t)nc * n_obs;
        hipLaunchKernelGGL(hmx::fill_zero_mag_kernel, dim3((unsigned)((total + 255) / 256)),
                           dim3(256), 0, st, out, total);
        HMX_LAUNCHED("fill_zero_mag_kernel");
        return HMX_OK;
    }
    if (!prisms || !magnetization_e || !magnetization_n || !magnetization_u)
        return hmx::fail(HMX_ERR_ARG, "NULL prism or magnetization array");
    const int cu = hmx::device_cu_count();
    if (cu < 0) return cu;
    const int64_t chunks = hmx::forward_chunks(n_obs, n_prisms, cu);
    const size_t partial_bytes = (size_t)chunks * (size_t)nc * (size_t)n_obs * sizeof(double);
    const bool order = hmx::use_point_order(n_obs, n_prisms);
    const size_t order_at = hmx::point_order_offset(partial_bytes);
    const size_t need = order ? order_at + hmx::point_order_bytes(n_obs) : partial_bytes;
    if (!workspace || workspace_bytes < need)
        return hmx::fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes,
                         need);
    double *partial = static_cast<double *>(workspace);
    // points in wavefront-coherent order (see point_order.hip)
    const int *perm = nullptr;
    if (order) {
        hmx::PointOrder po;
        const int ost = hmx::point_order(easting, northing, upward, n_obs,
                                         static_cast<char *>(workspace) + order_at, &po, st);
        if (ost) return ost;
        easting = po.e;
        northing = po.n;
        upward = po.u;
        perm = po.perm;
    }
    int rc;
    switch (component_mask) {
#define HMX_MAG_CASE(C)                                                                       \
    case C:                                                                                   \
        rc = hmx::launch_magnetic<C>(easting, northing, upward, n_obs, prisms,                \
                                     magnetization_e, magnetization_n, magnetization_u,       \
                                     n_prisms, chunks, partial, st);                          \
        break;
        HMX_MAG_CASE(1u)
        HMX_MAG_CASE(2u)
        HMX_MAG_CASE(3u)
        HMX_MAG_CASE(4u)
        HMX_MAG_CASE(5u)
        HMX_MAG_CASE(6u)
        HMX_MAG_CASE(7u)
#undef HMX_MAG_CASE
    default:
        return hmx::fail(HMX_ERR_ARG, "invalid component mask 0x%x", component_mask);
    }
    if (rc) return rc;
    const int64_t total = (int64_t)nc * n_obs;
    hipLaunchKernelGGL(hmx::reduce_magnetic_kernel, dim3((unsigned)((total + 255) / 256)),
                       dim3(256), 0, st, partial, chunks, nc, n_obs, perm, out);
    HMX_LAUNCHED("reduce_magnetic_kernel");
    return HMX_OK;
}

}  // extern "C"
