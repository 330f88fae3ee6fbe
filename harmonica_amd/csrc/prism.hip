// Prism forward-modelling kernels (explicit prisms and prism layers).
//
// Launch shape: a workgroup owns 256 observation points (4 wavefronts, one
// point per lane, FP64 accumulators in VGPRs) and one chunk of the prism
// list.  Prisms are staged tile by tile into LDS with coalesced loads and
// read back as wave-uniform broadcasts, so every lane of a wavefront
// evaluates the same prism against its own point.  Each chunk writes its
// partial sums to the workspace; a second kernel reduces the chunks in
// order and applies harmonica's units, so results are deterministic.
#include <mutex>

#include "hmx_internal.h"
#include "prism_device.h"

namespace hmx {
namespace {

using dev::FieldSet;

// One staged prism (64 bytes): bounds, G * density, skip flag and the edge
// it shares with the previous prism of the tile (dev::shared_edge).
struct alignas(16) StagedPrism {
    double w, e, s, n, b, t, grho;
    int skip, share;
};

// Stage prisms [t0, t0 + cnt) of an (m, 6) row-major array.
__device__ __forceinline__ void stage_prisms(StagedPrism *tile, const double *__restrict__ prisms,
                                             const double *__restrict__ density, int64_t t0,
                                             int cnt) {
    double *flat = reinterpret_cast<double *>(tile);
    const double *src = prisms + 6 * t0;
    for (int idx = threadIdx.x; idx < 6 * cnt; idx += blockDim.x) {
        int p = idx / 6, c = idx - 6 * p;
        flat[8 * p + c] = src[idx];
    }
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        tile[p].grho = dev::kG * density[t0 + p];
        tile[p].skip = 0;
        tile[p].share = 0;
    }
}

// Stage prisms [t0, t0 + cnt) of a layer, enumerated easting-major as the
// reference loops (layer.py:590-594): p = j * n_north + k.  Bounds are
// centre -/+ half the first spacing (layer.py:584-585,592-593,609-610);
// skip rules of layer.py:595-607.
__device__ __forceinline__ void stage_layer(StagedPrism *tile, const double *__restrict__ ce,
                                            int n_east, const double *__restrict__ cn,
                                            int n_north, const double *__restrict__ bottom,
                                            const double *__restrict__ top,
                                            const double *__restrict__ density, double threshold,
                                            int64_t t0, int cnt) {
    const double half_e = (ce[1] - ce[0]) / 2;
    const double half_n = (cn[1] - cn[0]) / 2;
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
        int64_t p = t0 + q;
        int j = (int)(p / n_north);
        int k = (int)(p - (int64_t)j * n_north);
        int64_t idx = (int64_t)k * n_east + j;
        double d = density[idx], b = bottom[idx], t = top[idx];
        bool skip = d == 0.0 || d != d || (t - b < threshold) || t != t || b != b;
        StagedPrism sp;
        sp.w = ce[j] - half_e;
        sp.e = ce[j] + half_e;
        sp.s = cn[k] - half_n;
        sp.n = cn[k] + half_n;
        sp.b = b;
        sp.t = t;
        sp.grho = dev::kG * d;
        sp.skip = skip ? 1 : 0;
        sp.share = 0;
        tile[q] = sp;
    }
}

struct ExplicitSource {
    const double *prisms, *density;
    __device__ void stage(StagedPrism *tile, int64_t t0, int cnt) const {
        stage_prisms(tile, prisms, density, t0, cnt);
    }
    static constexpr bool may_skip = false;
};

struct LayerSource {
    const double *ce, *cn, *bottom, *top, *density;
    int n_east, n_north;
    double threshold;
    __device__ void stage(StagedPrism *tile, int64_t t0, int cnt) const {
        stage_layer(tile, ce, n_east, cn, n_north, bottom, top, density, threshold, t0, cnt);
    }
    static constexpr bool may_skip = true;
};

static_assert(kTile <= 128, "tile_pairs marks at most 128 prisms");

template <unsigned M, class Src>
__global__ __launch_bounds__(kObsPerBlock) void prism_forward_kernel(
    const double *__restrict__ oe, const double *__restrict__ on, const double *__restrict__ ou,
    int n_obs, Src src, int64_t n_src, int64_t chunk_len, double *__restrict__ partial,
    uint32_t *__restrict__ flags) {
    using F = FieldSet<M>;
    __shared__ StagedPrism tile[kTile];
    __shared__ dm::Tables tables;
    dm::load_tables(&tables);
    const int i = blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = i < n_obs;
    const int ii = valid ? i : n_obs - 1;
    const double e = oe[ii], n = on[ii], u = ou[ii];
    double acc[F::count];
#pragma unroll
    for (int s = 0; s < F::count; ++s) acc[s] = 0.0;
    uint32_t sflags = 0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_src ? c0 + chunk_len : n_src;
    for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
        const int cnt = (int)(c1 - t0 < kTile ? c1 - t0 : kTile);
        __syncthreads();
        src.stage(tile, t0, cnt);
        __syncthreads();
        if constexpr (M == 0x8u) {  // edges shared with the previous prism (g_u reuse)
            int sh[(kTile + kObsPerBlock - 1) / kObsPerBlock];
#pragma unroll
            for (int k = 0; k * kObsPerBlock < kTile; ++k) {
                const int q = threadIdx.x + k * kObsPerBlock;
                sh[k] = (q > 0 && q < cnt) ? dev::shared_edge(tile[q - 1], tile[q]) : 0;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k * kObsPerBlock < kTile; ++k) {
                const int q = threadIdx.x + k * kObsPerBlock;
                if (q < cnt) tile[q].share = sh[k];
            }
            __syncthreads();
        }
        dev::tile_pairs<M, Src::may_skip>(tile, cnt, e, n, u, acc, sflags, tables);
    }
    if (valid) {
#pragma unroll
        for (int s = 0; s < F::count; ++s)
            partial[((int64_t)blockIdx.y * F::count + s) * n_obs + i] = acc[s];
        if (F::tensor && flags && sflags) atomicOr(flags, sflags);
    }
}

// Singular-point scan: bit f of *flags is set when some (point, prism)
// pair is a singular point of tensor field f (replaces the serial scans
// _any_singular_point_g_* of gravity.py:272-449).
__global__ __launch_bounds__(kObsPerBlock) void singular_scan_kernel(
    const double *__restrict__ oe, const double *__restrict__ on, const double *__restrict__ ou,
    int n_obs, const double *__restrict__ prisms, int64_t n_src, int64_t chunk_len,
    uint32_t mask, uint32_t *__restrict__ flags) {
    __shared__ double tile[kTile * 6];
    const int i = blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = i < n_obs;
    const int ii = valid ? i : n_obs - 1;
    const double e = oe[ii], n = on[ii], u = ou[ii];
    uint32_t found = 0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_src ? c0 + chunk_len : n_src;
    for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
        const int cnt = (int)(c1 - t0 < kTile ? c1 - t0 : kTile);
        __syncthreads();
        for (int idx = threadIdx.x; idx < 6 * cnt; idx += blockDim.x) tile[idx] = prisms[6 * t0 + idx];
        __syncthreads();
        for (int p = 0; p < cnt; ++p) {
            const double *b = tile + 6 * p;
            const bool in_e = b[0] <= e && e <= b[1], in_n = b[2] <= n && n <= b[3];
            const bool in_u = b[4] <= u && u <= b[5];
            const bool at_e = e == b[0] || e == b[1], at_n = n == b[2] || n == b[3];
            const bool at_u = u == b[4] || u == b[5];
            const bool edge_e = in_e && at_n && at_u, edge_n = in_n && at_e && at_u;
            const bool edge_u = in_u && at_e && at_n;
            found |= (edge_n || edge_u) ? 1u << 4 : 0u;
            found |= (edge_e || edge_u) ? 1u << 5 : 0u;
            found |= (edge_e || edge_n) ? 1u << 6 : 0u;
            found |= edge_u ? 1u << 7 : 0u;
            found |= edge_n ? 1u << 8 : 0u;
            found |= edge_e ? 1u << 9 : 0u;
        }
    }
    found &= mask;
    if (valid && found) atomicOr(flags, found);
}

// out[s][perm[i]] = units(sum over chunks in order of partial[c][s][i])
// (perm NULL: the identity).
__global__ void reduce_chunks_kernel(const double *__restrict__ partial, int64_t chunks, int nf,
                                     int64_t n_obs, uint32_t mask, const int *__restrict__ perm,
                                     double *__restrict__ out) {
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
    out[(int64_t)s * n_obs + (perm ? perm[i] : i)] = dev::to_harmonica_units(f, sum);
}

// Zero fill with a kernel (stays valid inside captured HIP graphs).
__global__ void fill_zero_kernel(double *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 0.0;
}

template <unsigned M, class Src>
int launch_forward(const double *oe, const double *on, const double *ou, int64_t n_obs,
                   const Src &src, int64_t n_src, int64_t chunks, double *partial,
                   uint32_t *flags, hipStream_t stream) {
    const int64_t chunk_len = (n_src + chunks - 1) / chunks;
    dim3 grid((unsigned)((n_obs + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
    hipLaunchKernelGGL((prism_forward_kernel<M, Src>), grid, dim3(kObsPerBlock), 0, stream, oe, on,
                       ou, (int)n_obs, src, n_src, chunk_len, partial, flags);
    HMX_LAUNCHED("prism_forward_kernel");
    return HMX_OK;
}

int launch_reduce(const double *partial, int64_t chunks, int nf, int64_t n_obs, uint32_t mask,
                  const int *perm, double *out, hipStream_t stream) {
    int64_t total = (int64_t)nf * n_obs;
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, partial, chunks, nf, n_obs, mask, perm, out);
    HMX_LAUNCHED("reduce_chunks_kernel");
    return HMX_OK;
}

// Field sets with their own kernel instantiation: each single field, the
// three accelerations, the six tensor terms and all ten.
template <class Src>
int dispatch_mask(uint32_t mask, const double *oe, const double *on, const double *ou,
                  int64_t n_obs, const Src &src, int64_t n_src, int64_t chunks, double *partial,
                  uint32_t *flags, hipStream_t stream) {
#define HMX_CASE(MASK)                                                                          \
    case MASK:                                                                                  \
        return launch_forward<MASK, Src>(oe, on, ou, n_obs, src, n_src, chunks, partial, flags, \
                                         stream);
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
        HMX_CASE(0x00eu)
        HMX_CASE(0x3f0u)
        HMX_CASE(0x3ffu)
    }
#undef HMX_CASE
    return fail(HMX_ERR_ARG, "no kernel for field mask 0x%x", mask);
}

bool has_kernel(uint32_t mask) {
    return __builtin_popcount(mask) == 1 || mask == 0x00eu || mask == 0x3f0u || mask == 0x3ffu;
}

// Run the forward model for any field mask: one launch when the mask has
// its own kernel, else one launch per field reusing the workspace.
// chunk_cap > 0 limits the number of source chunks (so that the partial
// sums fit a fixed scratch buffer).  With may_order, calls that pass
// use_point_order evaluate the points in wavefront-coherent order, sorted
// into the workspace after the room for all fields' partial sums (the
// layout hmx_forward_workspace_bytes sizes).
template <class Src>
int run_forward(const double *oe, const double *on, const double *ou, int64_t n_obs,
                const Src &src, int64_t n_src, uint32_t mask, double *out, uint32_t *flags,
                void *workspace, size_t workspace_bytes, hipStream_t stream,
                int64_t chunk_cap = 0, bool may_order = true) {
    if (n_obs <= 0) return HMX_OK;
    int cu = device_cu_count();
    if (cu < 0) return cu;
    const int nf = popcount10(mask);
    int64_t chunks = forward_chunks(n_obs, n_src, cu);
    if (chunk_cap > 0 && chunks > chunk_cap) chunks = chunk_cap;
    if (n_src == 0) {
        // No sources: the reference returns zeros.
        int64_t total = (int64_t)nf * n_obs;
        hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                           stream, out, total);
        HMX_LAUNCHED("fill_zero_kernel");
        return HMX_OK;
    }
    const size_t need_one = (size_t)chunks * (size_t)n_obs * sizeof(double);
    double *partial = static_cast<double *>(workspace);
    const bool order = may_order && use_point_order(n_obs, n_src);
    const size_t order_at = point_order_offset(need_one * nf);
    const size_t need = order ? order_at + point_order_bytes(n_obs)
                              : (has_kernel(mask) ? need_one * nf : need_one);
    if (!workspace || workspace_bytes < need)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    const int *perm = nullptr;
    if (order) {
        PointOrder po;
        int st = point_order(oe, on, ou, n_obs, static_cast<char *>(workspace) + order_at, &po,
                             stream);
        if (st) return st;
        oe = po.e;
        on = po.n;
        ou = po.u;
        perm = po.perm;
    }
    if (has_kernel(mask)) {
        int st = dispatch_mask<Src>(mask, oe, on, ou, n_obs, src, n_src, chunks, partial, flags,
                                    stream);
        if (st) return st;
        return launch_reduce(partial, chunks, nf, n_obs, mask, perm, out, stream);
    }
    int s = 0;
    for (int f = 0; f < 10; ++f) {
        if (!((mask >> f) & 1u)) continue;
        uint32_t m = 1u << f;
        int st = dispatch_mask<Src>(m, oe, on, ou, n_obs, src, n_src, chunks, partial, flags,
                                    stream);
        if (st) return st;
        st = launch_reduce(partial, chunks, 1, n_obs, m, perm, out + (int64_t)s * n_obs, stream);
        if (st) return st;
        ++s;
    }
    return HMX_OK;
}

int check_common(int64_t n_obs, int64_t n_src, uint32_t mask, const void *out) {
    if (n_obs < 0 || n_src < 0) return fail(HMX_ERR_ARG, "negative size");
    if (n_obs > INT32_MAX || n_src > INT32_MAX)
        return fail(HMX_ERR_ARG, "more than 2^31-1 observation points or sources");
    if (mask == 0 || (mask & ~HMX_ALL_FIELDS)) return fail(HMX_ERR_ARG, "invalid field mask 0x%x", mask);
    if (n_obs > 0 && !out) return fail(HMX_ERR_ARG, "out is NULL");
    return HMX_OK;
}

// Scratch of hmx_prism_gravity_static: the chunk partial sums, in a
// file-scope device array (one per device, held until the process ends).
// Host code uses its per-device address from hipGetSymbolAddress, never the
// array's name.
constexpr int64_t kStaticScratchDoubles = (int64_t)1 << 22;  // 32 MiB
__device__ __align__(256) double scratch[kStaticScratchDoubles];

std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch array (looked up once per
// device, with that device current).
int static_scratch(double **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(HMX_ERR_NODEV, "no HIP device: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(HMX_ERR_NODEV, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(HMX_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return HMX_OK;
}

}  // namespace
}  // namespace hmx

extern "C" {

int hmx_prism_gravity_static(const double *easting, const double *northing,
                             const double *upward, int64_t n_obs, const double *prisms,
                             const double *density, int64_t n_prisms, uint32_t field_mask,
                             double *out, uint32_t *singular_flags, void *stream) {
    using namespace hmx;
    clear_error();
    int st = check_common(n_obs, n_prisms, field_mask, out);
    if (st) return st;
    if (n_obs > 0 && (!easting || !northing || !upward))
        return fail(HMX_ERR_ARG, "NULL coordinate array");
    if (n_prisms > 0 && (!prisms || !density)) return fail(HMX_ERR_ARG, "NULL prism array");
    // one launch holds the partial sums of all its fields for at least one
    // chunk: n_obs x (fields of one launch) doubles
    const int64_t per_chunk = n_obs * (has_kernel(field_mask) ? popcount10(field_mask) : 1);
    if (per_chunk > kStaticScratchDoubles)
        return fail(HMX_ERR_ARG,
                    "hmx_prism_gravity_static: %lld observation points x fields exceed its scratch "
                    "of %lld doubles (use hmx_prism_gravity with a workspace)",
                    (long long)per_chunk, (long long)kStaticScratchDoubles);
    if (n_obs == 0) return HMX_OK;
    double *ws = nullptr;
    st = static_scratch(&ws);
    if (st) return st;
    ExplicitSource src{prisms, density};
    return run_forward(easting, northing, upward, n_obs, src, n_prisms, field_mask, out,
                       singular_flags, ws, (size_t)kStaticScratchDoubles * sizeof(double),
                       (hipStream_t)stream, kStaticScratchDoubles / per_chunk, false);
}

int hmx_prism_gravity(const double *easting, const double *northing, const double *upward,
                      int64_t n_obs, const double *prisms, const double *density,
                      int64_t n_prisms, uint32_t field_mask, double *out,
                      uint32_t *singular_flags, void *workspace, size_t workspace_bytes,
                      void *stream) {
    hmx::clear_error();
    int st = hmx::check_common(n_obs, n_prisms, field_mask, out);
    if (st) return st;
    if (n_obs > 0 && (!easting || !northing || !upward))
        return hmx::fail(HMX_ERR_ARG, "NULL coordinate array");
    if (n_prisms > 0 && (!prisms || !density)) return hmx::fail(HMX_ERR_ARG, "NULL prism array");
    hmx::ExplicitSource src{prisms, density};
    return hmx::run_forward(easting, northing, upward, n_obs, src, n_prisms, field_mask, out,
                            singular_flags, workspace, workspace_bytes, (hipStream_t)stream);
}

int hmx_prism_singular_scan(const double *easting, const double *northing, const double *upward,
                            int64_t n_obs, const double *prisms, int64_t n_prisms,
                            uint32_t field_mask, uint32_t *flags, void *stream) {
    hmx::clear_error();
    int st = hmx::check_common(n_obs, n_prisms, field_mask, flags);
    if (st) return st;
    if (n_obs == 0 || n_prisms == 0) return HMX_OK;
    if (!easting || !northing || !upward || !prisms)
        return hmx::fail(HMX_ERR_ARG, "NULL input array");
    int cu = hmx::device_cu_count();
    if (cu < 0) return cu;
    const int64_t chunks = hmx::forward_chunks(n_obs, n_prisms, cu);
    const int64_t chunk_len = (n_prisms + chunks - 1) / chunks;
    dim3 grid((unsigned)((n_obs + hmx::kObsPerBlock - 1) / hmx::kObsPerBlock), (unsigned)chunks);
    hipLaunchKernelGGL(hmx::singular_scan_kernel, grid, dim3(hmx::kObsPerBlock), 0,
                       (hipStream_t)stream, easting, northing, upward, (int)n_obs, prisms,
                       n_prisms, chunk_len, field_mask, flags);
    HMX_LAUNCHED("singular_scan_kernel");
    return HMX_OK;
}

int hmx_prism_layer_gravity(const double *easting, const double *northing, const double *upward,
                            int64_t n_obs, const double *centre_easting, int64_t n_east,
                            const double *centre_northing, int64_t n_north,
                            const double *bottom, const double *top, const double *density,
                            double thickness_threshold, uint32_t field_mask, double *out,
                            uint32_t *singular_flags, void *workspace, size_t workspace_bytes,
                            void *stream) {
    hmx::clear_error();
    if (n_east < 2 || n_north < 2)
        return hmx::fail(HMX_ERR_ARG, "a prism layer needs at least 2 centres per direction");
    int st = hmx::check_common(n_obs, n_east * n_north, field_mask, out);
    if (st) return st;
    if (n_obs > 0 && (!easting || !northing || !upward))
        return hmx::fail(HMX_ERR_ARG, "NULL coordinate array");
    if (!centre_easting || !centre_northing || !bottom || !top || !density)
        return hmx::fail(HMX_ERR_ARG, "NULL layer array");
    hmx::LayerSource src{centre_easting, centre_northing, bottom, top, density,
                         (int)n_east, (int)n_north, thickness_threshold};
    return hmx::run_forward(easting, northing, upward, n_obs, src, n_east * n_north, field_mask,
                            out, singular_flags, workspace, workspace_bytes, (hipStream_t)stream);
}

}  // extern "C"
