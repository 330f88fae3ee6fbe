// Many independent g_z prism models of uneven sizes in one persistent
// launch (no counterpart in the reference, where each model is a separate
// prism_gravity call on the Numba thread pool).
//
// Plan (host): every model b with n_b points and m_b prisms is cut into
// items (b, observation tile of 256 points, prism chunk of at most
// prisms_per_item prisms); item slots are numbered model-major, tile,
// chunk; the processing order sorts items by decreasing cost so the
// longest ones start first.
// Kernel: a grid sized to the device's resident capacity; each workgroup
// takes the next item from an atomic counter until the list is exhausted
// (every workgroup reaches the exit), evaluates its tile against its chunk
// (prisms staged in LDS, FP64 closed form) and stores 256 partial sums at
// the item's slot.
// Reduce: each point sums its model's chunk partials in chunk order
// (deterministic), applies harmonica's sign and mGal.
#include <algorithm>
#include <numeric>
#include <vector>

#include "hmx_internal.h"
#include "prism_device.h"

namespace hmx {
namespace {

constexpr int kItemWords = 5;  // model, tile, p0, p1, slot

struct alignas(16) BPrism {
    double w, e, s, n, b, t, grho;
    int skip, share;  // never set: batched models are not marked for edge sharing
};

static_assert(kTile <= 128, "tile_pairs marks at most 128 prisms");

__global__ __launch_bounds__(kObsPerBlock) void batched_gz_kernel(
    const double *__restrict__ oe, const double *__restrict__ on, const double *__restrict__ ou,
    const int64_t *__restrict__ obs_off, const double *__restrict__ prisms,
    const double *__restrict__ density, const int64_t *__restrict__ items, int64_t n_items,
    double *__restrict__ partial, unsigned long long *__restrict__ counter) {
    __shared__ BPrism tile[kTile];
    __shared__ dm::Tables tables;
    __shared__ long long item_id;
    dm::load_tables(&tables);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) item_id = (long long)atomicAdd(counter, 1ull);
        __syncthreads();
        const long long k = item_id;
        if (k >= n_items) break;
        const int64_t *it = items + kItemWords * k;
        const int64_t b = it[0], T = it[1], p0 = it[2], p1 = it[3], slot = it[4];
        const int64_t first = obs_off[b], end = obs_off[b + 1];
        const int64_t i = first + T * kObsPerBlock + threadIdx.x;
        const bool valid = i < end;
        const int64_t ii = valid ? i : end - 1;
        const double e = oe[ii], n = on[ii], u = ou[ii];
        double acc[1] = {0.0};
        uint32_t sflags = 0;
        for (int64_t t0 = p0; t0 < p1; t0 += kTile) {
            const int cnt = (int)(p1 - t0 < kTile ? p1 - t0 : kTile);
            __syncthreads();
            double *flat = reinterpret_cast<double *>(tile);
            for (int idx = threadIdx.x; idx < 6 * cnt; idx += blockDim.x) {
                const int q = idx / 6, c = idx - 6 * q;
                flat[8 * q + c] = prisms[6 * t0 + idx];
            }
            for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
                tile[q].grho = dev::kG * density[t0 + q];
                tile[q].skip = 0;
                tile[q].share = 0;
            }
            __syncthreads();
            dev::tile_pairs<0x8u, false>(tile, cnt, e, n, u, acc, sflags, tables);
        }
        partial[slot * kObsPerBlock + threadIdx.x] = valid ? acc[0] : 0.0;
    }
}

__global__ void batched_reduce_kernel(const double *__restrict__ partial,
                                      const int64_t *__restrict__ obs_off, int64_t n_models,
                                      const int64_t *__restrict__ item_base,
                                      const int64_t *__restrict__ chunks, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= obs_off[n_models]) return;
    int64_t lo = 0, hi = n_models;  // find b with obs_off[b] <= i < obs_off[b + 1]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (obs_off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    const int64_t b = lo, local = i - obs_off[b];
    const int64_t T = local / kObsPerBlock, t = local - T * kObsPerBlock;
    const int64_t C = chunks[b], base = item_base[b] + T * C;
    double sum = 0.0;
    for (int64_t c = 0; c < C; ++c) sum += partial[(base + c) * kObsPerBlock + t];
    out[i] = dev::to_harmonica_units(3, sum);
}

__global__ void zero_counter_kernel(unsigned long long *c) {
    if (threadIdx.x == 0) *c = 0ull;
}

}  // namespace
}  // namespace hmx

extern "C" {

int hmx_batched_plan_size(const int64_t *obs_offsets, const int64_t *prism_offsets,
                          int64_t n_models, int64_t prisms_per_item, int64_t *n_items) {
    using namespace hmx;
    clear_error();
    if (!obs_offsets || !prism_offsets || !n_items) return fail(HMX_ERR_ARG, "NULL array");
    if (n_models <= 0 || prisms_per_item <= 0) return fail(HMX_ERR_ARG, "invalid sizes");
    int64_t total = 0;
    for (int64_t b = 0; b < n_models; ++b) {
        const int64_t nb = obs_offsets[b + 1] - obs_offsets[b];
        const int64_t mb = prism_offsets[b + 1] - prism_offsets[b];
        if (nb < 0 || mb < 0) return fail(HMX_ERR_ARG, "offsets not increasing at model %lld", (long long)b);
        const int64_t tiles = (nb + kObsPerBlock - 1) / kObsPerBlock;
        const int64_t ch = mb > 0 ? (mb + prisms_per_item - 1) / prisms_per_item : 0;
        total += tiles * ch;
    }
    *n_items = total;
    return HMX_OK;
}

int hmx_batched_plan(const int64_t *obs_offsets, const int64_t *prism_offsets, int64_t n_models,
                     int64_t prisms_per_item, int64_t *items, int64_t capacity, int64_t *n_items,
                     int64_t *item_base, int64_t *chunks) {
    using namespace hmx;
    int64_t total = 0;
    int st = hmx_batched_plan_size(obs_offsets, prism_offsets, n_models, prisms_per_item, &total);
    if (st) return st;
    if (!items || !item_base || !chunks) return fail(HMX_ERR_ARG, "NULL array");
    if (capacity < total) return fail(HMX_ERR_ARG, "item capacity %lld < %lld", (long long)capacity,
                                      (long long)total);
    std::vector<int64_t> tmp((size_t)total * kItemWords);
    std::vector<int64_t> cost((size_t)total);
    int64_t slot = 0;
    for (int64_t b = 0; b < n_models; ++b) {
        const int64_t nb = obs_offsets[b + 1] - obs_offsets[b];
        const int64_t mb = prism_offsets[b + 1] - prism_offsets[b];
        const int64_t tiles = (nb + kObsPerBlock - 1) / kObsPerBlock;
        const int64_t ch = mb > 0 ? (mb + prisms_per_item - 1) / prisms_per_item : 0;
        item_base[b] = slot;
        chunks[b] = ch;
        for (int64_t T = 0; T < tiles; ++T) {
            const int64_t pts = std::min<int64_t>(kObsPerBlock, nb - T * kObsPerBlock);
            for (int64_t c = 0; c < ch; ++c) {
                const int64_t p0 = prism_offsets[b] + c * prisms_per_item;
                const int64_t p1 = std::min(p0 + prisms_per_item, prism_offsets[b + 1]);
                int64_t *it = tmp.data() + kItemWords * slot;
                it[0] = b;
                it[1] = T;
                it[2] = p0;
                it[3] = p1;
                it[4] = slot;
                // a workgroup costs its prism count whatever its valid points
                cost[(size_t)slot] = (p1 - p0) * (pts > 0 ? 1 : 0);
                ++slot;
            }
        }
    }
    std::vector<int64_t> order((size_t)total);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b2) { return cost[(size_t)a] > cost[(size_t)b2]; });
    for (int64_t k = 0; k < total; ++k)
        std::copy_n(tmp.data() + kItemWords * order[(size_t)k], kItemWords, items + kItemWords * k);
    *n_items = total;
    return HMX_OK;
}

int hmx_batched_workspace_bytes(int64_t n_items, size_t *bytes) {
    using namespace hmx;
    if (!bytes) return fail(HMX_ERR_ARG, "bytes is NULL");
    if (n_items < 0) return fail(HMX_ERR_ARG, "negative size");
    *bytes = 256 + (size_t)n_items * kObsPerBlock * sizeof(double);
    return HMX_OK;
}

int hmx_batched_prism_gz(const double *easting, const double *northing, const double *upward,
                         const int64_t *obs_offsets, const double *prisms, const double *density,
                         const int64_t *items, int64_t n_items, const int64_t *item_base,
                         const int64_t *chunks, int64_t n_models, int64_t n_obs_total,
                         int blocks_per_cu, double *out, void *workspace, size_t workspace_bytes,
                         void *stream) {
    using namespace hmx;
    clear_error();
    if (n_models <= 0 || n_items < 0 || n_obs_total < 0) return fail(HMX_ERR_ARG, "invalid sizes");
    if (blocks_per_cu <= 0 || blocks_per_cu > 64) return fail(HMX_ERR_ARG, "blocks_per_cu out of range");
    if (!easting || !northing || !upward || !obs_offsets || !out || !item_base || !chunks)
        return fail(HMX_ERR_ARG, "NULL array");
    if (n_items > 0 && (!prisms || !density || !items)) return fail(HMX_ERR_ARG, "NULL array");
    size_t need = 0;
    int st = hmx_batched_workspace_bytes(n_items, &need);
    if (st) return st;
    if (!workspace || workspace_bytes < need)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    const int cu = device_cu_count();
    if (cu < 0) return cu;
    hipStream_t s = (hipStream_t)stream;
    auto *counter = static_cast<unsigned long long *>(workspace);
    double *partial = reinterpret_cast<double *>(static_cast<char *>(workspace) + 256);
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, s, counter);
    HMX_LAUNCHED("zero_counter_kernel");
    if (n_items > 0) {
        const int64_t grid64 = std::min<int64_t>((int64_t)cu * blocks_per_cu, n_items);
        hipLaunchKernelGGL(batched_gz_kernel, dim3((unsigned)grid64), dim3(kObsPerBlock), 0, s,
                           easting, northing, upward, obs_offsets, prisms, density, items, n_items,
                           partial, counter);
        HMX_LAUNCHED("batched_gz_kernel");
    }
    if (n_obs_total > 0) {
        hipLaunchKernelGGL(batched_reduce_kernel, dim3((unsigned)((n_obs_total + 255) / 256)),
                           dim3(256), 0, s, partial, obs_offsets, n_models, item_base, chunks, out);
        HMX_LAUNCHED("batched_reduce_kernel");
    }
    return HMX_OK;
}

}  // extern "C"
