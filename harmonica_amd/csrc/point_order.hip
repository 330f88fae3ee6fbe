// Wavefront-coherent order of the observation points of the pairwise prism
// kernels.  A prism kernel picks, per wavefront and prism, the cheapest
// closed form that is valid for all 64 points of the wavefront (lean, x-only
// or general), so a wavefront whose points are spread over the survey area
// pays for the general form on most prisms.  The points are therefore
// evaluated in boustrophedon strip order: the bounding box of the points is
// cut into ns strips along northing and nc cells per strip along easting
// (cells 4 times taller than wide, about n / 4 cells), strips are walked
// alternately east- and westward, and a stable LSD radix sort of the cell
// indices (8-bit digits) gives the permutation; points of one cell keep the
// caller's order.  64 consecutive points then cover about 16 cells of one
// strip, a patch about 4 times wider than tall.  The kernels read gathered
// copies of the coordinates and the reduction writes each value back to the
// caller's position, so results are independent of the order up to the
// rounding of the chosen forms, and deterministic.
#include <cmath>

#include "hmx_internal.h"

namespace hmx {
namespace {

constexpr int kSortThreads = 256;  // 4 wavefronts
constexpr int kSortItems = 8;      // keys per thread
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kDigitBits = 8;
constexpr int kBins = 1 << kDigitBits;
constexpr int kScanThreads = 1024;
constexpr int kBoundsThreads = 1024;
static_assert(kSortThreads == kBins, "one thread per digit in the sort kernels");

struct OrderParams {
    double emin, nmin, inv_w, inv_h;
    int64_t ns, nc;
};

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

inline int64_t order_cells(int64_t n_obs) { return n_obs / 4 > 1 ? n_obs / 4 : 1; }

inline int64_t sort_blocks(int64_t n_obs) { return (n_obs + kSortTile - 1) / kSortTile; }

// Byte offsets of the order region's parts.
struct OrderLayout {
    size_t params, keys0, keys1, vals0, vals1, hist, e, n, u, total;
    explicit OrderLayout(int64_t n_obs) {
        const size_t n = (size_t)n_obs;
        size_t at = 0;
        auto take = [&](size_t bytes) {
            const size_t off = at;
            at += align256(bytes);
            return off;
        };
        params = take(sizeof(OrderParams));
        keys0 = take(n * sizeof(uint32_t));
        keys1 = take(n * sizeof(uint32_t));
        vals0 = take(n * sizeof(int));
        vals1 = take(n * sizeof(int));
        hist = take((size_t)kBins * (size_t)sort_blocks(n_obs) * sizeof(uint32_t));
        e = take(n * sizeof(double));
        this->n = take(n * sizeof(double));
        u = take(n * sizeof(double));
        total = at;
    }
};

// Bounding box of the points (NaN coordinates ignored) and the strip and
// cell counts: ns * nc <= cells, cells of height h = span_n / ns and width
// w = span_e / nc with h ~ 4 w.  One workgroup.
__global__ __launch_bounds__(kBoundsThreads) void order_bounds_kernel(
    const double *__restrict__ e, const double *__restrict__ n, int64_t n_obs, int64_t cells,
    OrderParams *__restrict__ par) {
    __shared__ double red[4][kBoundsThreads];
    double emin = INFINITY, emax = -INFINITY, nmin = INFINITY, nmax = -INFINITY;
    for (int64_t i = threadIdx.x; i < n_obs; i += kBoundsThreads) {
        const double x = e[i], y = n[i];
        emin = fmin(emin, x);
        emax = fmax(emax, x);
        nmin = fmin(nmin, y);
        nmax = fmax(nmax, y);
    }
    red[0][threadIdx.x] = emin;
    red[1][threadIdx.x] = -emax;
    red[2][threadIdx.x] = nmin;
    red[3][threadIdx.x] = -nmax;
    __syncthreads();
    for (int off = kBoundsThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    emin = red[0][0];
    emax = -red[1][0];
    nmin = red[2][0];
    nmax = -red[3][0];
    const double span_e = emax - emin, span_n = nmax - nmin;
    int64_t ns = 1, nc = 1;
    if (std::isfinite(span_e) && std::isfinite(span_n)) {
        if (span_e > 0.0 && span_n > 0.0) {
            const double r = floor(sqrt((double)cells * span_n / (4.0 * span_e)));
            ns = r < 1.0 ? 1 : (r > (double)cells ? cells : (int64_t)r);
            nc = cells / ns > 1 ? cells / ns : 1;
        } else if (span_n > 0.0) {
            ns = cells;
        } else if (span_e > 0.0) {
            nc = cells;
        }
    }
    par->emin = std::isfinite(emin) ? emin : 0.0;
    par->nmin = std::isfinite(nmin) ? nmin : 0.0;
    par->inv_w = nc > 1 ? (double)nc / span_e : 0.0;
    par->inv_h = ns > 1 ? (double)ns / span_n : 0.0;
    par->ns = ns;
    par->nc = nc;
}

// Cell index in [0, cnt) of v on a row of cnt cells starting at lo (0 for
// NaN, clamped at the ends).
__device__ __forceinline__ int64_t cell_of(double v, double lo, double inv, int64_t cnt) {
    const double f = floor((v - lo) * inv);
    return f >= 0.0 ? (f < (double)cnt ? (int64_t)f : cnt - 1) : 0;
}

// keys[i] = boustrophedon cell index of point i, vals[i] = i.
__global__ void order_keys_kernel(const double *__restrict__ e, const double *__restrict__ n,
                                  int64_t n_obs, const OrderParams *__restrict__ par,
                                  uint32_t *__restrict__ keys, int *__restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_obs) return;
    const OrderParams p = *par;
    const int64_t s = cell_of(n[i], p.nmin, p.inv_h, p.ns);
    int64_t c = cell_of(e[i], p.emin, p.inv_w, p.nc);
    if (s & 1) c = p.nc - 1 - c;
    keys[i] = (uint32_t)(s * p.nc + c);
    vals[i] = (int)i;
}

// hist[d * nblocks + b] = number of keys of tile b whose digit at `shift`
// is d.
__global__ __launch_bounds__(kSortThreads) void sort_hist_kernel(const uint32_t *__restrict__ keys,
                                                               int64_t n, int shift,
                                                               uint32_t *__restrict__ hist,
                                                               int64_t nblocks) {
    __shared__ uint32_t cnt[kBins];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = base + (int64_t)j * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kBins - 1)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive prefix sum of hist[0, total) in place.  One workgroup: each
// thread sums a contiguous run, the run totals are scanned in LDS, then each
// thread rewrites its run.
__global__ __launch_bounds__(kScanThreads) void sort_scan_kernel(uint32_t *__restrict__ hist,
                                                               int64_t total) {
    __shared__ uint32_t part[kScanThreads];
    const int64_t per = (total + kScanThreads - 1) / kScanThreads;
    const int64_t b = (int64_t)threadIdx.x * per;
    const int64_t e = b + per < total ? b + per : total;
    uint32_t s = 0;
    for (int64_t i = b; i < e; ++i) s += hist[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t v = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (int64_t i = b; i < e; ++i) {
        const uint32_t v = hist[i];
        hist[i] = run;
        run += v;
    }
}

// Stable scatter of one radix pass.  Keys of tile b are taken in index
// order (slot j, wavefront, lane); a key's destination is the scanned
// histogram entry of its digit and tile, plus the keys of the same digit
// before it in the tile: those in earlier (slot, wavefront) groups (counted
// per group in LDS and prefix-summed per digit) and those on lower lanes of
// its own group (the lanes that agree with it on all 8 digit bits, from 8
// ballots).
__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(
    const uint32_t *__restrict__ keys_in, const int *__restrict__ vals_in, int64_t n, int shift,
    const uint32_t *__restrict__ hist, int64_t nblocks, uint32_t *__restrict__ keys_out,
    int *__restrict__ vals_out) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t cnt[kSortItems][kWaves][kBins];
    for (int k = threadIdx.x; k < kSortItems * kWaves * kBins; k += kSortThreads)
        (&cnt[0][0][0])[k] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    uint32_t key[kSortItems];
    int val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = base + (int64_t)j * kSortThreads + threadIdx.x;
        const bool ok = i < n;
        key[j] = ok ? keys_in[i] : 0u;
        val[j] = ok ? vals_in[i] : 0;
        const uint32_t d = (key[j] >> shift) & (kBins - 1);
        uint64_t same = __ballot(ok);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t on = __ballot(bit);
            same &= bit ? on : ~on;
        }
        rank[j] = ok ? __popcll(same & below) : -1;
        if (ok && rank[j] == 0) cnt[j][wave][d] = (uint32_t)__popcll(same);
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        uint32_t run = 0u;
        for (int j = 0; j < kSortItems; ++j)
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = cnt[j][w][d];
                cnt[j][w][d] = run;
                run += c;
            }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        if (rank[j] < 0) continue;
        const uint32_t d = (key[j] >> shift) & (kBins - 1);
        const uint32_t pos = hist[(int64_t)d * nblocks + blockIdx.x] + cnt[j][wave][d] + (uint32_t)rank[j];
        keys_out[pos] = key[j];
        vals_out[pos] = val[j];
    }
}

// oe[i] = e[perm[i]] (and n, u).
__global__ void order_gather_kernel(const double *__restrict__ e, const double *__restrict__ n,
                                    const double *__restrict__ u, const int *__restrict__ perm,
                                    int64_t n_obs, double *__restrict__ oe,
                                    double *__restrict__ on, double *__restrict__ ou) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_obs) return;
    const int64_t k = perm[i];
    oe[i] = e[k];
    on[i] = n[k];
    ou[i] = u[k];
}

}  // namespace

size_t point_order_bytes(int64_t n_obs) { return OrderLayout(n_obs).total; }

int point_order(const double *e, const double *n, const double *u, int64_t n_obs, void *ws,
                PointOrder *po, hipStream_t s) {
    const OrderLayout L(n_obs);
    char *base = static_cast<char *>(ws);
    auto *par = reinterpret_cast<OrderParams *>(base + L.params);
    uint32_t *keys[2] = {reinterpret_cast<uint32_t *>(base + L.keys0),
                         reinterpret_cast<uint32_t *>(base + L.keys1)};
    int *vals[2] = {reinterpret_cast<int *>(base + L.vals0), reinterpret_cast<int *>(base + L.vals1)};
    auto *hist = reinterpret_cast<uint32_t *>(base + L.hist);
    auto *oe = reinterpret_cast<double *>(base + L.e);
    auto *on = reinterpret_cast<double *>(base + L.n);
    auto *ou = reinterpret_cast<double *>(base + L.u);
    const int64_t cells = order_cells(n_obs), nblocks = sort_blocks(n_obs);
    int bits = 0;
    while (bits < 32 && ((uint64_t)(cells - 1) >> bits) != 0) ++bits;
    const int passes = (bits + kDigitBits - 1) / kDigitBits;
    const unsigned pt_blocks = (unsigned)((n_obs + 255) / 256);
    hipLaunchKernelGGL(order_bounds_kernel, dim3(1), dim3(kBoundsThreads), 0, s, e, n, n_obs, cells,
                       par);
    HMX_LAUNCHED("order_bounds_kernel");
    hipLaunchKernelGGL(order_keys_kernel, dim3(pt_blocks), dim3(256), 0, s, e, n, n_obs, par,
                       keys[0], vals[0]);
    HMX_LAUNCHED("order_keys_kernel");
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kDigitBits;
        hipLaunchKernelGGL(sort_hist_kernel, dim3((unsigned)nblocks), dim3(kSortThreads), 0, s,
                           keys[cur], n_obs, shift, hist, nblocks);
        HMX_LAUNCHED("sort_hist_kernel");
        hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, hist,
                           (int64_t)kBins * nblocks);
        HMX_LAUNCHED("sort_scan_kernel");
        hipLaunchKernelGGL(sort_scatter_kernel, dim3((unsigned)nblocks), dim3(kSortThreads), 0, s,
                           keys[cur], vals[cur], n_obs, shift, hist, nblocks, keys[cur ^ 1],
                           vals[cur ^ 1]);
        HMX_LAUNCHED("sort_scatter_kernel");
        cur ^= 1;
    }
    hipLaunchKernelGGL(order_gather_kernel, dim3(pt_blocks), dim3(256), 0, s, e, n, u, vals[cur],
                       n_obs, oe, on, ou);
    HMX_LAUNCHED("order_gather_kernel");
    po->e = oe;
    po->n = on;
    po->u = ou;
    po->perm = vals[cur];
    return HMX_OK;
}

}  // namespace hmx
