// Internal helpers shared by the libhmx translation units: error reporting,
// device queries and the launch-shape rules of the pairwise kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/hmx.h"

namespace hmx {

// Record a printf-style message for hmx_last_error() and return `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Clear the calling thread's last error.
void clear_error();

// Devices the per-device caches of the library cover.
constexpr int kMaxDevices = 64;

// Number of compute units of the current device (cached per device), or a
// negative status if no device is usable.
int device_cu_count();

}  // namespace hmx

#define HMX_HIP(expr)                                                              \
    do {                                                                           \
        hipError_t hmx_e_ = (expr);                                                \
        if (hmx_e_ != hipSuccess)                                                  \
            return ::hmx::fail(HMX_ERR_HIP, "%s failed: %s", #expr,                \
                               hipGetErrorString(hmx_e_));                         \
    } while (0)

// After a kernel launch: report launch-configuration errors.
#define HMX_LAUNCHED(name)                                                         \
    do {                                                                           \
        hipError_t hmx_e_ = hipGetLastError();                                     \
        if (hmx_e_ != hipSuccess)                                                  \
            return ::hmx::fail(HMX_ERR_HIP, "launch of %s failed: %s", name,       \
                               hipGetErrorString(hmx_e_));                         \
    } while (0)

namespace hmx {

// Observation points per workgroup of the pairwise kernels: 4 wavefronts of
// 64 lanes, one observation point per lane.
constexpr int kObsPerBlock = 256;
// Sources staged in LDS per pass.
constexpr int kTile = 128;

// Split of the source dimension into chunks so that a launch has enough
// workgroups to fill all compute units several times over.  Each chunk's
// partial sums go to the workspace and are reduced in chunk order, so the
// result does not depend on scheduling.
inline int64_t forward_chunks(int64_t n_obs, int64_t n_src, int n_cu) {
    int64_t obs_blocks = (n_obs + kObsPerBlock - 1) / kObsPerBlock;
    if (obs_blocks < 1) obs_blocks = 1;
    int64_t target = 40LL * (n_cu > 0 ? n_cu : 256);
    int64_t chunks = (target + obs_blocks - 1) / obs_blocks;
    int64_t max_by_src = (n_src + kTile - 1) / kTile;
    if (chunks > max_by_src) chunks = max_by_src;
    if (chunks > 1024) chunks = 1024;
    if (chunks < 1) chunks = 1;
    return chunks;
}

inline int popcount10(uint32_t m) { return __builtin_popcount(m & HMX_ALL_FIELDS); }

// Wavefront-coherent point order of the prism kernels (point_order.hip):
// gathered copies of the coordinates in boustrophedon strip order and the
// permutation perm (sorted position -> caller's index) for the reduction.
struct PointOrder {
    const double *e, *n, *u;
    const int *perm;
};

// Whether a prism call sorts its points: enough points for the strips to
// matter and enough pairs that the sort's dozen small launches are
// negligible next to the pair loop.
inline bool use_point_order(int64_t n_obs, int64_t n_src) {
    // n_obs * n_src >= 2^30, without forming the product
    return n_obs >= 2048 && n_src >= ((((int64_t)1 << 30) + n_obs - 1) / n_obs);
}

// Bytes of the order's workspace region for n_obs points.
size_t point_order_bytes(int64_t n_obs);

// Sort the points into the region `ws` (point_order_bytes(n_obs) bytes,
// 256-byte aligned) on stream s.
int point_order(const double *e, const double *n, const double *u, int64_t n_obs, void *ws,
                PointOrder *po, hipStream_t s);

// Offset of the order region in a forward workspace whose partial sums take
// partial_bytes.
inline size_t point_order_offset(size_t partial_bytes) { return (partial_bytes + 255) & ~(size_t)255; }

// Least squares min |G c - y|^2 (+ alpha |c|^2, per filter) through the
// one-sided Jacobi SVD (lsq_svd.hip) of A (L x k, column-major, L >= k,
// overwritten): A = G (n x m, transposed = 0) or A = G^T (transposed = 1).
// W (k x k), vec (3 k doubles), counters (2 ints); y has n entries and
// coef m.  Returns the sweeps taken (negative if max_sweeps stopped them)
// and the number of kept components.  Synchronous.
int svd_lsq(double *A, int64_t L, int64_t k, int transposed, const double *y, double alpha,
            int filter, double cutoff, double *W, double *vec, int *counters, double *coef,
            int max_sweeps, int *sweeps, int *rank, hipStream_t s);
constexpr int kSvdMaxSweeps = 60;

// One window of the gradient-boosted fit, gathered on the device: its data
// (w may be NULL) and its sources.
struct EqsWindow {
    const double *e, *n, *u, *w;
    int64_t nd;
    const double *pe, *pn, *pu;
    int64_t m;
};
// Bytes of the buffers of a batch of `batch` windows of at most n_max data
// and m_max sources (FP64 mode, damping > 0; eqs_fit.hip).
size_t eqs_batch_bytes(int64_t n_max, int64_t m_max, int64_t batch);
// Forms and factorises the normal matrices of a batch of windows; status
// (host, batch ints) receives 0 or the 1-based column of a failed pivot,
// seconds (host, 2) the times of the normal matrices and of the factors.
int eqs_factor_batch(const EqsWindow *wins, int64_t batch, int64_t n_max, int64_t m_max,
                     double damping, void *buffer, size_t buffer_bytes, int *status,
                     double *seconds, hipStream_t s);
// The coefficients of window b of a factorised batch for the data `data`
// (the window's residue): right-hand side, solve and CG refinement as
// hmx_eqs_fit; info (host, 11 doubles) as hmx_eqs_fit's.
int eqs_solve_factored(const EqsWindow &w, const double *data, double damping, void *buffer,
                       int64_t b, int64_t batch, int64_t n_max, int64_t m_max, int max_iter,
                       double rtol, double *coefs, double *info, hipStream_t s);
// Largest fits that take the SVD route when the Cholesky route cannot match
// the reference's solver: at most this many sources and matrix entries.
constexpr int64_t kSvdMaxSrc = 8192;
constexpr int64_t kSvdMaxElems = (int64_t)1 << 27;

}  // namespace hmx
