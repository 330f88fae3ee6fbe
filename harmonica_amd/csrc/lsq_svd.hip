// Least squares through the singular value decomposition, for the
// least-squares problems that the Cholesky route of the equivalent-source
// fit cannot solve the way the reference's solvers do.
//
// The reference fits through verde.least_squares
// (src/harmonica/_equivalent_sources/cartesian.py:280), i.e. scikit-learn:
//   * LinearRegression (damping None): scipy.linalg.lstsq, the
//     minimum-norm solution with singular values s <= cond * s_max dropped
//     (cond = max(n, m) * 2^-52);
//   * Ridge (damping alpha): a Cholesky solve of X^T X + alpha I; when it
//     fails, c = V diag(s / (s^2 + alpha)) U^T y over the singular values
//     s > 1e-15 (sklearn/linear_model/_ridge.py, _solve_svd); with more
//     columns than rows the dual form, whose fallback keeps the components
//     with s^2 + alpha > 2^-52 (s_max^2 + alpha) (scipy.linalg.lstsq of
//     X X^T + alpha I).
// Every one of these is c = V diag(f(s)) U^T y for a filter f of the
// singular values of G (n x m), computed here from the one-sided
// (Hestenes) Jacobi SVD of A = G (n >= m) or A = G^T (m > n), so that A
// has at least as many rows L as columns k: pairs of columns of A are
// rotated until all are mutually orthogonal, A W = Z diag(s), with W
// (k x k) accumulated from the identity.  With a_j = s_j z_j and
// t_j = d_j f(s_j) / s_j,
//   n >= m (W = V, Z = U):  d_j = a_j . y = s_j (U^T y)_j,  c = W t,
//   m > n  (W = U, Z = V):  d_j = w_j . y = (U^T y)_j,      c = A t.
// One-sided Jacobi is backward stable
// and, for matrices that are well conditioned up to a scaling of their
// columns, accurate in the relative sense even for small singular values.
//
// Parallel order: the round-robin (circle) schedule; round t of a sweep
// pairs the K (= k rounded up to even) column slots so that every slot
// occurs once, and the K - 1 rounds of a sweep pair every two columns once.
// One workgroup rotates one pair.  A sweep without rotations ends the
// iteration.
#include <cmath>

#include "hmx_internal.h"

namespace hmx {
namespace {

constexpr int kJT = 256;  // threads per column pair / column

// Slots of pair k in round t of the circle schedule over M slots: slot 0
// stays, slots 1..M-1 advance by one position per round.
__device__ __forceinline__ void circle_pair(int64_t M, int64_t t, int64_t k, int64_t &p,
                                            int64_t &q) {
    const int64_t r = M - 1;
    p = k == 0 ? 0 : (k - 1 + t) % r + 1;
    const int64_t kq = M - 1 - k;
    q = (kq - 1 + t) % r + 1;
}

// Workgroup sums of three values in a fixed order; every thread gets them.
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c) {
    __shared__ double red[3][kJT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[0][wave] = a;
        red[1][wave] = b;
        red[2][wave] = c;
    }
    __syncthreads();
    a = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    b = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    c = (red[2][0] + red[2][1]) + (red[2][2] + red[2][3]);
}

__global__ void eye_kernel(double *__restrict__ V, int64_t m) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m * m) V[k] = (k / m == k % m) ? 1.0 : 0.0;
}

__global__ void zero_counters_kernel(int *c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

// Round t: workgroup k orthogonalises the columns of its pair.  With
// a = |g_p|^2, b = |g_q|^2, c = g_p . g_q, zeta = (b - a) / (2 c) and
// tn the root of tn^2 + 2 zeta tn - 1 = 0 of smaller magnitude, the
// rotation cs = 1 / sqrt(1 + tn^2), sn = cs tn makes
//   g_p' = cs g_p - sn g_q,   g_q' = sn g_p + cs g_q
// orthogonal (the same rotation is applied to columns p, q of V).  Pairs
// with |c| <= tol sqrt(a) sqrt(b) are left alone.
__global__ __launch_bounds__(kJT) void jacobi_rotate_kernel(double *__restrict__ G, int64_t n,
                                                            double *__restrict__ V, int64_t m,
                                                            int64_t M, int64_t t, double tol,
                                                            int *__restrict__ rotations) {
    int64_t p, q;
    circle_pair(M, t, blockIdx.x, p, q);
    if (p >= m || q >= m) return;  // the padding slot of an odd m
    if (p > q) {
        const int64_t tmp = p;
        p = q;
        q = tmp;
    }
    double *gp = G + p * n, *gq = G + q * n;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kJT) {
        const double x = gp[i], y = gq[i];
        a = fma(x, x, a);
        b = fma(y, y, b);
        c = fma(x, y, c);
    }
    block_sum3(a, b, c);
    if (!(fabs(c) > tol * sqrt(a) * sqrt(b))) return;
    const double zeta = (b - a) / (2.0 * c);
    const double az = fabs(zeta);
    const double tn = az > 1e150 ? 0.5 / zeta : copysign(1.0, zeta) / (az + sqrt(fma(az, az, 1.0)));
    const double cs = 1.0 / sqrt(fma(tn, tn, 1.0)), sn = cs * tn;
    for (int64_t i = threadIdx.x; i < n; i += kJT) {
        const double x = gp[i], y = gq[i];
        gp[i] = fma(cs, x, -sn * y);
        gq[i] = fma(sn, x, cs * y);
    }
    double *vp = V + p * m, *vq = V + q * m;
    for (int64_t i = threadIdx.x; i < m; i += kJT) {
        const double x = vp[i], y = vq[i];
        vp[i] = fma(cs, x, -sn * y);
        vq[i] = fma(sn, x, cs * y);
    }
    if (threadIdx.x == 0) atomicAdd(rotations, 1);
}

// sigma_j = |g_j|, dots_j = g_j . y (y == NULL: 0; one workgroup per
// column).
__global__ __launch_bounds__(kJT) void col_norm_dot_kernel(const double *__restrict__ G,
                                                           int64_t n, const double *__restrict__ y,
                                                           double *__restrict__ sigma,
                                                           double *__restrict__ dots) {
    const double *g = G + (int64_t)blockIdx.x * n;
    double a = 0.0, c = 0.0, unused = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kJT) {
        const double x = g[i];
        a = fma(x, x, a);
        if (y) c = fma(x, y[i], c);
    }
    block_sum3(a, c, unused);
    if (threadIdx.x == 0) {
        sigma[blockIdx.x] = sqrt(a);
        dots[blockIdx.x] = c;
    }
}

// t_j = dots_j f(sigma_j) / sigma_j for the filter (one workgroup of 1024
// threads; the largest singular value first, in a fixed order); counts the
// kept components in *rank.
__global__ __launch_bounds__(1024) void svd_filter_kernel(const double *__restrict__ sigma,
                                                          const double *__restrict__ dots,
                                                          int64_t m, double alpha, int filter,
                                                          double cutoff, double *__restrict__ t,
                                                          int *__restrict__ rank) {
    __shared__ double red[1024];
    __shared__ int cnt[1024];
    double mx = 0.0;
    for (int64_t j = threadIdx.x; j < m; j += 1024) mx = fmax(mx, sigma[j]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double smax = red[0];
    int kept = 0;
    for (int64_t j = threadIdx.x; j < m; j += 1024) {
        const double s = sigma[j];
        double f = 0.0;  // f(s) / s
        if (filter == HMX_SVD_RIDGE) {
            if (s > cutoff) f = 1.0 / fma(s, s, alpha);
        } else if (filter == HMX_SVD_LSTSQ) {
            if (s > cutoff * smax) f = 1.0 / (s * s);
        } else {  // HMX_SVD_DUAL_LSTSQ
            const double lam = fma(s, s, alpha);
            if (s > 0.0 && lam > cutoff * fma(smax, smax, alpha)) f = 1.0 / lam;
        }
        kept += f != 0.0;
        t[j] = dots[j] * f;
    }
    cnt[threadIdx.x] = kept;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) cnt[threadIdx.x] += cnt[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *rank = cnt[0];
}

// out_i = sum_j A[j * rows + i] t_j for i < rows, j < cols (A column-major;
// j in order).
__global__ __launch_bounds__(256) void gemv_cols_kernel(const double *__restrict__ A,
                                                        int64_t rows, int64_t cols,
                                                        const double *__restrict__ t,
                                                        double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= rows) return;
    double acc = 0.0;
    for (int64_t j = 0; j < cols; ++j) acc = fma(A[j * rows + i], t[j], acc);
    out[i] = acc;
}

// The transpose of G (n x m, column-major) as an m x n column-major matrix:
// B[i * m + j] = G[j * n + i], through 32 x 32 tiles in LDS.
__global__ __launch_bounds__(256) void transpose_kernel(const double *__restrict__ G, int64_t n,
                                                        int64_t m, double *__restrict__ B) {
    __shared__ double tile[32][33];
    const int64_t j0 = (int64_t)blockIdx.x * 32, i0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = i0 + tx, j = j0 + r;
        if (i < n && j < m) tile[r][tx] = G[j * n + i];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t j = j0 + tx, i = i0 + r;
        if (i < n && j < m) B[i * m + j] = tile[tx][r];
    }
}

inline unsigned grid_for(int64_t n, int per) {
    const int64_t b = (n + per - 1) / per;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

int svd_lsq(double *A, int64_t L, int64_t k, int transposed, const double *y, double alpha,
            int filter, double cutoff, double *W, double *vec, int *counters, double *coef,
            int max_sweeps, int *sweeps, int *rank, hipStream_t s) {
    if (L <= 0 || k <= 0) return fail(HMX_ERR_ARG, "empty problem");
    if (k > L) return fail(HMX_ERR_ARG, "the Jacobi matrix needs rows >= columns");
    if (filter != HMX_SVD_RIDGE && filter != HMX_SVD_LSTSQ && filter != HMX_SVD_DUAL_LSTSQ)
        return fail(HMX_ERR_ARG, "invalid filter %d", filter);
    double *sigma = vec, *dots = vec + k, *t = vec + 2 * k;
    hipLaunchKernelGGL(eye_kernel, dim3(grid_for(k * k, 256)), dim3(256), 0, s, W, k);
    HMX_LAUNCHED("eye_kernel");
    const int64_t K = k + (k & 1);
    // Orthogonality threshold of a pair: sqrt(L) units of roundoff.
    const double tol = std::sqrt((double)L) * std::ldexp(1.0, -53);
    int done = 0, sw = 0;
    for (sw = 0; sw < max_sweeps && K >= 2; ++sw) {
        hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counters, 1);
        HMX_LAUNCHED("zero_counters_kernel");
        for (int64_t r = 0; r < K - 1; ++r) {
            hipLaunchKernelGGL(jacobi_rotate_kernel, dim3((unsigned)(K / 2)), dim3(kJT), 0, s, A,
                               L, W, k, K, r, tol, counters);
            HMX_LAUNCHED("jacobi_rotate_kernel");
        }
        int rot = 0;
        HMX_HIP(hipMemcpyAsync(&rot, counters, sizeof(int), hipMemcpyDeviceToHost, s));
        HMX_HIP(hipStreamSynchronize(s));
        if (rot == 0) {
            done = 1;
            ++sw;
            break;
        }
    }
    if (K < 2) done = 1;
    *sweeps = done ? sw : -sw;  // negative: stopped at max_sweeps
    if (!transposed) {
        // s_j = |a_j|, dots_j = a_j . y = s_j (U^T y)_j
        hipLaunchKernelGGL(col_norm_dot_kernel, dim3((unsigned)k), dim3(kJT), 0, s, A, L, y,
                           sigma, dots);
        HMX_LAUNCHED("col_norm_dot_kernel");
    } else {
        // s_j = |a_j|; dots_j = w_j . y = (U^T y)_j (the norms of W's
        // columns go to t, overwritten below)
        hipLaunchKernelGGL(col_norm_dot_kernel, dim3((unsigned)k), dim3(kJT), 0, s, A, L,
                           (const double *)nullptr, sigma, dots);
        HMX_LAUNCHED("col_norm_dot_kernel");
        hipLaunchKernelGGL(col_norm_dot_kernel, dim3((unsigned)k), dim3(kJT), 0, s, W, k, y, t,
                           dots);
        HMX_LAUNCHED("col_norm_dot_kernel");
    }
    hipLaunchKernelGGL(svd_filter_kernel, dim3(1), dim3(1024), 0, s, sigma, dots, k, alpha, filter,
                       cutoff, t, counters + 1);
    HMX_LAUNCHED("svd_filter_kernel");
    if (!transposed)
        hipLaunchKernelGGL(gemv_cols_kernel, dim3(grid_for(k, 256)), dim3(256), 0, s, W, k, k, t,
                           coef);
    else
        hipLaunchKernelGGL(gemv_cols_kernel, dim3(grid_for(L, 256)), dim3(256), 0, s, A, L, k, t,
                           coef);
    HMX_LAUNCHED("gemv_cols_kernel");
    HMX_HIP(hipMemcpyAsync(rank, counters + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    HMX_HIP(hipStreamSynchronize(s));
    return HMX_OK;
}

}  // namespace hmx

extern "C" {

int hmx_lsq_svd_workspace_bytes(int64_t n, int64_t m, size_t *bytes) {
    hmx::clear_error();
    if (!bytes) return hmx::fail(HMX_ERR_ARG, "bytes is NULL");
    if (n <= 0 || m <= 0) return hmx::fail(HMX_ERR_ARG, "empty problem");
    // n >= m: W (m x m); m > n: W (n x n) and the transpose of G
    const size_t k = (size_t)(n >= m ? m : n);
    const size_t extra = n >= m ? 0 : (size_t)n * m;
    *bytes = (k * k + extra + 3 * k) * sizeof(double) + 256;
    return HMX_OK;
}

int hmx_lsq_svd(double *G, int64_t n, int64_t m, const double *y, double damping, int filter,
                double cutoff, double *coef, double *sigma, double *info, void *workspace,
                size_t workspace_bytes, void *stream) {
    using namespace hmx;
    clear_error();
    if (!G || !y || !coef || !info) return fail(HMX_ERR_ARG, "NULL array");
    if (n <= 0 || m <= 0) return fail(HMX_ERR_ARG, "empty problem");
    if (!(damping >= 0.0)) return fail(HMX_ERR_ARG, "damping must be >= 0");
    if (!(cutoff >= 0.0)) return fail(HMX_ERR_ARG, "cutoff must be >= 0");
    size_t need = 0;
    int st = hmx_lsq_svd_workspace_bytes(n, m, &need);
    if (st) return st;
    if (!workspace || workspace_bytes < need)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    const bool transposed = m > n;
    const int64_t k = transposed ? n : m, L = transposed ? m : n;
    double *W = static_cast<double *>(workspace);
    double *A = G;
    double *vec = W + (size_t)k * k;
    if (transposed) {
        A = vec;
        vec += (size_t)n * m;
        hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)((m + 31) / 32), (unsigned)((n + 31) / 32)),
                           dim3(256), 0, s, G, n, m, A);
        HMX_LAUNCHED("transpose_kernel");
    }
    int *counters = reinterpret_cast<int *>(vec + 3 * k);
    int sweeps = 0, rank = 0;
    st = svd_lsq(A, L, k, transposed, y, damping, filter, cutoff, W, vec, counters, coef,
                 kSvdMaxSweeps, &sweeps, &rank, s);
    if (st) return st;
    if (sigma) {  // the k singular values, then zeros
        HMX_HIP(hipMemsetAsync(sigma, 0, (size_t)m * sizeof(double), s));
        HMX_HIP(hipMemcpyAsync(sigma, vec, (size_t)k * sizeof(double), hipMemcpyDeviceToDevice, s));
        HMX_HIP(hipStreamSynchronize(s));
    }
    info[0] = (double)sweeps;
    info[1] = (double)rank;
    return HMX_OK;
}

}  // extern "C"
