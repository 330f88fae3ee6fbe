// Equivalent-source least squares on the GPU.
//
// Restates verde.base.least_squares as EquivalentSources.fit calls it
// (src/harmonica/_equivalent_sources/cartesian.py:279-280): columns of the
// Jacobian J (J_ij = 1 / |x_i - p_j|) are scaled by their population
// standard deviation s_j (zero -> 1), the damped, weighted normal equations
//     A c' = b,   A = S J^T W J S + alpha I,   b = S J^T W d,   S = diag(1/s)
// are solved, and c = S c'.
//
// Pipeline (all on the caller's stream):
//   col_stats        s_j from two FP64 passes over J generated on the fly;
//   scaled_jacobian  X = W^(1/2) J S materialised row-major (n x m), FP64 or
//                    FP32;
//   syrk_upper       A = X^T X (+ shift I) on MFMA (v_mfma_f64_16x16x4f64, or
//                    v_mfma_f32_16x16x4f32 with FP32 accumulation), upper
//                    triangle, row-major; also the Cholesky trailing update;
//   potrf_diag / trsm_rowpanel / syrk_upper
//                    right-looking blocked Cholesky A = U^T U in FP64;
//   trsv_*           U^T y = r, U x = y, blocked;
//   apply_A          FP64 matrix-free A p (J regenerated on the fly) for the
//                    residuals of the refinement / preconditioned CG.
#include <cmath>
#include <type_traits>
#include <vector>

#include "devmath.h"
#include "hmx_internal.h"

namespace hmx {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kNB = 128;    // Cholesky block size
constexpr int kST = 128;    // SYRK block tile (kST x kST per workgroup of 256 threads)
constexpr int kKC = 16;     // SYRK K chunk staged in LDS
constexpr int kPad = 144;   // LDS row stride (doubles/floats) of the SYRK panels

__device__ __forceinline__ double inv_dist(double ae, double an, double au, double be, double bn,
                                           double bu) {
    const double de = ae - be, dn = an - bn, du = au - bu;
    return dm::div(1.0, dm::sqrt_nn(de * de + dn * dn + du * du));
}

// ------------------------------------------------------------ column stats
// mean_j and population std s_j of column j of J, two passes in FP64.
__global__ __launch_bounds__(256) void col_stats_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    int64_t n_data, const double *__restrict__ pe, const double *__restrict__ pn,
    const double *__restrict__ pu, int64_t n_src, double *__restrict__ scale) {
    __shared__ double4_t tile[256];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t jj = j < n_src ? j : n_src - 1;
    const double se = pe[jj], sn = pn[jj], su = pu[jj];
    double sum = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        const double mean = sum / (double)n_data;
        double acc = 0.0;
        for (int64_t t0 = 0; t0 < n_data; t0 += 256) {
            const int cnt = (int)(n_data - t0 < 256 ? n_data - t0 : 256);
            __syncthreads();
            if (threadIdx.x < cnt) {
                double4_t v;
                v.x = e[t0 + threadIdx.x];
                v.y = n[t0 + threadIdx.x];
                v.z = u[t0 + threadIdx.x];
                v.w = 0.0;
                tile[threadIdx.x] = v;
            }
            __syncthreads();
            for (int q = 0; q < cnt; ++q) {
                const double4_t o = tile[q];
                const double g = inv_dist(o.x, o.y, o.z, se, sn, su);
                if (pass == 0) {
                    acc += g;
                } else {
                    const double dv = g - mean;
                    acc = fma(dv, dv, acc);
                }
            }
        }
        sum = acc;
    }
    if (j < n_src) {
        double s = sqrt(sum / (double)n_data);
        scale[j] = s == 0.0 ? 1.0 : s;
    }
}

// -------------------------------------------------------- scaled jacobian
// X[i * ld + j] = sqrt(w_i) * J_ij / s_j (w == NULL -> 1).
template <class T>
__global__ __launch_bounds__(256) void scaled_jacobian_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    const double *__restrict__ w, int64_t n_data, const double *__restrict__ pe,
    const double *__restrict__ pn, const double *__restrict__ pu,
    const double *__restrict__ scale, int64_t n_src, T *__restrict__ X, int64_t ld) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n_src) return;
    const double se = pe[j], sn = pn[j], su = pu[j], inv_s = 1.0 / scale[j];
    for (int64_t i = blockIdx.y; i < n_data; i += gridDim.y) {
        const double sw = w ? sqrt(w[i]) : 1.0;
        X[i * ld + j] = (T)(sw * inv_dist(e[i], n[i], u[i], se, sn, su) * inv_s);
    }
}

// -------------------------------------------------------------- MFMA SYRK
template <class T>
struct Mfma;
// Operand layout of both 16x16x4 MFMAs (A 16x4, B 4x16): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16].  Element v of lane l's result is
// D[row(l, v)][l % 16] with row = l / 16 + 4 v for f64 and 4 (l / 16) + v
// for f32 (checked by tests/test_gpu_linalg.py through hmx_mfma_probe).
template <>
struct Mfma<double> {
    typedef double4_t acc_t;
    static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float> {
    typedef float4_t acc_t;
    static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// C = beta * C + sign * A^T B + shift * I on the selected tiles, with A
// (K x M) and B (K x N) row-major (leading dimensions lda, ldb) and C (M x N)
// row-major (ldc).  UPPER: M == N and only the upper triangle (r <= c) is
// computed (the symmetric rank-K update A^T A when B == A).  A workgroup of
// 4 wavefronts owns a kST x kST tile of C (2 x 2 waves of 64 x 64, each
// 4 x 4 MFMA 16x16x4 tiles); the K dimension is staged through LDS in
// chunks of kKC rows.  Every workgroup reads all its inputs before it writes
// its tile, and tiles are disjoint, so C may alias B when M <= kST (used by
// the in-place panel solve of the Cholesky factorisation).
template <class T, bool UPPER>
__global__ __launch_bounds__(256) void gemm_tn_kernel(const T *__restrict__ A, int64_t lda,
                                                      const T *B, int64_t ldb, int64_t K,
                                                      int64_t M, int64_t N, double *C, int64_t ldc,
                                                      double beta, double sign, double shift,
                                                      int64_t tiles_n) {
    using acc_t = typename Mfma<T>::acc_t;
    constexpr bool kFold = !std::is_same<T, double>::value;
    __shared__ T pr[kKC * kPad];
    __shared__ T pc[kKC * kPad];
    int64_t rb, cb;
    if (UPPER) {
        int64_t b = blockIdx.x, rem = tiles_n;
        rb = 0;
        while (b >= rem) {
            b -= rem;
            ++rb;
            --rem;
        }
        cb = rb + b;
    } else {
        rb = blockIdx.x / tiles_n;
        cb = blockIdx.x - rb * tiles_n;
    }
    const int64_t r0 = rb * kST, c0 = cb * kST;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    // FP32 MFMA accumulators are folded into FP64 after every K chunk, so
    // rounding accumulates over kKC rows only (FP32 inputs, FP64 sums).
    acc_t acc[4][4];
    double4_t dacc[kFold ? 4 : 1][kFold ? 4 : 1];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = acc_t{0, 0, 0, 0};
    if constexpr (kFold) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) dacc[a][c] = double4_t{0, 0, 0, 0};
    }
    for (int64_t k0 = 0; k0 < K; k0 += kKC) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < kKC * kST; idx += 256) {
            const int kk = idx / kST, cc = idx - kk * kST;
            const int64_t k = k0 + kk;
            const bool kin = k < K;
            const int64_t gr = r0 + cc, gc = c0 + cc;
            pr[kk * kPad + cc] = (kin && gr < M) ? A[k * lda + gr] : T(0);
            pc[kk * kPad + cc] = (kin && gc < N) ? B[k * ldb + gc] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kKC; ks += 4) {
            const int kr = ks + (lane >> 4);
            T a[4], bb[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                a[t] = pr[kr * kPad + 64 * wr + 16 * t + (lane & 15)];
                bb[t] = pc[kr * kPad + 64 * wc + 16 * t + (lane & 15)];
            }
#pragma unroll
            for (int ti = 0; ti < 4; ++ti)
#pragma unroll
                for (int tj = 0; tj < 4; ++tj)
                    acc[ti][tj] = Mfma<T>::step(a[ti], bb[tj], acc[ti][tj]);
        }
        if constexpr (kFold) {
#pragma unroll
            for (int ti = 0; ti < 4; ++ti)
#pragma unroll
                for (int tj = 0; tj < 4; ++tj) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) dacc[ti][tj][v] += (double)acc[ti][tj][v];
                    acc[ti][tj] = acc_t{0, 0, 0, 0};
                }
        }
    }
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t r = r0 + 64 * wr + 16 * ti + Mfma<T>::row(lane, v);
                const int64_t c = c0 + 64 * wc + 16 * tj + (lane & 15);
                if (r < M && c < N && (!UPPER || r <= c)) {
                    const double sum = kFold ? (double)dacc[kFold ? ti : 0][kFold ? tj : 0][v]
                                             : (double)acc[ti][tj][v];
                    double val = sign * sum;
                    if (r == c) val += shift;
                    C[r * ldc + c] = beta == 0.0 ? val : fma(beta, C[r * ldc + c], val);
                }
            }
}

template <class T>
int launch_syrk(const T *X, int64_t ldx, int64_t K, int64_t m, double *C, int64_t ldc, double beta,
                double sign, double shift, hipStream_t s) {
    if (m <= 0) return HMX_OK;
    const int64_t tiles = (m + kST - 1) / kST;
    const int64_t blocks = tiles * (tiles + 1) / 2;
    hipLaunchKernelGGL((gemm_tn_kernel<T, true>), dim3((unsigned)blocks), dim3(256), 0, s, X, ldx, X,
                       ldx, K, m, m, C, ldc, beta, sign, shift, tiles);
    HMX_LAUNCHED("gemm_tn_kernel<upper>");
    return HMX_OK;
}

int launch_gemm_tn(const double *A, int64_t lda, const double *B, int64_t ldb, int64_t K,
                   int64_t M, int64_t N, double *C, int64_t ldc, double beta, double sign,
                   hipStream_t s) {
    if (M <= 0 || N <= 0) return HMX_OK;
    const int64_t tm = (M + kST - 1) / kST, tn = (N + kST - 1) / kST;
    hipLaunchKernelGGL((gemm_tn_kernel<double, false>), dim3((unsigned)(tm * tn)), dim3(256), 0, s,
                       A, lda, B, ldb, K, M, N, C, ldc, beta, sign, 0.0, tn);
    HMX_LAUNCHED("gemm_tn_kernel");
    return HMX_OK;
}

// ---------------------------------------------------------------- Cholesky
// Factor the nb x nb diagonal block at D (upper, row-major, ld) in place as
// U^T U and write U^-1 to Dinv (nb x nb, row-major, ld kNB) and its
// transpose to DinvT.  One workgroup; the block lives in LDS.  On a
// non-positive pivot, *info is set (once) to its 1-based global column and
// the block is left unfinished.
__global__ __launch_bounds__(1024) void potrf_diag_kernel(double *__restrict__ D, int64_t ld,
                                                          int nb, int64_t col0,
                                                          double *__restrict__ Dinv,
                                                          double *__restrict__ DinvT,
                                                          int *__restrict__ info) {
    constexpr int S = kNB + 1;
    __shared__ double L[kNB * S];
    __shared__ double col[kNB];
    __shared__ int bad;
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const int r = idx / nb, c = idx - r * nb;
        L[r * S + c] = D[r * ld + c];
    }
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        const double d = L[k * S + k];
        if (!(d > 0.0)) {
            if (threadIdx.x == 0) {
                bad = 1;
                atomicCAS(info, 0, (int)(col0 + k + 1));
            }
        }
        __syncthreads();
        if (bad) return;
        const double piv = sqrt(d);
        const double inv = 1.0 / piv;
        for (int c = k + 1 + threadIdx.x; c < nb; c += blockDim.x) L[k * S + c] *= inv;
        __syncthreads();
        if (threadIdx.x == 0) L[k * S + k] = piv;
        const int rem = nb - k - 1;
        for (int idx = threadIdx.x; idx < rem * rem; idx += blockDim.x) {
            const int r = k + 1 + idx / rem, c = k + 1 + idx % rem;
            if (r <= c) L[r * S + c] = fma(-L[k * S + r], L[k * S + c], L[r * S + c]);
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const int r = idx / nb, c = idx - r * nb;
        D[r * ld + c] = r <= c ? L[r * S + c] : 0.0;
    }
    // In-place inverse of the upper triangular factor, column by column:
    // X[j][j] = 1 / U[j][j], X[i][j] = -X[j][j] sum_{k=i}^{j-1} X[i][k] U[k][j].
    for (int j = 0; j < nb; ++j) {
        __syncthreads();
        for (int i = threadIdx.x; i <= j; i += blockDim.x) col[i] = L[i * S + j];
        __syncthreads();
        const double djj = 1.0 / col[j];
        for (int i = threadIdx.x; i < j; i += blockDim.x) {
            double acc = 0.0;
            for (int k = i; k < j; ++k) acc = fma(L[i * S + k], col[k], acc);
            L[i * S + j] = -djj * acc;
        }
        if (threadIdx.x == 0) L[j * S + j] = djj;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < kNB * kNB; idx += blockDim.x) {
        const int r = idx / kNB, c = idx - r * kNB;
        const double v = (r < nb && c < nb && r <= c) ? L[r * S + c] : 0.0;
        Dinv[r * kNB + c] = v;
        DinvT[c * kNB + r] = v;
    }
}

// ------------------------------------------------------------------ solves
// One step of the blocked forward substitution U^T y = r for the block at
// rows [p, p + nb): every workgroup forms y_b = U_bb^-T r_b from DinvT (the
// 1-based workgroup 0 stores it), then updates its slice of the remaining
// right-hand side r[c] -= sum_k U[p + k][c] y_b[k], c in [p + nb, m).
__global__ __launch_bounds__(256) void trsv_fwd_step_kernel(const double *__restrict__ U,
                                                            int64_t ld, int64_t m, int64_t p,
                                                            int nb,
                                                            const double *__restrict__ DinvT,
                                                            double *__restrict__ r,
                                                            double *__restrict__ y) {
    __shared__ double yb[kNB];
    __shared__ double rb[kNB];
    if (threadIdx.x < nb) rb[threadIdx.x] = r[p + threadIdx.x];
    __syncthreads();
    if (threadIdx.x < nb) {
        // y_b[k] = sum_l Dinv[l][k] r_b[l] = sum_l DinvT[k][l] r_b[l]
        double acc = 0.0;
        const double *row = DinvT + threadIdx.x * kNB;
        for (int l = 0; l <= (int)threadIdx.x; ++l) acc = fma(row[l], rb[l], acc);
        yb[threadIdx.x] = acc;
        if (blockIdx.x == 0) y[p + threadIdx.x] = acc;
    }
    __syncthreads();
    const int64_t base = p + nb;
    for (int64_t c = base + (int64_t)blockIdx.x * 256 + threadIdx.x; c < m;
         c += (int64_t)gridDim.x * 256) {
        double v = r[c];
        for (int k = 0; k < nb; ++k) v = fma(-U[(p + k) * ld + c], yb[k], v);
        r[c] = v;
    }
}

// One step of the blocked backward substitution U x = y for the block at
// rows [p, p + nb): x_b = U_bb^-1 y_b, then y[i] -= sum_c U[i][p + c] x_b[c]
// for i in [0, p).
__global__ __launch_bounds__(256) void trsv_bwd_step_kernel(const double *__restrict__ U,
                                                            int64_t ld, int64_t p, int nb,
                                                            const double *__restrict__ Dinv,
                                                            double *__restrict__ y,
                                                            double *__restrict__ x) {
    __shared__ double xb[kNB];
    __shared__ double yb[kNB];
    if (threadIdx.x < nb) yb[threadIdx.x] = y[p + threadIdx.x];
    __syncthreads();
    if (threadIdx.x < nb) {
        double acc = 0.0;
        const double *row = Dinv + threadIdx.x * kNB;
        for (int l = threadIdx.x; l < nb; ++l) acc = fma(row[l], yb[l], acc);
        xb[threadIdx.x] = acc;
        if (blockIdx.x == 0) x[p + threadIdx.x] = acc;
    }
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < p; i += (int64_t)gridDim.x * 256) {
        double v = y[i];
        const double *row = U + i * ld + p;
        for (int c = 0; c < nb; ++c) v = fma(-row[c], xb[c], v);
        y[i] = v;
    }
}

// ------------------------------------------------------ matrix-free A p
// t_i = sum_j J_ij p_j / s_j (the unweighted data-space image of p).
__global__ __launch_bounds__(256) void j_apply_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    int64_t n_data, const double *__restrict__ pe, const double *__restrict__ pn,
    const double *__restrict__ pu, const double *__restrict__ p,
    const double *__restrict__ scale, int64_t n_src, double *__restrict__ t) {
    __shared__ double4_t tile[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t ii = i < n_data ? i : n_data - 1;
    const double oe = e[ii], on = n[ii], ou = u[ii];
    double acc = 0.0;
    for (int64_t t0 = 0; t0 < n_src; t0 += 256) {
        const int cnt = (int)(n_src - t0 < 256 ? n_src - t0 : 256);
        __syncthreads();
        if (threadIdx.x < cnt) {
            const int64_t j = t0 + threadIdx.x;
            double4_t v;
            v.x = pe[j];
            v.y = pn[j];
            v.z = pu[j];
            v.w = p[j] / scale[j];
            tile[threadIdx.x] = v;
        }
        __syncthreads();
        for (int q = 0; q < cnt; ++q) {
            const double4_t s = tile[q];
            acc = fma(s.w, inv_dist(oe, on, ou, s.x, s.y, s.z), acc);
        }
    }
    if (i < n_data) t[i] = acc;
}

// out_j = (1 / s_j) sum_i J_ij w_i t_i + alpha p_j  (w, p may be NULL).
__global__ __launch_bounds__(256) void jt_apply_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    const double *__restrict__ w, int64_t n_data, const double *__restrict__ pe,
    const double *__restrict__ pn, const double *__restrict__ pu, const double *__restrict__ t,
    const double *__restrict__ scale, int64_t n_src, double alpha, const double *__restrict__ p,
    double *__restrict__ out) {
    __shared__ double4_t tile[256];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t jj = j < n_src ? j : n_src - 1;
    const double se = pe[jj], sn = pn[jj], su = pu[jj];
    double acc = 0.0;
    for (int64_t t0 = 0; t0 < n_data; t0 += 256) {
        const int cnt = (int)(n_data - t0 < 256 ? n_data - t0 : 256);
        __syncthreads();
        if (threadIdx.x < cnt) {
            const int64_t i = t0 + threadIdx.x;
            double4_t v;
            v.x = e[i];
            v.y = n[i];
            v.z = u[i];
            v.w = w ? w[i] * t[i] : t[i];
            tile[threadIdx.x] = v;
        }
        __syncthreads();
        for (int q = 0; q < cnt; ++q) {
            const double4_t o = tile[q];
            acc = fma(o.w, inv_dist(o.x, o.y, o.z, se, sn, su), acc);
        }
    }
    if (j < n_src) {
        double v = acc / scale[j];
        if (p) v = fma(alpha, p[j], v);
        out[j] = v;
    }
}

// ------------------------------------------------------ small vector ops
__global__ void axpy_kernel(int64_t n, double a, const double *__restrict__ x,
                            double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = fma(a, x[i], y[i]);
}

// y = x + b * y
__global__ void xpby_kernel(int64_t n, const double *__restrict__ x, double b,
                            double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = fma(b, y[i], x[i]);
}

__global__ void sub_kernel(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                           double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

__global__ void div_kernel(int64_t n, const double *__restrict__ x, const double *__restrict__ s,
                           double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] / s[i];
}

__global__ void copy_kernel(int64_t n, const double *__restrict__ x, double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

// Partial dot products / max-abs (mode 0 / 1): one value per workgroup,
// combined on the host in workgroup order (deterministic).
__global__ __launch_bounds__(256) void reduce_partial_kernel(int mode, int64_t n, int64_t stride,
                                                             const double *__restrict__ x,
                                                             const double *__restrict__ y,
                                                             double *__restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double v = x[i * stride];
        acc = mode == 0 ? fma(v, y[i], acc) : fmax(acc, fabs(v));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            red[threadIdx.x] = mode == 0 ? red[threadIdx.x] + red[threadIdx.x + s]
                                         : fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out_j = sum_i X[i][j]^2 (FP64), the diagonal of X^T X.
template <class T>
__global__ __launch_bounds__(256) void colsumsq_kernel(const T *__restrict__ X, int64_t ld,
                                                       int64_t n_rows, int64_t m,
                                                       double *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    double acc = 0.0;
    for (int64_t i = 0; i < n_rows; ++i) {
        const double v = (double)X[i * ld + j];
        acc = fma(v, v, acc);
    }
    out[j] = acc;
}

__global__ void zero_int_kernel(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

// ----------------------------------------------------------- host driver
inline unsigned blocks_for(int64_t n, int per = 256) {
    const int64_t b = (n + per - 1) / per;
    return (unsigned)(b < 1 ? 1 : b);
}

struct Fit {
    const double *e, *n, *u, *w;
    int64_t nd;
    const double *pe, *pn, *pu;
    int64_t m;
    double alpha;
    hipStream_t s;
    double *scale, *A, *Dinv, *DinvT, *tmp, *part;
    double host_part[256];
    int *dinfo;

    int reduce(int mode, const double *x, const double *y, int64_t n, int64_t stride, double *out) {
        const int nb = 256;
        hipLaunchKernelGGL(reduce_partial_kernel, dim3(nb), dim3(256), 0, s, mode, n, stride, x, y, part);
        HMX_LAUNCHED("reduce_partial_kernel");
        HMX_HIP(hipMemcpyAsync(host_part, part, nb * sizeof(double), hipMemcpyDeviceToHost, s));
        HMX_HIP(hipStreamSynchronize(s));
        double acc = 0.0;
        for (int i = 0; i < nb; ++i) acc = mode == 0 ? acc + host_part[i] : std::fmax(acc, host_part[i]);
        *out = acc;
        return HMX_OK;
    }

    // out = A p (FP64, matrix-free); t receives J S p (data space).
    int apply(const double *p, double *t, double *out) {
        hipLaunchKernelGGL(j_apply_kernel, dim3(blocks_for(nd)), dim3(256), 0, s, e, n, u, nd, pe, pn,
                           pu, p, scale, m, t);
        HMX_LAUNCHED("j_apply_kernel");
        hipLaunchKernelGGL(jt_apply_kernel, dim3(blocks_for(m)), dim3(256), 0, s, e, n, u, w, nd, pe,
                           pn, pu, t, scale, m, alpha, p, out);
        HMX_LAUNCHED("jt_apply_kernel");
        return HMX_OK;
    }

    int cholesky(int *pivot) {
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, dinfo);
        HMX_LAUNCHED("zero_int_kernel");
        for (int64_t p = 0; p < m; p += kNB) {
            const int nb = (int)(m - p < kNB ? m - p : kNB);
            double *D = A + p * m + p;
            const int64_t blk = p / kNB;
            hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(1024), 0, s, D, m, nb, p,
                               Dinv + blk * kNB * kNB, DinvT + blk * kNB * kNB, dinfo);
            HMX_LAUNCHED("potrf_diag_kernel");
            const int64_t rest = m - p - nb;
            if (rest <= 0) break;
            // U12 = U11^-T A12 = (U11^-1)^T A12, in place.
            int st = launch_gemm_tn(Dinv + blk * kNB * kNB, kNB, D + nb, m, nb, nb, rest, D + nb, m,
                                    0.0, 1.0, s);
            if (st) return st;
            st = launch_syrk<double>(D + nb, m, nb, rest, D + nb * m + nb, m, 1.0, -1.0, 0.0, s);
            if (st) return st;
        }
        HMX_HIP(hipMemcpyAsync(pivot, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
        HMX_HIP(hipStreamSynchronize(s));
        return HMX_OK;
    }

    // x = A^-1 r with the factor; r is overwritten (work), y is scratch.
    int solve(double *r, double *y, double *x) {
        for (int64_t p = 0; p < m; p += kNB) {
            const int nb = (int)(m - p < kNB ? m - p : kNB);
            const int64_t rest = m - p - nb;
            const unsigned g = blocks_for(rest > 0 ? rest : 1, 1024);
            hipLaunchKernelGGL(trsv_fwd_step_kernel, dim3(g), dim3(256), 0, s, A, m, m, p, nb,
                               DinvT + (p / kNB) * kNB * kNB, r, y);
            HMX_LAUNCHED("trsv_fwd_step_kernel");
        }
        const int64_t last = ((m - 1) / kNB) * kNB;
        for (int64_t p = last; p >= 0; p -= kNB) {
            const int nb = (int)(m - p < kNB ? m - p : kNB);
            const unsigned g = blocks_for(p > 0 ? p : 1, 1024);
            hipLaunchKernelGGL(trsv_bwd_step_kernel, dim3(g), dim3(256), 0, s, A, m, p, nb,
                               Dinv + (p / kNB) * kNB * kNB, y, x);
            HMX_LAUNCHED("trsv_bwd_step_kernel");
        }
        return HMX_OK;
    }
};

// One 16x16x4 MFMA of one wavefront with per-lane operands a[l], b[l];
// out[4 l + v] = element v of lane l's result (layout check).
__global__ void mfma_probe_kernel(int dtype, const double *__restrict__ a,
                                  const double *__restrict__ b, double *__restrict__ out) {
    const int l = threadIdx.x;
    if (dtype == HMX_F64) {
        double4_t c = {0, 0, 0, 0};
        c = Mfma<double>::step(a[l], b[l], c);
        for (int v = 0; v < 4; ++v) out[4 * l + v] = c[v];
    } else {
        float4_t c = {0, 0, 0, 0};
        c = Mfma<float>::step((float)a[l], (float)b[l], c);
        for (int v = 0; v < 4; ++v) out[4 * l + v] = (double)c[v];
    }
}

}  // namespace
}  // namespace hmx

extern "C" {

int hmx_mfma_probe(int dtype, const double *a, const double *b, double *out, void *stream) {
    using namespace hmx;
    clear_error();
    if (dtype != HMX_F64 && dtype != HMX_F32) return fail(HMX_ERR_ARG, "invalid dtype %d", dtype);
    if (!a || !b || !out) return fail(HMX_ERR_ARG, "NULL array");
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, dtype, a, b, out);
    HMX_LAUNCHED("mfma_probe_kernel");
    return HMX_OK;
}

int hmx_gemm_tn(int dtype, const void *A, int64_t lda, const void *B, int64_t ldb, int64_t K,
                int64_t M, int64_t N, double *C, int64_t ldc, double beta, double sign,
                double shift, int upper, void *stream) {
    using namespace hmx;
    clear_error();
    if (K < 0 || M < 0 || N < 0) return fail(HMX_ERR_ARG, "negative size");
    if (lda < M || ldb < N || ldc < N) return fail(HMX_ERR_ARG, "leading dimension too small");
    if (upper && M != N) return fail(HMX_ERR_ARG, "upper requires M == N");
    if (M == 0 || N == 0) return HMX_OK;
    if (!A || !B || !C) return fail(HMX_ERR_ARG, "NULL array");
    hipStream_t s = (hipStream_t)stream;
    const int64_t tm = (M + kST - 1) / kST, tn = (N + kST - 1) / kST;
    const unsigned blocks = (unsigned)(upper ? tn * (tn + 1) / 2 : tm * tn);
    if (dtype == HMX_F64) {
        if (upper)
            hipLaunchKernelGGL((gemm_tn_kernel<double, true>), dim3(blocks), dim3(256), 0, s,
                               (const double *)A, lda, (const double *)B, ldb, K, M, N, C, ldc,
                               beta, sign, shift, tn);
        else
            hipLaunchKernelGGL((gemm_tn_kernel<double, false>), dim3(blocks), dim3(256), 0, s,
                               (const double *)A, lda, (const double *)B, ldb, K, M, N, C, ldc,
                               beta, sign, shift, tn);
    } else if (dtype == HMX_F32) {
        if (upper)
            hipLaunchKernelGGL((gemm_tn_kernel<float, true>), dim3(blocks), dim3(256), 0, s,
                               (const float *)A, lda, (const float *)B, ldb, K, M, N, C, ldc, beta,
                               sign, shift, tn);
        else
            hipLaunchKernelGGL((gemm_tn_kernel<float, false>), dim3(blocks), dim3(256), 0, s,
                               (const float *)A, lda, (const float *)B, ldb, K, M, N, C, ldc, beta,
                               sign, shift, tn);
    } else {
        return fail(HMX_ERR_ARG, "invalid dtype %d", dtype);
    }
    HMX_LAUNCHED("gemm_tn_kernel");
    return HMX_OK;
}

int hmx_cholesky_workspace_bytes(int64_t m, size_t *bytes) {
    using namespace hmx;
    if (!bytes) return fail(HMX_ERR_ARG, "bytes is NULL");
    if (m <= 0) return fail(HMX_ERR_ARG, "empty matrix");
    const size_t nblk = (size_t)((m + kNB - 1) / kNB);
    *bytes = 2 * nblk * kNB * kNB * sizeof(double) + (2 * (size_t)m + 256) * sizeof(double) + 256;
    return HMX_OK;
}

int hmx_cholesky_solve(double *A, int64_t m, double *x, int *info, void *workspace,
                       size_t workspace_bytes, void *stream) {
    using namespace hmx;
    clear_error();
    size_t need = 0;
    int st = hmx_cholesky_workspace_bytes(m, &need);
    if (st) return st;
    if (!A || !x || !info) return fail(HMX_ERR_ARG, "NULL array");
    if (!workspace || workspace_bytes < need)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    const size_t nblk = (size_t)((m + kNB - 1) / kNB);
    Fit f;
    f.m = m;
    f.s = (hipStream_t)stream;
    f.A = A;
    f.Dinv = static_cast<double *>(workspace);
    f.DinvT = f.Dinv + nblk * kNB * kNB;
    double *wk = f.DinvT + nblk * kNB * kNB;
    double *wk2 = wk + m;
    f.part = wk2 + m;
    f.dinfo = reinterpret_cast<int *>(f.part + 256);
    st = f.cholesky(info);
    if (st) return st;
    if (*info) return HMX_OK;
    hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(m)), dim3(256), 0, f.s, m, x, wk);
    HMX_LAUNCHED("copy_kernel");
    st = f.solve(wk, wk2, x);
    if (st) return st;
    HMX_HIP(hipStreamSynchronize(f.s));
    return HMX_OK;
}

int hmx_eqs_fit_workspace_bytes(int64_t n_data, int64_t n_src, int precision_mode,
                                size_t *bytes) {
    using namespace hmx;
    if (!bytes) return fail(HMX_ERR_ARG, "bytes is NULL");
    if (n_data <= 0 || n_src <= 0) return fail(HMX_ERR_ARG, "empty problem");
    if (precision_mode != HMX_FIT_FP64 && precision_mode != HMX_FIT_FP32_MFMA)
        return fail(HMX_ERR_ARG, "invalid precision mode %d", precision_mode);
    const size_t m = (size_t)n_src, n = (size_t)n_data;
    const size_t nblk = (m + kNB - 1) / kNB;
    const size_t x_elem = precision_mode == HMX_FIT_FP32_MFMA ? sizeof(float) : sizeof(double);
    *bytes = m * m * sizeof(double)                      // A, then its factor U
             + n * m * x_elem                            // X = W^1/2 J S
             + 2 * nblk * kNB * kNB * sizeof(double)     // inverses of diagonal blocks
             + (9 * m + n + 256) * sizeof(double) + 256;  // 9 m-vectors, n-vector, partials, info
    return HMX_OK;
}

int hmx_eqs_fit(const double *easting, const double *northing, const double *upward,
                int64_t n_data, const double *data, const double *weights,
                const double *p_easting, const double *p_northing, const double *p_upward,
                int64_t n_src, double damping, int precision_mode, int max_iter, double rtol,
                double *coefs, double *info, void *workspace, size_t workspace_bytes,
                void *stream) {
    using namespace hmx;
    clear_error();
    if (n_data <= 0 || n_src <= 0) return fail(HMX_ERR_ARG, "empty problem");
    if (n_data > INT32_MAX || n_src > INT32_MAX) return fail(HMX_ERR_ARG, "problem too large");
    if (!easting || !northing || !upward || !data || !p_easting || !p_northing || !p_upward ||
        !coefs || !info)
        return fail(HMX_ERR_ARG, "NULL array");
    if (!(damping >= 0.0)) return fail(HMX_ERR_ARG, "damping must be >= 0");
    if (max_iter < 0) return fail(HMX_ERR_ARG, "max_iter must be >= 0");
    size_t need = 0;
    int st = hmx_eqs_fit_workspace_bytes(n_data, n_src, precision_mode, &need);
    if (st) return st;
    if (!workspace || workspace_bytes < need)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    const int64_t m = n_src, nd = n_data;
    const bool mixed = precision_mode == HMX_FIT_FP32_MFMA;
    const size_t nblk = (size_t)((m + kNB - 1) / kNB);

    char *ws = static_cast<char *>(workspace);
    Fit f;
    f.e = easting;
    f.n = northing;
    f.u = upward;
    f.w = weights;
    f.nd = nd;
    f.pe = p_easting;
    f.pn = p_northing;
    f.pu = p_upward;
    f.m = m;
    f.alpha = damping;
    f.s = s;
    f.A = reinterpret_cast<double *>(ws);
    ws += (size_t)m * m * sizeof(double);
    void *X = ws;
    ws += (size_t)nd * m * (mixed ? sizeof(float) : sizeof(double));
    f.Dinv = reinterpret_cast<double *>(ws);
    f.DinvT = f.Dinv + nblk * kNB * kNB;
    double *vec = f.DinvT + nblk * kNB * kNB;
    f.scale = vec;
    double *b = vec + m, *x = b + m, *r = x + m, *z = r + m, *pd = z + m, *q = pd + m, *wk = q + m,
           *wk2 = wk + m;
    f.tmp = wk2 + m;  // n_data
    f.part = f.tmp + nd;
    f.dinfo = reinterpret_cast<int *>(f.part + 256);
    const unsigned gm = blocks_for(m), gy = (unsigned)(nd < 8192 ? nd : 8192);

    // 1. column scales, right-hand side b = S J^T W d (FP64).
    hipLaunchKernelGGL(col_stats_kernel, dim3(gm), dim3(256), 0, s, easting, northing, upward, nd,
                       p_easting, p_northing, p_upward, m, f.scale);
    HMX_LAUNCHED("col_stats_kernel");
    hipLaunchKernelGGL(jt_apply_kernel, dim3(gm), dim3(256), 0, s, easting, northing, upward,
                       weights, nd, p_easting, p_northing, p_upward, data, f.scale, m, 0.0,
                       (const double *)nullptr, b);
    HMX_LAUNCHED("jt_apply_kernel");

    // 2. X = W^1/2 J S and the normal matrix in the mode's precision.
    if (mixed) {
        hipLaunchKernelGGL(scaled_jacobian_kernel<float>, dim3(gm, gy), dim3(256), 0, s, easting,
                           northing, upward, weights, nd, p_easting, p_northing, p_upward, f.scale,
                           m, static_cast<float *>(X), m);
    } else {
        hipLaunchKernelGGL(scaled_jacobian_kernel<double>, dim3(gm, gy), dim3(256), 0, s, easting,
                           northing, upward, weights, nd, p_easting, p_northing, p_upward, f.scale,
                           m, static_cast<double *>(X), m);
    }
    HMX_LAUNCHED("scaled_jacobian_kernel");

    // The normal matrix of FP32 inputs (FP64 sums) differs from the FP64 one
    // by about 2^-24 of its entries; with damping small against its largest
    // eigenvalue that can make it indefinite.  A failed factorisation is
    // retried with a shift tau (then tenfold larger), and the factor of the
    // (shifted) low-precision matrix serves as the preconditioner of CG on
    // the exact FP64 operator.
    double tau = 0.0, dmax = 0.0;
    auto diag_max = [&](double *out) -> int {
        if (mixed)
            hipLaunchKernelGGL(colsumsq_kernel<float>, dim3(gm), dim3(256), 0, s,
                               static_cast<const float *>(X), m, nd, m, wk);
        else
            hipLaunchKernelGGL(colsumsq_kernel<double>, dim3(gm), dim3(256), 0, s,
                               static_cast<const double *>(X), m, nd, m, wk);
        HMX_LAUNCHED("colsumsq_kernel");
        return f.reduce(1, wk, nullptr, m, 1, out);
    };
    int pivot = 0, attempts = 0;
    for (;;) {
        if (mixed)
            st = launch_syrk<float>(static_cast<const float *>(X), m, nd, m, f.A, m, 0.0, 1.0,
                                    damping + tau, s);
        else
            st = launch_syrk<double>(static_cast<const double *>(X), m, nd, m, f.A, m, 0.0, 1.0,
                                     damping + tau, s);
        if (st) return st;
        st = f.cholesky(&pivot);
        if (st) return st;
        if (pivot == 0) break;
        if (++attempts > 12)
            return fail(HMX_ERR_NUMERIC, "normal matrix not positive definite (column %d)", pivot);
        if (dmax == 0.0) {
            st = diag_max(&dmax);
            if (st) return st;
        }
        // First shift: about the rounding of the normal matrix in the mode's
        // input precision (2^-26 for FP32 inputs with FP64 sums, 1e-14 for
        // FP64); then tenfold.
        tau = tau > 0.0 ? 10.0 * tau : (mixed ? std::ldexp(dmax, -26) : 1e-14 * dmax);
    }

    // 3. x = (factor)^-1 b.
    hipLaunchKernelGGL(copy_kernel, dim3(gm), dim3(256), 0, s, m, b, wk);
    HMX_LAUNCHED("copy_kernel");
    st = f.solve(wk, wk2, x);
    if (st) return st;

    // 4. Preconditioned CG on the exact FP64 operator A = S J^T W J S + alpha I
    // (starting from x), stopping when the data-space update
    // |a| |J S p| falls below rtol |J S x| or after max_iter iterations.
    int iters = 0;
    double last_rel = 0.0;
    // A shifted factorisation is only a preconditioner: refine even when
    // the caller asked for a direct solve.
    const int allowed = (tau > 0.0 && max_iter == 0) ? 200 : max_iter;
    if (allowed > 0) {
        double *t_dat = f.tmp;
        st = f.apply(x, t_dat, q);  // q = A x
        if (st) return st;
        hipLaunchKernelGGL(sub_kernel, dim3(gm), dim3(256), 0, s, m, b, q, r);  // r = b - A x
        HMX_LAUNCHED("sub_kernel");
        double ynorm = 0.0;
        st = f.reduce(0, t_dat, t_dat, nd, 1, &ynorm);  // |J S x|^2
        if (st) return st;
        hipLaunchKernelGGL(copy_kernel, dim3(gm), dim3(256), 0, s, m, r, wk);
        HMX_LAUNCHED("copy_kernel");
        st = f.solve(wk, wk2, z);
        if (st) return st;
        hipLaunchKernelGGL(copy_kernel, dim3(gm), dim3(256), 0, s, m, z, pd);
        HMX_LAUNCHED("copy_kernel");
        double rz = 0.0;
        st = f.reduce(0, r, z, m, 1, &rz);
        if (st) return st;
        for (iters = 1; iters <= allowed; ++iters) {
            st = f.apply(pd, t_dat, q);  // q = A p, t = J S p
            if (st) return st;
            double pq = 0.0, tt = 0.0;
            st = f.reduce(0, pd, q, m, 1, &pq);
            if (st) return st;
            st = f.reduce(0, t_dat, t_dat, nd, 1, &tt);
            if (st) return st;
            if (!(pq > 0.0)) break;
            const double a = rz / pq;
            hipLaunchKernelGGL(axpy_kernel, dim3(gm), dim3(256), 0, s, m, a, pd, x);
            HMX_LAUNCHED("axpy_kernel");
            hipLaunchKernelGGL(axpy_kernel, dim3(gm), dim3(256), 0, s, m, -a, q, r);
            HMX_LAUNCHED("axpy_kernel");
            last_rel = std::fabs(a) * std::sqrt(tt) / std::sqrt(ynorm > 0 ? ynorm : 1.0);
            if (last_rel <= rtol) break;
            hipLaunchKernelGGL(copy_kernel, dim3(gm), dim3(256), 0, s, m, r, wk);
            HMX_LAUNCHED("copy_kernel");
            st = f.solve(wk, wk2, z);
            if (st) return st;
            double rz_new = 0.0;
            st = f.reduce(0, r, z, m, 1, &rz_new);
            if (st) return st;
            hipLaunchKernelGGL(xpby_kernel, dim3(gm), dim3(256), 0, s, m, z, rz_new / rz, pd);
            HMX_LAUNCHED("xpby_kernel");
            rz = rz_new;
        }
        if (iters > allowed) iters = allowed;
    }

    // 5. c = S c'.
    hipLaunchKernelGGL(div_kernel, dim3(gm), dim3(256), 0, s, m, x, f.scale, coefs);
    HMX_LAUNCHED("div_kernel");
    HMX_HIP(hipStreamSynchronize(s));
    info[0] = (double)iters;
    info[1] = last_rel;
    info[2] = tau;
    info[3] = (double)attempts;
    info[4] = dmax;
    return HMX_OK;
}

}  // extern "C"
