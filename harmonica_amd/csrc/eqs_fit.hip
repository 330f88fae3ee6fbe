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
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef int int32x4_t __attribute__((ext_vector_type(4)));  // also 16 packed int8

constexpr int kNB = 128;    // Cholesky block size
constexpr int kOuter = 1024;  // Cholesky outer block (K of the big trailing updates)
constexpr int kST = 128;    // SYRK block tile (kST x kST per workgroup of 256 threads)
constexpr int kKC = 16;     // SYRK K chunk staged in LDS
constexpr int kPad = 144;   // LDS row stride (doubles/floats) of the SYRK panels

__device__ __forceinline__ double inv_dist(double ae, double an, double au, double be, double bn,
                                           double bu) {
    const double de = ae - be, dn = an - bn, du = au - bu;
    return dm::rsqrt_pos(de * de + dn * dn + du * du);
}

// ------------------------------------------------------------ column stats
// The reductions over data (column statistics, J^T products) and over
// sources (J products) are split over blockIdx.y, so that outputs of a few
// thousand elements (the windows of the gradient-boosted fit) still fill
// the device: split y covers indices [y chunk, (y + 1) chunk) and writes
// its partial sum to part[y * n_out + k]; a finish kernel adds the splits
// in order.
inline int64_t reduce_splits(int64_t n_out, int64_t n_red) {
    const int64_t blocks = (n_out + 255) / 256;
    int64_t splits = (1024 + blocks - 1) / blocks;  // about 1024 workgroups in all
    const int64_t most = (n_red + 511) / 512;       // at least 512 terms per split
    splits = splits < most ? splits : most;
    return splits < 1 ? 1 : splits;
}
// Doubles of the partial sums for any reduce_splits(n_out, .) x n_out.
inline size_t split_doubles(int64_t n_out) { return (size_t)(1024 + 1) * 256 + (size_t)n_out; }

// Column moments of J over the split's data rows: pass 0 the sums of J_ij,
// pass 1 the sums of (J_ij - mean_j)^2.
__global__ __launch_bounds__(256) void col_moment_part_kernel(
    int pass, const double *__restrict__ e, const double *__restrict__ n,
    const double *__restrict__ u, int64_t n_data, const double *__restrict__ pe,
    const double *__restrict__ pn, const double *__restrict__ pu, int64_t n_src,
    const double *__restrict__ mean, int64_t chunk, double *__restrict__ part) {
    __shared__ double4_t tile[256];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t jj = j < n_src ? j : n_src - 1;
    const double se = pe[jj], sn = pn[jj], su = pu[jj];
    const double mu = pass ? mean[jj] : 0.0;
    const int64_t i0 = (int64_t)blockIdx.y * chunk;
    const int64_t i1 = i0 + chunk < n_data ? i0 + chunk : n_data;
    double acc = 0.0;
    for (int64_t t0 = i0; t0 < i1; t0 += 256) {
        const int cnt = (int)(i1 - t0 < 256 ? i1 - t0 : 256);
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
                const double dv = g - mu;
                acc = fma(dv, dv, acc);
            }
        }
    }
    if (j < n_src) part[(int64_t)blockIdx.y * n_src + j] = acc;
}

// mean_j (pass 0) or the population std s_j (pass 1; 1 where it is 0) of
// column j of J from its split sums.
__global__ void col_moment_finish_kernel(int pass, const double *__restrict__ part,
                                         int64_t splits, int64_t n_src, int64_t n_data,
                                         double *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_src) return;
    double sum = 0.0;
    for (int64_t y = 0; y < splits; ++y) sum += part[y * n_src + j];
    if (pass == 0) {
        out[j] = sum / (double)n_data;
    } else {
        const double sd = sqrt(sum / (double)n_data);
        out[j] = sd == 0.0 ? 1.0 : sd;
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

// G[j * n_data + i] = sqrt(w_i) * J_ij / s_j: the scaled jacobian
// column-major (columns contiguous), for the SVD route.
__global__ __launch_bounds__(256) void scaled_jacobian_cm_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    const double *__restrict__ w, int64_t n_data, const double *__restrict__ pe,
    const double *__restrict__ pn, const double *__restrict__ pu,
    const double *__restrict__ scale, int64_t n_src, double *__restrict__ G) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_data) return;
    const double oe = e[i], on = n[i], ou = u[i], sw = w ? sqrt(w[i]) : 1.0;
    for (int64_t j = blockIdx.y; j < n_src; j += gridDim.y)
        G[j * n_data + i] = sw * inv_dist(oe, on, ou, pe[j], pn[j], pu[j]) * (1.0 / scale[j]);
}

// y_i = sqrt(w_i) d_i (w == NULL -> d).
__global__ void sqrtw_times_kernel(int64_t n, const double *__restrict__ w,
                                   const double *__restrict__ d, double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = w ? sqrt(w[i]) * d[i] : d[i];
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

// Rows [k0, k0 + kKC) x columns [col0, col0 + kST) of a row-major K x cols
// panel, fetched 8 consecutive columns per thread (16-byte loads when the
// panel is aligned, zero outside the panel).
template <class T>
struct Panel {
    const T *X;
    int64_t ld, rows, cols;
    bool vec;  // X and ld * sizeof(T) are multiples of 16 bytes
    // Unchecked 16-byte loads: the caller guarantees the whole kKC x kST
    // block is inside the panel and vec is set.
    __device__ __forceinline__ void fetch_fast(int64_t k0, int64_t col0, int kk, int c8,
                                               T (&v)[8]) const {
        const T *p = X + (k0 + kk) * ld + col0 + 8 * c8;
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double2 d = reinterpret_cast<const double2 *>(p)[i];
                v[2 * i] = d.x;
                v[2 * i + 1] = d.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const float4 f = reinterpret_cast<const float4 *>(p)[i];
                v[4 * i] = f.x;
                v[4 * i + 1] = f.y;
                v[4 * i + 2] = f.z;
                v[4 * i + 3] = f.w;
            }
        }
    }
    __device__ __forceinline__ void fetch(int64_t k0, int64_t col0, int kk, int c8,
                                          T (&v)[8]) const {
        const int64_t k = k0 + kk, c = col0 + 8 * c8;
        if (k < rows && c + 8 <= cols && vec) {
            const T *p = X + k * ld + c;
            if constexpr (sizeof(T) == 8) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double2 d = reinterpret_cast<const double2 *>(p)[i];
                    v[2 * i] = d.x;
                    v[2 * i + 1] = d.y;
                }
            } else {
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const float4 f = reinterpret_cast<const float4 *>(p)[i];
                    v[4 * i] = f.x;
                    v[4 * i + 1] = f.y;
                    v[4 * i + 2] = f.z;
                    v[4 * i + 3] = f.w;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (k < rows && c + i < cols) ? X[k * ld + c + i] : T(0);
        }
    }
};

template <class T>
Panel<T> make_panel(const T *X, int64_t ld, int64_t rows, int64_t cols) {
    const bool vec = ((uintptr_t)X % 16 == 0) && (((size_t)ld * sizeof(T)) % 16 == 0);
    return Panel<T>{X, ld, rows, cols, vec};
}

// The K loop of one kST x kST tile (see gemm_tn_kernel).  FAST: every chunk
// is inside both panels and they are aligned, so the loads are unchecked
// 16-byte loads; the loop body is straight-line code either way, which
// keeps the accumulators in place across iterations.
template <class T, bool FAST, bool UPPER>
__device__ __forceinline__ void tile_k_loop(const Panel<T> &A, const Panel<T> &B, int64_t K,
                                            int64_t r0, int64_t c0, T (*lds)[2][kKC * kPad],
                                            int64_t M, int64_t N, double *C, int64_t ldc,
                                            double beta, double sign, double shift) {
    using acc_t = typename Mfma<T>::acc_t;
    constexpr bool kFold = !std::is_same<T, double>::value;
    acc_t acc[4][4];
    double4_t dacc[kFold ? 4 : 1][kFold ? 4 : 1];
    // FP64 accumulators are carried across the whole K loop; their zeros
    // come out of the matrix unit itself so they live in the MFMA result
    // registers from the start (a plain zero would be a VGPR value and cost
    // a copy per iteration).  FP32 accumulators restart every chunk.
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            acc[a][c] = kFold ? acc_t{0, 0, 0, 0}
                              : Mfma<T>::step(T(0), T(0), acc_t{0, 0, 0, 0});
    if constexpr (kFold) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) dacc[a][c] = double4_t{0, 0, 0, 0};
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int kk = threadIdx.x >> 4, c8 = threadIdx.x & 15;  // staging slot
    T ra[8], rc[8];
    auto fetch = [&](int64_t k0) {
        if constexpr (FAST) {
            A.fetch_fast(k0, r0, kk, c8, ra);
            B.fetch_fast(k0, c0, kk, c8, rc);
        } else {
            A.fetch(k0, r0, kk, c8, ra);
            B.fetch(k0, c0, kk, c8, rc);
        }
    };
    auto store = [&](int buf) {
        T *pa = &lds[buf][0][kk * kPad + 8 * c8];
        T *pb = &lds[buf][1][kk * kPad + 8 * c8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            pa[i] = ra[i];
            pb[i] = rc[i];
        }
    };
    // Chunk count as a wave-uniform 32-bit value (K < 2^31 * kKC), so the
    // loop control stays scalar.
    const int chunks = __builtin_amdgcn_readfirstlane((int)((K + kKC - 1) / kKC));
    if (chunks > 0) {
        fetch(0);
        store(0);
    }
    __syncthreads();
    for (int ch = 0; ch < chunks; ++ch) {
        const int cur = ch & 1;
        const bool more = ch + 1 < chunks;
        if (more) fetch((int64_t)(ch + 1) * kKC);
        const T *pr = lds[cur][0];
        const T *pc = lds[cur][1];
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
        if (more) store(cur ^ 1);
        __syncthreads();
    }
    // Epilogue (kept in this instantiation, so the accumulators never merge
    // across the two loop versions), one row of 16 x 16 tiles at a time: with
    // beta != 0 the row's 16 old values of C are loaded first, from in-range
    // (clamped) addresses without branches, so their loads overlap, while
    // only a quarter of them is held at once.
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
        double old[4][4];
        if (beta != 0.0) {
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    int64_t r = r0 + 64 * wr + 16 * ti + Mfma<T>::row(lane, v);
                    int64_t c = c0 + 64 * wc + 16 * tj + (lane & 15);
                    r = r < M ? r : M - 1;
                    c = c < N ? c : N - 1;
                    old[tj][v] = C[r * ldc + c];
                }
        }
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
                    C[r * ldc + c] = beta == 0.0 ? val : fma(beta, old[tj][v], val);
                }
            }
    }
}

// Upper-triangle tiles in an XCD-aware order, for the bf16 split-product
// kernel, whose inputs stream from beyond L2 (the FP64 GEMM, bound by its
// MFMA rate, runs 3% slower in this order and keeps the row-major one).
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2, and the 32 that one XCD runs at a time are consecutive in that XCD's
// share of the ids; they take one super-block of 4 row tiles x 8 column
// tiles, so the panels they stream through K are shared in that L2 (12
// panels per super-block instead of 33 for a row of 32 tiles).
// Super-blocks that reach the upper triangle run row-major; tiles of a
// super-block below the diagonal or past the edge get no work (false).
constexpr int kXcds = 8, kSbRows = 4, kSbCols = 8;

__host__ __device__ inline int64_t superblock_first_col(int64_t sr) {
    return sr * kSbRows / kSbCols;  // first super-column that reaches the diagonal
}

__device__ __forceinline__ bool upper_tile_xcd(int64_t id, int64_t tiles, int64_t &rb,
                                               int64_t &cb) {
    const int64_t x = id % kXcds, q = id / kXcds;
    int64_t S = kXcds * (q / (kSbRows * kSbCols)) + x;
    const int64_t p = q % (kSbRows * kSbCols);
    const int64_t C = (tiles + kSbCols - 1) / kSbCols;
    int64_t sr = 0;
    for (;;) {
        const int64_t cnt = C - superblock_first_col(sr);
        if (cnt <= 0) return false;  // past the last super-block
        if (S < cnt) break;
        S -= cnt;
        ++sr;
    }
    rb = sr * kSbRows + p / kSbCols;
    cb = (superblock_first_col(sr) + S) * kSbCols + p % kSbCols;
    return rb < tiles && cb < tiles && cb >= rb;
}

// Workgroups for upper_tile_xcd over `tiles` x `tiles` tiles.
inline unsigned upper_tile_xcd_blocks(int64_t tiles) {
    const int64_t R = (tiles + kSbRows - 1) / kSbRows, C = (tiles + kSbCols - 1) / kSbCols;
    int64_t n = 0;
    for (int64_t sr = 0; sr < R; ++sr) n += C - superblock_first_col(sr);
    n = (n + kXcds - 1) / kXcds * kXcds;
    return (unsigned)(n * kSbRows * kSbCols);
}

// C = beta * C + sign * A^T B + shift * I on the selected tiles, with A
// (K x M) and B (K x N) row-major panels and C (M x N) row-major (ldc).
// UPPER: M == N and only the upper triangle (r <= c) is computed (the
// symmetric rank-K update A^T A when B == A).  A workgroup of 4 wavefronts
// owns a kST x kST tile of C (2 x 2 waves of 64 x 64, each 4 x 4 MFMA
// 16x16x4 tiles).  The K dimension goes through two LDS buffers in chunks
// of kKC rows: the global loads of chunk c + 1 are issued before the MFMAs
// of chunk c and written to the other buffer after them, so their latency
// overlaps the matrix work and one barrier per chunk suffices.  Tiles fully
// inside aligned panels (with K a multiple of kKC) run the unchecked loop.
// FP32 MFMA accumulators are folded into FP64 after every chunk (FP32
// inputs, FP64 sums).  Every workgroup reads all its inputs before it
// writes its tile, and tiles are disjoint, so C may alias B when M <= kST
// (used by the in-place panel solve of the Cholesky factorisation).
template <class T, bool UPPER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void gemm_tn_kernel(Panel<T> A, Panel<T> B, int64_t K,
                                                      int64_t M, int64_t N, double *C, int64_t ldc,
                                                      double beta, double sign, double shift,
                                                      int64_t tiles_n, int64_t bsa, int64_t bsb,
                                                      int64_t bsc,
                                                      const int64_t *__restrict__ dims) {
    __shared__ T lds[2][2][kKC * kPad];
    // problem blockIdx.z of a strided batch (element strides of A, B, C)
    A.X += (int64_t)blockIdx.z * bsa;
    B.X += (int64_t)blockIdx.z * bsb;
    C += (int64_t)blockIdx.z * bsc;
    int64_t rb, cb;
    if (UPPER) {  // row-major over the upper triangle (FP64 MFMA is compute-bound)
        int64_t bq = blockIdx.x, rem = tiles_n;
        rb = 0;
        while (bq >= rem) {
            bq -= rem;
            ++rb;
            --rem;
        }
        cb = rb + bq;
    } else {
        rb = blockIdx.x / tiles_n;
        cb = blockIdx.x - rb * tiles_n;
    }
    const int64_t r0 = rb * kST, c0 = cb * kST;
    if (dims) {
        // Problem z uses only its first dims[2z] rows (the rest of its panels
        // are zero rows) and its first dims[2z + 1] rows and columns of C (the
        // rest is zero, plus the shift on the diagonal): the K loop stops at
        // its own rows, and tiles wholly past its size run no K loop at all.
        const int64_t kz = dims[2 * blockIdx.z], mz = dims[2 * blockIdx.z + 1];
        const int64_t kr = (kz + kKC - 1) / kKC * kKC;
        K = (r0 < mz && c0 < mz) ? (kr < K ? kr : K) : 0;
    }
    const bool fast = A.vec && B.vec && r0 + kST <= M && c0 + kST <= N && K % kKC == 0;
    if (fast)
        tile_k_loop<T, true, UPPER>(A, B, K, r0, c0, lds, M, N, C, ldc, beta, sign, shift);
    else
        tile_k_loop<T, false, UPPER>(A, B, K, r0, c0, lds, M, N, C, ldc, beta, sign, shift);
}

// batch > 1: a strided batch of equal-shaped problems (element strides
// bsa, bsb, bsc between consecutive A, B, C; multiples of 2 keep the
// 16-byte loads aligned).
template <class T>
int launch_gemm(const T *A, int64_t lda, const T *B, int64_t ldb, int64_t K, int64_t M,
                int64_t N, double *C, int64_t ldc, double beta, double sign, double shift,
                bool upper, hipStream_t s, int batch = 1, int64_t bsa = 0, int64_t bsb = 0,
                int64_t bsc = 0, const int64_t *dims = nullptr) {
    if (M <= 0 || N <= 0 || batch <= 0) return HMX_OK;
    const int64_t tm = (M + kST - 1) / kST, tn = (N + kST - 1) / kST;
    const Panel<T> pa = make_panel(A, lda, K, M), pb = make_panel(B, ldb, K, N);
    if (upper) {
        hipLaunchKernelGGL((gemm_tn_kernel<T, true>),
                           dim3((unsigned)(tn * (tn + 1) / 2), 1, (unsigned)batch), dim3(256), 0,
                           s, pa, pb, K, M, N, C, ldc, beta, sign, shift, tn, bsa, bsb, bsc, dims);
        HMX_LAUNCHED("gemm_tn_kernel<upper>");
    } else {
        hipLaunchKernelGGL((gemm_tn_kernel<T, false>), dim3((unsigned)(tm * tn), 1, (unsigned)batch),
                           dim3(256), 0, s, pa, pb, K, M, N, C, ldc, beta, sign, shift, tn, bsa,
                           bsb, bsc, dims);
        HMX_LAUNCHED("gemm_tn_kernel");
    }
    return HMX_OK;
}

// dims (device, 2 per problem, or NULL): the rows and the size each problem
// of the batch really has (see gemm_tn_kernel).
template <class T>
int launch_syrk(const T *X, int64_t ldx, int64_t K, int64_t m, double *C, int64_t ldc, double beta,
                double sign, double shift, hipStream_t s, int batch = 1, int64_t bsx = 0,
                int64_t bsc = 0, const int64_t *dims = nullptr) {
    return launch_gemm<T>(X, ldx, X, ldx, K, m, m, C, ldc, beta, sign, shift, true, s, batch, bsx,
                          bsx, bsc, dims);
}

int launch_gemm_tn(const double *A, int64_t lda, const double *B, int64_t ldb, int64_t K,
                   int64_t M, int64_t N, double *C, int64_t ldc, double beta, double sign,
                   hipStream_t s, int batch = 1, int64_t bsa = 0, int64_t bsb = 0,
                   int64_t bsc = 0) {
    return launch_gemm<double>(A, lda, B, ldb, K, M, N, C, ldc, beta, sign, 0.0, false, s, batch,
                               bsa, bsb, bsc);
}

// ---------------------------------------------------------------- Cholesky
// v from another lane of the same group of 8 aligned lanes, by a DPP move
// of its two 32-bit halves (kCtrl: quad_perm [1,0,3,2] = 0xb1 for lane ^ 1,
// quad_perm [2,3,0,1] = 0x4e for lane ^ 2, row_half_mirror = 0x141 for
// lane 7 - i of the group); a VALU operation, no trip through the LDS unit.
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, kCtrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), kCtrl, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over each group of 8 aligned lanes, in every lane of the group: the
// pair and quad sums, then the other quad's sum (the same value in its four
// lanes), so the result is the same as with xor shuffles 1, 2, 4.
__device__ __forceinline__ double sum8_f64(double v) {
    v += dpp_f64<0xb1>(v);
    v += dpp_f64<0x4e>(v);
    v += dpp_f64<0x141>(v);
    return v;
}

// Factor the nb x nb diagonal block at D (upper, row-major, ld) in place as
// U^T U and write U^-1 to Dinv (nb x nb, row-major, ld kNB) and its
// transpose to DinvT.  One workgroup; the block lives in LDS.  On a
// non-positive pivot, *info is set (once) to its 1-based global column and
// the block is left unfinished.
__global__ __launch_bounds__(1024) void potrf_diag_kernel(double *__restrict__ D, int64_t ld,
                                                          int nb, int64_t col0,
                                                          double *__restrict__ Dinv,
                                                          double *__restrict__ DinvT,
                                                          int *__restrict__ info, int64_t bsd,
                                                          int64_t bsinv, int info_stride) {
    constexpr int S = kNB + 1;
    // problem blockIdx.x of a strided batch
    D += (int64_t)blockIdx.x * bsd;
    Dinv += (int64_t)blockIdx.x * bsinv;
    DinvT += (int64_t)blockIdx.x * bsinv;
    info += (int64_t)blockIdx.x * info_stride;
    __shared__ double L[kNB * S];
    __shared__ double piv[kNB];
    const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5;  // 32 x 32 threads
    // upper triangle of the block; the rest of L is zero (the inverse below
    // multiplies it by zeros, and it may not hold finite values in D)
    for (int r = ty; r < kNB; r += 32)
        for (int c = tx; c < kNB; c += 32)
            L[r * S + c] = (r < nb && c < nb && r <= c) ? D[r * ld + c] : 0.0;
    __syncthreads();
    // Right-looking factorisation, one barrier per column: every thread
    // reads the pivot d (the same value, so the test is uniform) and updates
    // its part of the trailing upper triangle by L[k][r] L[k][c] / d; row k
    // itself is read only in step k and is scaled by its pivot at the end.
    for (int k = 0; k < nb; ++k) {
        const double d = L[k * S + k];
        if (!(d > 0.0)) {
            if (tid == 0) atomicCAS(info, 0, (int)(col0 + k + 1));
            return;
        }
        if (tid == 0) piv[k] = sqrt(d);
        const double invd = 1.0 / d;
        for (int r = k + 1 + ty; r < nb; r += 32) {
            const double lr = L[k * S + r] * invd;
            for (int c = r + tx; c < nb; c += 32) L[r * S + c] = fma(-lr, L[k * S + c], L[r * S + c]);
        }
        __syncthreads();
    }
    for (int r = ty; r < nb; r += 32) {
        const double inv = 1.0 / piv[r];
        for (int c = r + tx; c < nb; c += 32) L[r * S + c] = c == r ? piv[r] : L[r * S + c] * inv;
    }
    __syncthreads();
    for (int r = ty; r < nb; r += 32)
        for (int c = tx; c < nb; c += 32) D[r * ld + c] = r <= c ? L[r * S + c] : 0.0;
    // reciprocals of the factor's diagonal, once (they multiply below)
    if (tid < nb) piv[tid] = 1.0 / L[tid * S + tid];
    __syncthreads();
    // Inverse X = U^-1 row by row, from X U = I:
    //   X[i][i] = 1 / U[i][i],  X[i][j] = -(sum_{k=i}^{j-1} X[i][k] U[k][j]) (1 / U[j][j]).
    // Rows are independent, so there is no barrier: row i belongs to the 8
    // adjacent lanes of its group, lane p keeping X[i][8 q + p] in xr[q]
    // (zero until computed, and for k < i), each sum split over the lanes
    // and combined by DPP moves.  Wavefront w holds rows 8 w .. 8 w + 7,
    // whose terms start at k = 8 w, so its loops skip q < w uniformly.
    {
        const int i = tid >> 3, p = tid & 7, w = tid >> 6;
        const bool live = i < nb;
        double xr[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) xr[q] = (live && 8 * q + p == i) ? piv[i] : 0.0;
#pragma unroll
        for (int jb = 0; jb < 16; ++jb) {
            if (jb < w) continue;  // uniform in the wavefront
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const int j = 8 * jb + jj;
                double acc0 = 0.0, acc1 = 0.0;  // two chains of FMAs
#pragma unroll
                for (int q = 0; q <= jb; ++q) {
                    if (q < w) continue;  // uniform in the wavefront
                    if (q & 1)
                        acc1 = fma(xr[q], L[(8 * q + p) * S + j], acc1);
                    else
                        acc0 = fma(xr[q], L[(8 * q + p) * S + j], acc0);
                }
                const double acc = sum8_f64(acc0 + acc1);
                if (live && j > i && j < nb && p == jj) xr[jb] = -acc * piv[j];
            }
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int k = 8 * q + p;
            const double v = (live && k < nb && k >= i) ? xr[q] : 0.0;
            Dinv[i * kNB + k] = v;
            DinvT[k * kNB + i] = v;
        }
    }
}

// ------------------------------------------------------------------ solves
// One-launch blocked substitutions.  Workgroup w of the forward solve
// (U^T y = r) owns block b = w of y; workgroup w of the backward solve
// (U x = y) owns block b = nb - 1 - w of x.  A block needs the solved
// blocks before it (forward) or after it (backward), which belong to
// workgroups of lower index.  The output vector is filled with a sentinel
// before the launch (kSolveSentinel, a signalling-NaN pattern that no
// arithmetic produces: NaN results are quiet NaNs), every solved entry is
// written with an agent-scope store, and a consumer lane polls the entries
// it needs with agent-scope loads until none is the sentinel: the poll
// returns the values themselves, so a dependency costs one round trip to
// the device-coherent level instead of a flag and then a load.  Each
// workgroup accumulates the product of a dependency with the block of U it
// prefetched into registers before polling; the dependencies are visited
// oldest first, so the wait that matters (the block just before) comes
// last, and the block's own triangle goes through its stored inverse.
// Workgroups of a launch are dispatched in index order on each XCD, so the
// lowest unfinished workgroup always has its dependencies done and the
// solve makes progress; every wait is still bounded (kTrsvSpinLimit polls)
// and a timed-out wait sets *err and goes on (with 0 for the value), so
// every wavefront ends.
constexpr int kTrsvSpinLimit = 1 << 22;
constexpr long long kSolveSentinel = 0x7ff4dead5eed0001LL;

__device__ __forceinline__ double load_agent(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_agent(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool is_sentinel(double v) {
    return __double_as_longlong(v) == kSolveSentinel;
}

// Starts the copy of a kNB x kNB block (row-major, 128 KB) into LDS
// straight from global memory (16 bytes per lane, 1 KB runs per wave); the
// copy is complete once every wave has waited for its vector memory
// operations and the workgroup has met at a barrier.  1024 threads.
__device__ __forceinline__ void block_to_lds(const double *__restrict__ src, double *dst) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int run = 8 * w + i;  // 128 runs of 128 doubles
        __builtin_amdgcn_global_load_lds((const void *)(src + run * kNB + 2 * l),
                                         (__attribute__((address_space(3))) void *)(dst + run * kNB),
                                         16, 0, 0);
    }
}

// The solved entry *p, polled until it is not the sentinel.
__device__ __forceinline__ double wait_value(const double *p, int *err) {
    double v = load_agent(p);
    int spins = 0;
    while (is_sentinel(v)) {
        if (++spins > kTrsvSpinLimit) {
            atomicOr(err, 1);
            return 0.0;
        }
        __builtin_amdgcn_s_sleep(1);
        v = load_agent(p);
    }
    return v;
}

// The solved entries *p0 (if v0) and *p1 (if v1), polled together until
// neither is the sentinel (0 for an entry not wanted).
__device__ __forceinline__ void wait_pair(const double *p0, bool v0, const double *p1, bool v1,
                                          int *err, double &x0, double &x1) {
    x0 = v0 ? load_agent(p0) : 0.0;
    x1 = v1 ? load_agent(p1) : 0.0;
    int spins = 0;
    while (is_sentinel(x0) || is_sentinel(x1)) {
        if (++spins > kTrsvSpinLimit) {
            atomicOr(err, 1);
            x0 = 0.0;
            x1 = 0.0;
            return;
        }
        __builtin_amdgcn_s_sleep(1);
        if (is_sentinel(x0)) x0 = load_agent(p0);
        if (is_sentinel(x1)) x1 = load_agent(p1);
    }
}

// Position of block (i, j), j > i, among the blocks above the diagonal of
// an nb x nb block matrix, stored row by row.
__host__ __device__ inline int64_t upper_block_index(int64_t i, int64_t j, int64_t nb) {
    return i * (nb - 1) - i * (i - 1) / 2 + (j - i - 1);
}

// The factor's blocks above the diagonal in FP32, each kNB x kNB block
// contiguous and row-major (zero past m), block (i, j) at
// upper_block_index(i, j, nb) * kNB^2: the substitutions of the mixed fits,
// where the factor is only a preconditioner, read these (half the bytes,
// in streams of 64 KB) instead of its FP64 rows (1 KB runs).
__global__ __launch_bounds__(256) void pack_upper_f32_kernel(const double *__restrict__ A,
                                                             int64_t ld, int64_t m, int nb,
                                                             float *__restrict__ Ub) {
    const int j = blockIdx.x, i = blockIdx.y;
    if (j <= i) return;
    float *dst = Ub + upper_block_index(i, j, nb) * kNB * kNB;
    const int c = threadIdx.x & (kNB - 1);
    const int64_t col = (int64_t)j * kNB + c;
    for (int rr = threadIdx.x >> 7; rr < kNB; rr += 2) {
        const int64_t row = (int64_t)i * kNB + rr;
        dst[rr * kNB + c] = (row < m && col < m) ? (float)A[row * ld + col] : 0.0f;
    }
}

// Forward: y_b = Dinv_b^T (r_b - sum_{k < b} U_kb^T y_k), U_kb = rows of
// block k, columns of block b.  Thread (g, t) = (tid / 128, tid % 128)
// takes column t of the block and rows 16 g .. 16 g + 15 of every block k
// (lanes 0-15 of its wavefront poll those 16 entries of y_k).  kBlocked:
// the blocks U_kb come from the FP32 copy Ub (pack_upper_f32_kernel).
template <bool kBlocked>
__global__ __launch_bounds__(1024) void trsv_fwd_sync_kernel(
    const double *__restrict__ U, const float *__restrict__ Ub, int64_t ld, int64_t m,
    const double *__restrict__ Dinv, const double *__restrict__ r, double *__restrict__ y,
    int *__restrict__ err) {
    __shared__ double dinv[kNB * kNB];  // Dinv_b, copied during the dependency loop
    __shared__ double part[8][kNB];
    __shared__ double z[kNB];
    const int b = blockIdx.x;
    const int t = threadIdx.x & (kNB - 1), g = threadIdx.x >> 7, lane = threadIdx.x & 63;
    const int64_t c0 = (int64_t)b * kNB, c = c0 + t;
    const bool cv = c < m;
    block_to_lds(Dinv + (int64_t)b * kNB * kNB, dinv);
    // the right-hand side entry, loaded now rather than on the chain
    const double rc = (g == 0 && cv) ? r[c] : 0.0;
    double u[16];
    // Columns past m (last block only) read column m - 1 (FP64) or the
    // copy's zero padding: finite values whose sums are dropped (z = 0 there).
    const int64_t nbk = (m + kNB - 1) / kNB;
    auto load_u = [&](int k) {
        if constexpr (kBlocked) {
            const float *blk = Ub + upper_block_index(k, b, nbk) * kNB * kNB + (16 * g) * kNB + t;
#pragma unroll
            for (int q = 0; q < 16; ++q) u[q] = (double)blk[q * kNB];
        } else {
            const double *col = U + ((int64_t)k * kNB + 16 * g) * ld + (cv ? c : m - 1);
#pragma unroll
            for (int q = 0; q < 16; ++q) u[q] = col[q * ld];
        }
    };
    double acc = 0.0;
    if (b > 0) load_u(0);
    for (int k = 0; k < b; ++k) {
        // blocks before b are whole (only the last block may be partial)
        const double yv = lane < 16 ? wait_value(y + (int64_t)k * kNB + 16 * g + lane, err) : 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) acc = fma(u[q], __shfl(yv, q, 64), acc);
        if (k + 1 < b) load_u(k + 1);
    }
    part[g][t] = acc;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's part of the Dinv_b copy
    __syncthreads();
    if (g == 0) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) s += part[q][t];
        z[t] = cv ? rc - s : 0.0;
    }
    __syncthreads();
    double d = 0.0;  // (Dinv_b^T z)[t] = sum_{l <= t} Dinv_b[l][t] z[l]
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int l = 16 * g + q;
        d = l <= t ? fma(dinv[l * kNB + t], z[l], d) : d;
    }
    part[g][t] = d;
    __syncthreads();
    if (g == 0 && cv) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) s += part[q][t];
        store_agent(y + c, s);
    }
}

// Backward: x_b = Dinv_b (y_b - sum_{k > b} U_bk x_k).  Wavefront v of the
// workgroup takes rows 8 v .. 8 v + 7 of the block, lane l columns l and
// l + 64 of every block k (partial sums per lane, added across the lanes
// once at the end); lane l polls entries l and l + 64 of x_k.  kBlocked:
// the blocks U_bk come from the FP32 copy Ub.
template <bool kBlocked>
__global__ __launch_bounds__(1024) void trsv_bwd_sync_kernel(
    const double *__restrict__ U, const float *__restrict__ Ub, int64_t ld, int64_t m,
    const double *__restrict__ DinvT, const double *__restrict__ y, double *__restrict__ x,
    int *__restrict__ err) {
    __shared__ double dinvt[kNB * kNB];  // DinvT_b, copied during the dependency loop
    __shared__ double part[8][kNB];
    __shared__ double z[kNB];
    const int nb = (int)((m + kNB - 1) / kNB);
    const int b = nb - 1 - (int)blockIdx.x;
    const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
    const int t = threadIdx.x & (kNB - 1), g = threadIdx.x >> 7;
    const int64_t r0 = (int64_t)b * kNB;
    block_to_lds(DinvT + (int64_t)b * kNB * kNB, dinvt);
    // Lanes 8r of wavefront v finish row 8 v + r of the block (see the sum
    // over lanes below); its right-hand side entry, loaded now rather than
    // on the chain.
    const int64_t my_row = r0 + 8 * v + (lane >> 3);
    const double yr = ((lane & 7) == 0 && my_row < m) ? y[my_row] : 0.0;
    double u[16];
    // Rows and columns past m read row or column m - 1: finite values that
    // meet x = 0 (columns) or whose sums are dropped (rows, z = 0 there).
    auto load_u = [&](int k) {
        if constexpr (kBlocked) {  // blocks right of the diagonal: rows all below m
            const float *blk = Ub + upper_block_index(b, k, nb) * kNB * kNB + (8 * v) * kNB + lane;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                u[2 * i] = (double)blk[i * kNB];
                u[2 * i + 1] = (double)blk[i * kNB + 64];
            }
        } else {
            const int64_t cb = (int64_t)k * kNB;
            const int64_t c0 = cb + lane < m ? cb + lane : m - 1;
            const int64_t c1 = cb + lane + 64 < m ? cb + lane + 64 : m - 1;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int64_t row = r0 + 8 * v + i;
                const double *p = U + (row < m ? row : m - 1) * ld;
                u[2 * i] = p[c0];
                u[2 * i + 1] = p[c1];
            }
        }
    };
    double acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0.0;
    if (b + 1 < nb) load_u(nb - 1);
    for (int k = nb - 1; k > b; --k) {
        // entries past m are never written: not polled
        const int64_t cb = (int64_t)k * kNB;
        double x0, x1;
        wait_pair(x + cb + lane, cb + lane < m, x + cb + lane + 64, cb + lane + 64 < m, err, x0,
                  x1);
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = fma(u[2 * i + 1], x1, fma(u[2 * i], x0, acc[i]));
        if (k - 1 > b) load_u(k - 1);
    }
    // Sums over the 64 lanes of the eight rows' partial sums, reduce-scatter
    // style: three exchange-and-add steps (lane offsets 32, 16, 8) halve the
    // rows a lane holds, keeping row 4 b5 + 2 b4 + b3 (b the lane's bits),
    // then three steps add within groups of 8 lanes, so that lanes 8r..8r+7
    // hold row r's sum: 10 shuffles on the chain instead of 48.
    double a4[4], a2[2];
    {
        const bool hi = lane & 32;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a4[i] = (hi ? acc[4 + i] : acc[i]) + __shfl_xor(hi ? acc[i] : acc[4 + i], 32, 64);
    }
    {
        const bool hi = lane & 16;
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a2[i] = (hi ? a4[2 + i] : a4[i]) + __shfl_xor(hi ? a4[i] : a4[2 + i], 16, 64);
    }
    const bool hi8 = lane & 8;
    double s = (hi8 ? a2[1] : a2[0]) + __shfl_xor(hi8 ? a2[0] : a2[1], 8, 64);
    s += __shfl_xor(s, 4, 64);
    s += __shfl_xor(s, 2, 64);
    s += __shfl_xor(s, 1, 64);
    if ((lane & 7) == 0) z[8 * v + (lane >> 3)] = my_row < m ? yr - s : 0.0;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's part of the DinvT_b copy
    __syncthreads();
    double d = 0.0;  // (Dinv_b z)[t] = sum_{l >= t} DinvT_b[l][t] z[l]
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int l = 16 * g + q;
        d = l >= t ? fma(dinvt[l * kNB + t], z[l], d) : d;
    }
    part[g][t] = d;
    __syncthreads();
    if (g == 0 && r0 + t < m) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) s += part[q][t];
        store_agent(x + r0 + t, s);
    }
}

// p[0, n) = the solves' sentinel.
__global__ void fill_sentinel_kernel(double *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = __longlong_as_double(kSolveSentinel);
}

__global__ void fill_int_kernel(int *__restrict__ p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------ matrix-free A p
// t_i = sum_j J_ij p_j / s_j (the unweighted data-space image of p): the
// split's sources [y chunk, (y + 1) chunk) into part[y * n_data + i].
__global__ __launch_bounds__(256) void j_apply_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    int64_t n_data, const double *__restrict__ pe, const double *__restrict__ pn,
    const double *__restrict__ pu, const double *__restrict__ p,
    const double *__restrict__ scale, int64_t n_src, int64_t chunk, double *__restrict__ part) {
    __shared__ double4_t tile[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t ii = i < n_data ? i : n_data - 1;
    const double oe = e[ii], on = n[ii], ou = u[ii];
    const int64_t j0 = (int64_t)blockIdx.y * chunk;
    const int64_t j1 = j0 + chunk < n_src ? j0 + chunk : n_src;
    double acc = 0.0;
    for (int64_t t0 = j0; t0 < j1; t0 += 256) {
        const int cnt = (int)(j1 - t0 < 256 ? j1 - t0 : 256);
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
    if (i < n_data) part[(int64_t)blockIdx.y * n_data + i] = acc;
}

// t_i = sum over the splits, in order, of part[y * n_data + i].
__global__ void j_apply_finish_kernel(const double *__restrict__ part, int64_t splits,
                                      int64_t n_data, double *__restrict__ t) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_data) return;
    double sum = 0.0;
    for (int64_t y = 0; y < splits; ++y) sum += part[y * n_data + i];
    t[i] = sum;
}

// sum_i J_ij w_i t_i (w may be NULL) over the split's data rows
// [y chunk, (y + 1) chunk) into part[y * n_src + j].
__global__ __launch_bounds__(256) void jt_apply_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    const double *__restrict__ w, int64_t n_data, const double *__restrict__ pe,
    const double *__restrict__ pn, const double *__restrict__ pu, const double *__restrict__ t,
    int64_t n_src, int64_t chunk, double *__restrict__ part) {
    __shared__ double4_t tile[256];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t jj = j < n_src ? j : n_src - 1;
    const double se = pe[jj], sn = pn[jj], su = pu[jj];
    const int64_t i0 = (int64_t)blockIdx.y * chunk;
    const int64_t i1 = i0 + chunk < n_data ? i0 + chunk : n_data;
    double acc = 0.0;
    for (int64_t t0 = i0; t0 < i1; t0 += 256) {
        const int cnt = (int)(i1 - t0 < 256 ? i1 - t0 : 256);
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
    if (j < n_src) part[(int64_t)blockIdx.y * n_src + j] = acc;
}

// out_j = (1 / s_j) (sum over the splits, in order, of part[y * n_src + j])
// + alpha p_j  (p may be NULL).
__global__ void jt_apply_finish_kernel(const double *__restrict__ part, int64_t splits,
                                       const double *__restrict__ scale, int64_t n_src,
                                       double alpha, const double *__restrict__ p,
                                       double *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_src) return;
    double sum = 0.0;
    for (int64_t y = 0; y < splits; ++y) sum += part[y * n_src + j];
    double v = sum / scale[j];
    if (p) v = fma(alpha, p[j], v);
    out[j] = v;
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

__global__ void fill_zero_f64_kernel(double *__restrict__ x, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

__global__ void copy_kernel(int64_t n, const double *__restrict__ x, double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

// Partial dot products / max-abs / min-abs (mode 0 / 1 / 2): one value per
// workgroup, combined on the host in workgroup order (deterministic).
__global__ __launch_bounds__(256) void reduce_partial_kernel(int mode, int64_t n, int64_t stride,
                                                             const double *__restrict__ x,
                                                             const double *__restrict__ y,
                                                             double *__restrict__ part) {
    __shared__ double red[256];
    double acc = mode == 2 ? INFINITY : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double v = x[i * stride];
        acc = mode == 0 ? fma(v, y[i], acc) : mode == 1 ? fmax(acc, fabs(v)) : fmin(acc, fabs(v));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            red[threadIdx.x] = mode == 0   ? red[threadIdx.x] + red[threadIdx.x + s]
                               : mode == 1 ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                           : fmin(red[threadIdx.x], red[threadIdx.x + s]);
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

// A[j][j] += tau for j < m (row-major, leading dimension m).
__global__ void add_diag_kernel(double *__restrict__ A, int64_t m, double tau) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) A[j * m + j] += tau;
}


// ------------------------------------------ split-precision normal matrix
// Exact-product schemes for X^T X on low-precision MFMA, in two slice
// formats (bf16 and int8, described below) that share the row groups, the
// group scales and the gram kernel.
//
// bf16 slices.  The rows are cut into
// groups G of kFoldRows; in each group, column j of X = W^1/2 J S is scaled
// by f_Gj = 2^(E + 1), where max_{k in G} |X_kj| < 2^E, so x = X_kj / f_Gj
// lies in (-1/2, 1/2), and cut into kSlices integer slices:
// sigma_0 = rint(256 x), sigma_1 = rint(256 (256 x - sigma_0)), ... with
// |sigma_t| <= 128, so that
//   X_kj = f_Gj sum_t sigma_t,kj 256^-(t+1)   (error below f_Gj 2^-33).
// Every sigma is an integer of at most 8 bits (exact in bf16) and every
// product of two has at most 14 bits, so the MFMA sums of one order over
// one group's 1024 rows are exact in FP32.  The normal matrix
//   (X^T X)_ij = sum_G f_Gi f_Gj sum_o 256^-(o+2) sum_{s+t=o} sum_{k in G}
//                sigma_s,ki sigma_t,kj
// keeps the orders o = s + t <= 3 (10 slice pairs); the dropped orders are
// below 2^-31 of sum_G f_Gi f_Gj 1024.  A scale per group rather than per
// column keeps the relative precision of columns whose entries fall off
// with distance (J ~ 1/r): far rows are sliced against their own group's
// largest entry.  Slices are stored k-packed: slice t, rows 8 g .. 8 g + 7
// of column j are the 8 consecutive bf16 at ((t * n_groups + g) * m + j) * 8,
// one 16-byte load per MFMA operand.
//
// int8 slices.  With the same f_Gj, x = X_kj / f_Gj in (-1/2, 1/2) is
// rounded to the integer q = rint(c x) with c = 0xFEFEFEFE (just below
// 2^32), so |q| <= 0x7F7F7F7F, and q is written in balanced base-256
// digits: the bytes b_t of q + 0x80808080 (which lies in [0, 2^32)) give
// d_t = b_t - 128 in [-128, 127], so that
//   X_kj = (2^32 / c) f_Gj sum_t d_t,kj 256^-(t+1)   (error below f_Gj / 2c).
// (c = 2^31 would also do, but would leave the leading digit in
// [-64, 64] and make the dropped orders 4 times larger.)  The digits are
// int8 exactly, products of two have at most 14 bits plus sign, and int32
// MFMA sums are exact (one group's order sum is below 4 x 1024 x 2^14 =
// 2^26).  The normal matrix is the bf16 formula with f_Gi f_Gj multiplied
// by (2^32 / c)^2, a factor applied in FP64 when a group is folded.
// Slice t, rows 16 g .. 16 g + 15 of column j are the 16 consecutive int8
// at ((t * n_units + g) * m + j) * 16 (byte q of the unit is row
// 16 g + q), one 16-byte load per MFMA operand of v_mfma_i32_16x16x64_i8,
// which takes twice the rows of the bf16 MFMA in the same issue time.
constexpr int kSlices = 4;
constexpr int kFoldRows = 1024;  // rows per group (one scale per group and column)
// Slice units (k-groups) per chunk of the gram kernel at most: the slices
// are padded with zero rows to a whole number of them.
constexpr int kSplitChunkUnits = 8;

struct SplitBf16 {
    using unit_t = bf16x8_t;  // one slice of 8 rows of one column
    using acc_t = float4_t;
    static constexpr int kRowsPerUnit = 8;
    static constexpr double kFoldScale = 1.0;  // multiplies f_Gi f_Gj
    __device__ static acc_t mfma(unit_t a, unit_t b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

struct SplitInt8 {
    using unit_t = int32x4_t;  // one slice of 16 rows of one column (16 int8)
    using acc_t = int32x4_t;
    static constexpr int kRowsPerUnit = 16;
    static constexpr double kDigitScale = 4278124286.0;  // c = 0xFEFEFEFE
    static constexpr double kFoldScale = (4294967296.0 / kDigitScale) * (4294967296.0 / kDigitScale);
    __device__ static acc_t mfma(unit_t a, unit_t b, acc_t c) {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
};

// A chunk of the gram kernel is 4 units (32 rows of bf16 slices, 64 of int8);
// a group of kFoldRows rows is a whole, even number of chunks.
template <class Tr>
constexpr int fold_chunks() {
    return kFoldRows / (4 * Tr::kRowsPerUnit);
}
static_assert(fold_chunks<SplitBf16>() % 2 == 0 && fold_chunks<SplitInt8>() % 2 == 0,
              "the gram kernel alternates its LDS buffers in pairs of chunks");

// f[G][j] = 2^(E + 1) with max_{k in group G} sqrt(w_k) J_kj / s_j < 2^E
// (1 if all zero), for the groups G of kFoldRows rows.
__global__ __launch_bounds__(256) void colscale_pow2_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    const double *__restrict__ w, int64_t n_data, const double *__restrict__ pe,
    const double *__restrict__ pn, const double *__restrict__ pu,
    const double *__restrict__ scale, int64_t n_src, double *__restrict__ f) {
    static_assert(kFoldRows % 256 == 0, "a fold group is a whole number of tiles");
    __shared__ double4_t tile[256];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t jj = j < n_src ? j : n_src - 1;
    const double se = pe[jj], sn = pn[jj], su = pu[jj];
    const double inv_s = 1.0 / scale[jj];
    auto put = [&](int64_t G, double mx) {
        if (j >= n_src) return;
        const double cmax = mx * inv_s;
        int ex = 0;
        (void)frexp(cmax, &ex);  // cmax < 2^ex
        f[G * n_src + j] = cmax > 0.0 ? ldexp(1.0, ex + 1) : 1.0;
    };
    double mx = 0.0;
    for (int64_t t0 = 0; t0 < n_data; t0 += 256) {
        const int cnt = (int)(n_data - t0 < 256 ? n_data - t0 : 256);
        __syncthreads();
        if (threadIdx.x < cnt) {
            const int64_t i = t0 + threadIdx.x;
            double4_t v;
            v.x = e[i];
            v.y = n[i];
            v.z = u[i];
            v.w = w ? sqrt(w[i]) : 1.0;
            tile[threadIdx.x] = v;
        }
        __syncthreads();
        for (int q = 0; q < cnt; ++q) {
            const double4_t o = tile[q];
            mx = fmax(mx, o.w * inv_dist(o.x, o.y, o.z, se, sn, su));
        }
        if ((t0 + 256) % kFoldRows == 0 || t0 + 256 >= n_data) {  // end of a group
            put(t0 / kFoldRows, mx);
            mx = 0.0;
        }
    }
}

// The kSlices slices of X / f (f per group of kFoldRows rows) in the
// k-packed layout (rows >= n_data are 0).
__global__ __launch_bounds__(256) void slice_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    const double *__restrict__ w, int64_t n_data, const double *__restrict__ pe,
    const double *__restrict__ pn, const double *__restrict__ pu,
    const double *__restrict__ scale, const double *__restrict__ f, int64_t n_src,
    int64_t n_groups, bf16x8_t *__restrict__ S) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n_src) return;
    const double se = pe[j], sn = pn[j], su = pu[j];
    const double inv_s = 1.0 / scale[j];
    for (int64_t g = blockIdx.y; g < n_groups; g += gridDim.y) {
        const double inv = inv_s / f[(8 * g / kFoldRows) * n_src + j];  // f: a power of 2
        bf16x8_t out[kSlices];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int64_t k = 8 * g + q;
            double x = 0.0;
            if (k < n_data) {
                const double sw = w ? sqrt(w[k]) : 1.0;
                x = sw * inv_dist(e[k], n[k], u[k], se, sn, su) * inv;
            }
#pragma unroll
            for (int t = 0; t < kSlices; ++t) {
                const double y = 256.0 * x;
                const double sig = rint(y);
                x = y - sig;
                out[t][q] = (__bf16)(float)sig;
            }
        }
#pragma unroll
        for (int t = 0; t < kSlices; ++t) S[(t * n_groups + g) * n_src + j] = out[t];
    }
}

// The kSlices int8 slices of X / f (f per group of kFoldRows rows) in the
// k-packed layout (rows >= n_data are 0).
__global__ __launch_bounds__(256) void slice_i8_kernel(
    const double *__restrict__ e, const double *__restrict__ n, const double *__restrict__ u,
    const double *__restrict__ w, int64_t n_data, const double *__restrict__ pe,
    const double *__restrict__ pn, const double *__restrict__ pu,
    const double *__restrict__ scale, const double *__restrict__ f, int64_t n_src,
    int64_t n_units, int32x4_t *__restrict__ S) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n_src) return;
    const double se = pe[j], sn = pn[j], su = pu[j];
    const double inv_s = 1.0 / scale[j];
    for (int64_t g = blockIdx.y; g < n_units; g += gridDim.y) {
        const double inv = inv_s / f[(16 * g / kFoldRows) * n_src + j];  // f: a power of 2
        uint32_t word[kSlices][4] = {};
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t k = 16 * g + q;
            double x = 0.0;
            if (k < n_data) {
                const double sw = w ? sqrt(w[k]) : 1.0;
                x = sw * inv_dist(e[k], n[k], u[k], se, sn, su) * inv;
            }
            const int32_t qi = (int32_t)rint(x * SplitInt8::kDigitScale);  // |qi| <= 0x7F7F7F7F
            // bytes of qi + 0x80808080, each minus 128 (as int8: XOR 0x80)
            const uint32_t d = ((uint32_t)qi + 0x80808080u) ^ 0x80808080u;
#pragma unroll
            for (int t = 0; t < kSlices; ++t)
                word[t][q >> 2] |= ((d >> (8 * (kSlices - 1 - t))) & 0xffu) << (8 * (q & 3));
        }
#pragma unroll
        for (int t = 0; t < kSlices; ++t)
            S[(t * n_units + g) * n_src + j] =
                int32x4_t{(int)word[t][0], (int)word[t][1], (int)word[t][2], (int)word[t][3]};
    }
}

// Orders O0..O1 of the split-precision normal matrix on the upper-triangle
// tiles: C = beta C + sum_G f_Gr f_Gc sum_o 256^-(o+2) sum_{s+t=o} S_s^T S_t
// (+ shift on the diagonal), f the (n_fold, m) group scales, for the
// slice format Tr.  A workgroup of 8 wavefronts owns a kST x kST tile
// (2 x 4 waves of 64 x 32, 4 x 2 MFMA 16x16 tiles each) and one exact
// accumulator set per order (FP32 for bf16 slices, int32 for int8), folded
// into FP64 at the end of every group of kFoldRows rows.  Slices 0..O1 of
// both panels go through two LDS buffers, loaded straight from global
// memory into LDS, with the loads of the next chunk in flight during the
// MFMAs.
template <class Tr, int O0, int O1>
__global__ __launch_bounds__(512) void gram_split_kernel(const typename Tr::unit_t *__restrict__ S,
                                                        int64_t n_groups, int64_t m,
                                                        const double *__restrict__ f,
                                                        double *__restrict__ C, int64_t ldc,
                                                        double beta, double shift,
                                                        int64_t tiles_n) {
    constexpr int kNO = O1 - O0 + 1;  // orders accumulated here
    constexpr int kNS = O1 + 1;       // slices needed on each side
    // k-steps of 4 k-groups (one MFMA deep) per chunk: two when two slices
    // per side leave room in LDS, so that each wait for a chunk's loads is
    // covered by twice the MFMAs.
    constexpr int kSub = kNS <= 2 ? 2 : 1;
    constexpr int kKG = 4 * kSub;  // k-groups per chunk
    static_assert(kSplitChunkUnits % kKG == 0, "slices are padded to whole chunks");
    constexpr int kFoldChunks = fold_chunks<Tr>() / kSub;
    static_assert(kFoldChunks % 2 == 0, "every fold starts in the first buffer");
    using unit_t = typename Tr::unit_t;
    using acc_t = typename Tr::acc_t;
    // Two staging buffers as separate arrays, so that the compiler can tell
    // the buffer being filled from the one being read.
    __shared__ unit_t lds0[2][kNS][kKG][kST];
    __shared__ unit_t lds1[2][kNS][kKG][kST];
    int64_t rb, cb;
    if (!upper_tile_xcd(blockIdx.x, tiles_n, rb, cb)) return;
    const int64_t r0 = rb * kST, c0 = cb * kST;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 2, wc = wave & 3;
    const unit_t zero_unit = {};
    acc_t acc[kNO][4][2];
    double4_t dacc[4][2];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) dacc[ti][tj] = double4_t{0, 0, 0, 0};
    // Staging: per side and slice, 4 k-groups x kST columns, one 16-byte unit
    // per thread, loaded straight into LDS (global_load_lds_dwordx4: lane l
    // of a wave writes 16 bytes at the wave's base + 16 l, so each wave fills
    // one contiguous run of 64 columns).  Columns past m load column m - 1;
    // the products they enter belong to output columns past m, never stored.
    const int kg = threadIdx.x >> 7, col = threadIdx.x & (kST - 1);
    const int col0 = col & ~63;
    // n_groups is a multiple of kSplitChunkUnits, so every chunk is whole.
    auto issue = [&](int ch, unit_t (&L)[2][kNS][kKG][kST]) {
#pragma unroll
        for (int q = 0; q < kSub; ++q) {
            const int64_t g = (int64_t)kKG * ch + 4 * q + kg;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                int64_t gc = (side ? c0 : r0) + col;
                gc = gc < m ? gc : m - 1;
#pragma unroll
                for (int t = 0; t < kNS; ++t)
                    __builtin_amdgcn_global_load_lds(
                        (const void *)(S + ((int64_t)t * n_groups + g) * m + gc),
                        (__attribute__((address_space(3))) void *)&L[side][t][4 * q + kg][col0],
                        16, 0, 0);
            }
        }
    };
    // MFMAs of one chunk from the staged buffer L.
    auto compute = [&](const unit_t (&L)[2][kNS][kKG][kST]) {
#pragma unroll
        for (int q = 0; q < kSub; ++q) {
            const int kq = 4 * q + (lane >> 4);
#pragma unroll
            for (int o = 0; o < kNO; ++o)
#pragma unroll
                for (int sa = 0; sa <= O0 + o; ++sa) {
                    const int sb = O0 + o - sa;
                    unit_t a[4], bb[2];
#pragma unroll
                    for (int t = 0; t < 4; ++t) a[t] = L[0][sa][kq][64 * wr + 16 * t + (lane & 15)];
#pragma unroll
                    for (int t = 0; t < 2; ++t) bb[t] = L[1][sb][kq][32 * wc + 16 * t + (lane & 15)];
#pragma unroll
                    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
                        for (int tj = 0; tj < 2; ++tj)
                            acc[o][ti][tj] = Tr::mfma(a[ti], bb[tj], acc[o][ti][tj]);
                }
        }
    };
    // The chunk's loads complete (this wave's), then every wave's.
    auto arrive = [&]() {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };
    const int chunks = __builtin_amdgcn_readfirstlane((int)(n_groups / kKG));
    if (chunks > 0) issue(0, lds0);
    arrive();
    for (int g0 = 0; g0 < chunks; g0 += kFoldChunks) {
        const int g1 = g0 + kFoldChunks < chunks ? g0 + kFoldChunks : chunks;
        // accumulators born in the matrix unit, so they stay in its registers
#pragma unroll
        for (int o = 0; o < kNO; ++o)
#pragma unroll
            for (int ti = 0; ti < 4; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
                    acc[o][ti][tj] = Tr::mfma(zero_unit, zero_unit, acc_t{0, 0, 0, 0});
        // kFoldChunks is even, so every fold starts in lds0; pairs of
        // chunks alternate lds0 and lds1 with the buffer known statically.
        int ch = g0;
        for (; ch + 1 < g1; ch += 2) {
            issue(ch + 1, lds1);
            compute(lds0);
            arrive();
            if (ch + 2 < chunks) issue(ch + 2, lds0);
            compute(lds1);
            arrive();
        }
        if (ch < g1) {  // odd tail (last fold only): chunk ch in lds0
            if (ch + 1 < chunks) issue(ch + 1, lds1);
            compute(lds0);
            arrive();
        }
        // fold: the group's exact order sums, scaled by f_Gr f_Gc (powers of 2)
        const double *fg = f + (int64_t)(g0 / kFoldChunks) * m;
        double fc[2];
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            const int64_t c = c0 + 32 * wc + 16 * tj + (lane & 15);
            fc[tj] = fg[c < m ? c : m - 1];
        }
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t r = r0 + 64 * wr + 16 * ti + 4 * (lane >> 4) + v;
                const double fr = fg[r < m ? r : m - 1];
#pragma unroll
                for (int tj = 0; tj < 2; ++tj) {
                    double sum = 0.0;
#pragma unroll
                    for (int o = 0; o < kNO; ++o)
                        sum = fma((double)acc[o][ti][tj][v], ldexp(1.0, -8 * (O0 + o + 2)), sum);
                    dacc[ti][tj][v] = fma(sum, fr * fc[tj] * Tr::kFoldScale, dacc[ti][tj][v]);
                }
            }
    }
    double old[4][2][4];
    if (beta != 0.0) {
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    int64_t r = r0 + 64 * wr + 16 * ti + 4 * (lane >> 4) + v;
                    int64_t c = c0 + 32 * wc + 16 * tj + (lane & 15);
                    r = r < m ? r : m - 1;
                    c = c < m ? c : m - 1;
                    old[ti][tj][v] = C[r * ldc + c];
                }
    }
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t r = r0 + 64 * wr + 16 * ti + 4 * (lane >> 4) + v;
                const int64_t c = c0 + 32 * wc + 16 * tj + (lane & 15);
                if (r < m && c < m && r <= c) {
                    double val = dacc[ti][tj][v];
                    if (r == c) val += shift;
                    C[r * ldc + c] = beta == 0.0 ? val : old[ti][tj][v] + val;
                }
            }
}

// ----------------------------------------------------------- host driver
inline unsigned blocks_for(int64_t n, int per = 256) {
    const int64_t b = (n + per - 1) / per;
    return (unsigned)(b < 1 ? 1 : b);
}

// Bytes of the substitutions' error word.
inline size_t solve_err_bytes() { return sizeof(int); }

// Whether a fit takes the SVD route when its Cholesky route cannot match
// the reference's solver (the route is sized for these fits only).
inline bool svd_route_fits(int64_t n_data, int64_t n_src) {
    return n_src <= kSvdMaxSrc && n_data * n_src <= kSvdMaxElems;
}

// Bytes of the scaled-jacobian region of the fit's workspace: the mode's
// jacobian (FP64, FP32, or slices with their group scales), enlarged where
// needed so that, with the copy of the normal matrix in front of it (mixed
// modes), it holds the FP64 column-major jacobian of the SVD route.
inline size_t fit_x_bytes(int64_t n_data, int64_t n_src, int mode) {
    const size_t m = (size_t)n_src, n = (size_t)n_data;
    size_t x_bytes = n * m * sizeof(double);  // X = W^1/2 J S
    if (mode == HMX_FIT_FP32_MFMA) x_bytes = n * m * sizeof(float);
    if (mode == HMX_FIT_BF16_SPLIT || mode == HMX_FIT_INT8_SPLIT) {
        // slices (16-byte units of 8 or 16 rows, padded to whole chunks of
        // kSplitChunkUnits units) and the group scales
        const size_t rows = mode == HMX_FIT_INT8_SPLIT ? 16 : 8;
        const size_t n_units = (n + kSplitChunkUnits * rows - 1) / (kSplitChunkUnits * rows) *
                               kSplitChunkUnits;
        x_bytes = kSlices * n_units * m * 16 +
                  (n_units * rows + kFoldRows - 1) / kFoldRows * m * sizeof(double);
    }
    if (svd_route_fits(n_data, n_src)) {
        const size_t avail = (mode == HMX_FIT_FP64 ? 0 : m * m * sizeof(double)) + x_bytes;
        const size_t g_bytes = n * m * sizeof(double);
        if (avail < g_bytes) x_bytes += g_bytes - avail;
    }
    return x_bytes;
}

// Stream events at the phase boundaries of a fit (destroyed on every
// return path); elapsed times in seconds.
struct PhaseEvents {
    hipEvent_t e[4] = {};
    PhaseEvents() {
        for (auto &x : e)
            if (hipEventCreate(&x) != hipSuccess) x = nullptr;
    }
    ~PhaseEvents() {
        for (auto &x : e)
            if (x) (void)hipEventDestroy(x);
    }
    void record(int k, hipStream_t s) {
        if (e[k]) (void)hipEventRecord(e[k], s);
    }
    double seconds(int a, int b) {
        float ms = 0.0f;
        if (!e[a] || !e[b] || hipEventElapsedTime(&ms, e[a], e[b]) != hipSuccess) return 0.0;
        return 1e-3 * ms;
    }
};

// Pivot ratio below which an unregularised fit (damping=None) is solved by
// the SVD route: min U_ii^2 / max U_ii^2 >= 1 / cond(X)^2, so above it the
// scaled jacobian X has cond(X) < 1e6 and lstsq's cut-off (max(n, m)
// 2^-52 s_max) drops nothing: the least-squares solution is unique and the
// refined Cholesky solution is it.
constexpr double kLstsqPivotRatio = 1e-12;

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
    // the error word of the substitutions' waits (device), and the same
    // pointer once a solve has been set up
    int *err_word = nullptr, *terr = nullptr;
    double *split = nullptr;  // partial sums of the split reductions
    // FP32 blocked copy of the factor's off-diagonal blocks, read by the
    // substitutions when set (use_upper_f32)
    const float *Ub = nullptr;

    // out = (1 / s) J^T W t + alpha p (W = diag(w_), w_ and p may be NULL).
    int jt(const double *w_, const double *t, double alpha_, const double *p, double *out) {
        const int64_t sp = reduce_splits(m, nd), ch = (nd + sp - 1) / sp;
        hipLaunchKernelGGL(jt_apply_kernel, dim3(blocks_for(m), (unsigned)sp), dim3(256), 0, s, e,
                           n, u, w_, nd, pe, pn, pu, t, m, ch, split);
        HMX_LAUNCHED("jt_apply_kernel");
        hipLaunchKernelGGL(jt_apply_finish_kernel, dim3(blocks_for(m)), dim3(256), 0, s, split, sp,
                           scale, m, alpha_, p, out);
        HMX_LAUNCHED("jt_apply_finish_kernel");
        return HMX_OK;
    }

    // scale_j = population std of column j of J (1 where it is 0), two passes.
    int col_stats() {
        const int64_t sp = reduce_splits(m, nd), ch = (nd + sp - 1) / sp;
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL(col_moment_part_kernel, dim3(blocks_for(m), (unsigned)sp),
                               dim3(256), 0, s, pass, e, n, u, nd, pe, pn, pu, m,
                               (const double *)scale, ch, split);
            HMX_LAUNCHED("col_moment_part_kernel");
            hipLaunchKernelGGL(col_moment_finish_kernel, dim3(blocks_for(m)), dim3(256), 0, s,
                               pass, split, sp, m, nd, scale);
            HMX_LAUNCHED("col_moment_finish_kernel");
        }
        return HMX_OK;
    }

    // Two dot products x1.y1 (n1 terms) and x2.y2 (n2 terms) with one trip
    // to the host: 128 partial sums each, added in order.
    int dot2(const double *x1, const double *y1, int64_t n1, const double *x2, const double *y2,
             int64_t n2, double *o1, double *o2) {
        const int nb = 128;
        hipLaunchKernelGGL(reduce_partial_kernel, dim3(nb), dim3(256), 0, s, 0, n1, (int64_t)1, x1,
                           y1, part);
        HMX_LAUNCHED("reduce_partial_kernel");
        hipLaunchKernelGGL(reduce_partial_kernel, dim3(nb), dim3(256), 0, s, 0, n2, (int64_t)1, x2,
                           y2, part + nb);
        HMX_LAUNCHED("reduce_partial_kernel");
        HMX_HIP(hipMemcpyAsync(host_part, part, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, s));
        HMX_HIP(hipStreamSynchronize(s));
        double a1 = 0.0, a2 = 0.0;
        for (int i = 0; i < nb; ++i) {
            a1 += host_part[i];
            a2 += host_part[nb + i];
        }
        *o1 = a1;
        *o2 = a2;
        return HMX_OK;
    }

    int reduce(int mode, const double *x, const double *y, int64_t n, int64_t stride, double *out) {
        const int nb = 256;
        hipLaunchKernelGGL(reduce_partial_kernel, dim3(nb), dim3(256), 0, s, mode, n, stride, x, y, part);
        HMX_LAUNCHED("reduce_partial_kernel");
        HMX_HIP(hipMemcpyAsync(host_part, part, nb * sizeof(double), hipMemcpyDeviceToHost, s));
        HMX_HIP(hipStreamSynchronize(s));
        double acc = mode == 2 ? INFINITY : 0.0;
        for (int i = 0; i < nb; ++i)
            acc = mode == 0   ? acc + host_part[i]
                  : mode == 1 ? std::fmax(acc, host_part[i])
                              : std::fmin(acc, host_part[i]);
        *out = acc;
        return HMX_OK;
    }

    // out = A p (FP64, matrix-free); t receives J S p (data space).
    int apply(const double *p, double *t, double *out) {
        const int64_t sp = reduce_splits(nd, m), ch = (m + sp - 1) / sp;
        hipLaunchKernelGGL(j_apply_kernel, dim3(blocks_for(nd), (unsigned)sp), dim3(256), 0, s, e, n,
                           u, nd, pe, pn, pu, p, scale, m, ch, split);
        HMX_LAUNCHED("j_apply_kernel");
        hipLaunchKernelGGL(j_apply_finish_kernel, dim3(blocks_for(nd)), dim3(256), 0, s, split, sp,
                           nd, t);
        HMX_LAUNCHED("j_apply_finish_kernel");
        return jt(w, t, alpha, p, out);
    }

    // Dimension of the factor (m unless padded; see the batched fits).
    int64_t mf = 0;
    int64_t fm() const { return mf ? mf : m; }

    // The second stream of the factorisation's lookahead and its two
    // events; the stream is drained before it is destroyed.
    struct SideStream {
        hipStream_t s2 = nullptr;
        hipEvent_t steps = nullptr, rest = nullptr;
        ~SideStream() {
            if (s2) {
                (void)hipStreamSynchronize(s2);
                (void)hipStreamDestroy(s2);
            }
            if (steps) (void)hipEventDestroy(steps);
            if (rest) (void)hipEventDestroy(rest);
        }
    };

    // Blocked right-looking factorisation A = U^T U (upper, in place) in two
    // levels: within an outer block of kOuter rows, kNB-row steps factor the
    // diagonal block, solve their row panel up to column m and update only
    // the remaining rows of the outer block; then one rank-kOuter update
    // (K = kOuter, the regime where the MFMA kernel is efficient) brings the
    // trailing matrix up to date.  batch > 1: the same on a strided batch of
    // matrices (element strides bs_a between them and bs_inv between their
    // diagonal-block inverses), every launch covering the whole batch;
    // pivot then receives one status per matrix (0 or the 1-based column of
    // its failed pivot, from the device array infos).
    // Lookahead (one matrix with more than two outer blocks left): the
    // rank-kOuter update is split into the next outer block's rows, updated
    // on the fit's stream, and the rows past them, updated on a second
    // stream while the next outer block's kNB-row steps (one-workgroup
    // diagonal factorisations and narrow panel products, which leave most
    // of the device idle) run on the fit's stream.  Rows [P + kOuter, n)
    // written by the side update of outer block P are next touched by the
    // update of outer block P + kOuter, which waits for it.
    int cholesky(int *pivot, int batch = 1, int64_t bs_a = 0, int64_t bs_inv = 0,
                 int *infos = nullptr) {
        const int64_t n = fm();
        int *dev_info = batch > 1 ? infos : dinfo;
        hipLaunchKernelGGL(fill_int_kernel, dim3(blocks_for(batch)), dim3(256), 0, s, dev_info,
                           (int64_t)batch, 0);
        HMX_LAUNCHED("fill_int_kernel");
        SideStream side;
        bool ahead = batch == 1 && n > 2 * (int64_t)kOuter;
        // The lookahead is an optimisation: without its stream or events
        // the factorisation runs on the fit's stream alone.
        if (ahead && (hipStreamCreateWithFlags(&side.s2, hipStreamNonBlocking) != hipSuccess ||
                      hipEventCreateWithFlags(&side.steps, hipEventDisableTiming) != hipSuccess ||
                      hipEventCreateWithFlags(&side.rest, hipEventDisableTiming) != hipSuccess)) {
            (void)hipGetLastError();  // clear the failed call's error
            ahead = false;
        }
        bool rest_pending = false;
        auto join = [&]() -> int {  // the fit's stream waits for the last side update
            if (!rest_pending) return HMX_OK;
            HMX_HIP(hipStreamWaitEvent(s, side.rest, 0));
            rest_pending = false;
            return HMX_OK;
        };
        for (int64_t P = 0; P < n; P += kOuter) {
            const int64_t Pend = P + kOuter < n ? P + kOuter : n;
            for (int64_t p = P; p < Pend; p += kNB) {
                const int nb = (int)(n - p < kNB ? n - p : kNB);
                double *D = A + p * n + p;
                const int64_t blk = p / kNB;
                hipLaunchKernelGGL(potrf_diag_kernel, dim3((unsigned)batch), dim3(1024), 0, s, D, n,
                                   nb, p, Dinv + blk * kNB * kNB, DinvT + blk * kNB * kNB,
                                   dev_info, bs_a, bs_inv, batch > 1 ? 1 : 0);
                HMX_LAUNCHED("potrf_diag_kernel");
                const int64_t rest = n - p - nb;
                if (rest <= 0) break;
                // U12 = U11^-T A12 = (U11^-1)^T A12, in place, up to column n.
                int st = launch_gemm_tn(Dinv + blk * kNB * kNB, kNB, D + nb, n, nb, nb, rest,
                                        D + nb, n, 0.0, 1.0, s, batch, bs_inv, bs_a, bs_a);
                if (st) return st;
                // Rows [p + nb, Pend) of the outer block, columns [p + nb, n)
                // (the few entries below the diagonal that this rectangle
                // covers are never read).
                const int64_t rows_in = Pend - p - nb;
                if (rows_in > 0) {
                    st = launch_gemm<double>(D + nb, n, D + nb, n, nb, rows_in, rest,
                                             D + nb * n + nb, n, 1.0, -1.0, 0.0, false, s, batch,
                                             bs_a, bs_a, bs_a);
                    if (st) return st;
                }
                if (batch == 1 && blk % 16 == 15) {  // stop early once a pivot has failed
                    HMX_HIP(hipMemcpyAsync(pivot, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
                    HMX_HIP(hipStreamSynchronize(s));
                    if (*pivot) return HMX_OK;
                }
            }
            if (Pend < n) {
                const int64_t K = Pend - P, rest = n - Pend;
                const double *panel = A + P * n + Pend;  // rows [P, Pend), columns [Pend, n)
                double *T = A + Pend * n + Pend;
                int st = HMX_OK;
                if (!ahead || rest <= kOuter) {
                    st = join();
                    if (st) return st;
                    st = launch_syrk<double>(panel, n, K, rest, T, n, 1.0, -1.0, 0.0, s, batch,
                                             bs_a, bs_a);
                    if (st) return st;
                } else {
                    // rows past the next outer block, on the side stream once
                    // this outer block's rows are final
                    HMX_HIP(hipEventRecord(side.steps, s));
                    HMX_HIP(hipStreamWaitEvent(side.s2, side.steps, 0));
                    st = join();  // the previous side update wrote the next outer block's rows
                    if (st) return st;
                    const int64_t top = kOuter, far = rest - kOuter;
                    st = launch_syrk<double>(panel + top, n, K, far, T + top * n + top, n, 1.0,
                                             -1.0, 0.0, side.s2);
                    if (st) return st;
                    HMX_HIP(hipEventRecord(side.rest, side.s2));
                    rest_pending = true;
                    // the next outer block's rows, all columns from Pend (the
                    // entries below the diagonal it covers are never read)
                    st = launch_gemm<double>(panel, n, panel, n, K, top, rest, T, n, 1.0, -1.0,
                                             0.0, false, s);
                    if (st) return st;
                }
            }
        }
        int st = join();
        if (st) return st;
        HMX_HIP(hipMemcpyAsync(pivot, dev_info, (size_t)batch * sizeof(int), hipMemcpyDeviceToHost,
                               s));
        HMX_HIP(hipStreamSynchronize(s));
        return HMX_OK;
    }

    // The error word at zero, before the first solve.
    int init_solves() {
        hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, err_word, (int64_t)1, 0);
        HMX_LAUNCHED("fill_int_kernel");
        terr = err_word;
        return HMX_OK;
    }

    // x = A^-1 r with the factor: U^T y = r (forward), then U x = y
    // (backward), one launch each (y is scratch; r is not modified).
    int solve(double *r, double *y, double *x) {
        const int64_t n = fm();
        const int nb = (int)((n + kNB - 1) / kNB);
        hipLaunchKernelGGL(fill_sentinel_kernel, dim3(blocks_for(n)), dim3(256), 0, s, y, n);
        HMX_LAUNCHED("fill_sentinel_kernel");
        if (Ub)
            hipLaunchKernelGGL(trsv_fwd_sync_kernel<true>, dim3(nb), dim3(1024), 0, s, A, Ub, n, n,
                               Dinv, r, y, terr);
        else
            hipLaunchKernelGGL(trsv_fwd_sync_kernel<false>, dim3(nb), dim3(1024), 0, s, A, Ub, n,
                               n, Dinv, r, y, terr);
        HMX_LAUNCHED("trsv_fwd_sync_kernel");
        hipLaunchKernelGGL(fill_sentinel_kernel, dim3(blocks_for(n)), dim3(256), 0, s, x, n);
        HMX_LAUNCHED("fill_sentinel_kernel");
        if (Ub)
            hipLaunchKernelGGL(trsv_bwd_sync_kernel<true>, dim3(nb), dim3(1024), 0, s, A, Ub, n, n,
                               DinvT, y, x, terr);
        else
            hipLaunchKernelGGL(trsv_bwd_sync_kernel<false>, dim3(nb), dim3(1024), 0, s, A, Ub, n,
                               n, DinvT, y, x, terr);
        HMX_LAUNCHED("trsv_bwd_sync_kernel");
        return HMX_OK;
    }

    // Bytes of the FP32 copy of the factor's blocks above the diagonal.
    int64_t upper_f32_bytes() const {
        const int64_t nb = (fm() + kNB - 1) / kNB;
        return nb * (nb - 1) / 2 * kNB * kNB * (int64_t)sizeof(float);
    }

    // Packs that copy into `dst` (upper_f32_bytes() bytes) and has the
    // substitutions read it from then on (the factor is then only a
    // preconditioner: the FP32 rounding of U moves the preconditioned
    // operator, not the solution CG converges to).
    int use_upper_f32(float *dst) {
        const int64_t n = fm();
        const int nb = (int)((n + kNB - 1) / kNB);
        if (nb > 1) {
            hipLaunchKernelGGL(pack_upper_f32_kernel, dim3(nb, nb), dim3(256), 0, s, A, n, n, nb,
                               dst);
            HMX_LAUNCHED("pack_upper_f32_kernel");
        }
        Ub = dst;
        return HMX_OK;
    }

    // x = (factor)^-1 b, then preconditioned CG on the exact FP64 operator
    // A = S J^T W J S + alpha I (starting from x), stopping when the
    // data-space update |a| |J S p| falls below rtol |J S x| or after
    // `allowed` iterations.  The vectors hold fm() doubles (the factor's
    // dimension); entries past m stay zero.
    int solve_refine(const double *b, double *x, double *r, double *z, double *pd, double *q,
                     double *wk, double *wk2, int allowed, double rtol, int &iters,
                     double &last_rel) {
        const unsigned gm = blocks_for(m);
        iters = 0;
        last_rel = 0.0;
        int st = init_solves();
        if (st) return st;
        if (fm() > m) {  // the factor's padding: zero right-hand sides
            hipLaunchKernelGGL(fill_zero_f64_kernel, dim3(blocks_for(fm())), dim3(256), 0, s, wk,
                               fm());
            HMX_LAUNCHED("fill_zero_f64_kernel");
        }
        hipLaunchKernelGGL(copy_kernel, dim3(gm), dim3(256), 0, s, m, b, wk);
        HMX_LAUNCHED("copy_kernel");
        st = solve(wk, wk2, x);
        if (st) return st;
        if (allowed <= 0) return HMX_OK;
        double *t_dat = tmp;
        st = apply(x, t_dat, q);  // q = A x
        if (st) return st;
        hipLaunchKernelGGL(sub_kernel, dim3(gm), dim3(256), 0, s, m, b, q, r);  // r = b - A x
        HMX_LAUNCHED("sub_kernel");
        hipLaunchKernelGGL(copy_kernel, dim3(gm), dim3(256), 0, s, m, r, wk);
        HMX_LAUNCHED("copy_kernel");
        st = solve(wk, wk2, z);
        if (st) return st;
        hipLaunchKernelGGL(copy_kernel, dim3(gm), dim3(256), 0, s, m, z, pd);
        HMX_LAUNCHED("copy_kernel");
        double ynorm = 0.0, rz = 0.0;  // |J S x|^2 (t_dat is still J S x) and r.z
        st = dot2(t_dat, t_dat, nd, r, z, m, &ynorm, &rz);
        if (st) return st;
        for (iters = 1; iters <= allowed; ++iters) {
            st = apply(pd, t_dat, q);  // q = A p, t = J S p
            if (st) return st;
            double pq = 0.0, tt = 0.0;
            st = dot2(pd, q, m, t_dat, t_dat, nd, &pq, &tt);
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
            st = solve(wk, wk2, z);
            if (st) return st;
            double rz_new = 0.0;
            st = reduce(0, r, z, m, 1, &rz_new);
            if (st) return st;
            hipLaunchKernelGGL(xpby_kernel, dim3(gm), dim3(256), 0, s, m, z, rz_new / rz, pd);
            HMX_LAUNCHED("xpby_kernel");
            rz = rz_new;
        }
        if (iters > allowed) iters = allowed;
        return HMX_OK;
    }

    // Fails if a wait of the substitutions timed out (synchronises).
    int check_solves() {
        if (!terr) return HMX_OK;  // no solve ran
        int e = 0;
        HMX_HIP(hipMemcpyAsync(&e, terr, sizeof(int), hipMemcpyDeviceToHost, s));
        HMX_HIP(hipStreamSynchronize(s));
        if (e) return fail(HMX_ERR_NUMERIC, "triangular solve: a dependency wait timed out");
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
    } else if (dtype == HMX_F32) {
        float4_t c = {0, 0, 0, 0};
        c = Mfma<float>::step((float)a[l], (float)b[l], c);
        for (int v = 0; v < 4; ++v) out[4 * l + v] = (double)c[v];
    } else {
        bf16x8_t va, vb;
        for (int e = 0; e < 8; ++e) {
            va[e] = (__bf16)(float)a[8 * l + e];
            vb[e] = (__bf16)(float)b[8 * l + e];
        }
        float4_t c = {0, 0, 0, 0};
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, c, 0, 0, 0);
        for (int v = 0; v < 4; ++v) out[4 * l + v] = (double)c[v];
    }
}

}  // namespace

// ------------------------------------------- batched window factorisation
// The gradient-boosted fit's windows in batches: the normal matrices of a
// batch of windows are formed and factorised together (every launch covers
// the batch), and each window is then solved from its stored factor in the
// sequential window loop.  The windows' matrices are padded to a common
// size: m_pad = the largest window's sources rounded up to 8, the scaled
// jacobians to n_max rows (zero rows and columns), and the shift added to
// the whole diagonal (damping > 0), so a padding column's factor row is
// sqrt(damping) on the diagonal and zero elsewhere and solves give zeros
// there.
struct BatchLayout {
    int64_t m_pad, n_max, nblk;
    size_t x_off, a_off, dinv_off, scale_off, vec_off, tmp_off, part_off, split_off, dinfo_off,
        err_off, infos_off, dims_off, total;
};

inline size_t al256(size_t b) { return (b + 255) / 256 * 256; }

BatchLayout batch_layout(int64_t n_max, int64_t m_max, int64_t batch) {
    BatchLayout L;
    L.m_pad = (m_max + 7) / 8 * 8;
    L.n_max = n_max;
    L.nblk = (L.m_pad + kNB - 1) / kNB;
    const size_t B = (size_t)batch, mp = (size_t)L.m_pad;
    size_t off = 0;
    L.x_off = off;
    off += al256(B * (size_t)n_max * mp * sizeof(double));
    L.a_off = off;
    off += al256(B * mp * mp * sizeof(double));
    L.dinv_off = off;
    off += al256(2 * B * (size_t)L.nblk * kNB * kNB * sizeof(double));
    L.scale_off = off;
    off += al256(B * mp * sizeof(double));
    L.vec_off = off;  // 9 vectors of m_pad (b, x, r, z, p, q, wk, wk2, spare)
    off += al256(9 * mp * sizeof(double));
    L.tmp_off = off;
    off += al256((size_t)n_max * sizeof(double));
    L.part_off = off;
    off += al256(256 * sizeof(double));
    L.split_off = off;
    off += al256(split_doubles(n_max > L.m_pad ? n_max : L.m_pad) * sizeof(double));
    L.dinfo_off = off;
    off += 256;
    L.err_off = off;
    off += al256(solve_err_bytes());
    L.infos_off = off;
    off += al256(B * sizeof(int));
    L.dims_off = off;  // per window: its data rows and its sources (int64)
    off += al256(2 * B * sizeof(int64_t));
    L.total = off;
    return L;
}

size_t eqs_batch_bytes(int64_t n_max, int64_t m_max, int64_t batch) {
    return batch_layout(n_max, m_max, batch).total;
}

namespace {

// Fit of window w with the batch buffers (vectors and factor of slot b).
Fit batch_fit(const EqsWindow &w, double damping, const BatchLayout &L, char *buf, int64_t b,
              hipStream_t s) {
    Fit f;
    f.e = w.e;
    f.n = w.n;
    f.u = w.u;
    f.w = w.w;
    f.nd = w.nd;
    f.pe = w.pe;
    f.pn = w.pn;
    f.pu = w.pu;
    f.m = w.m;
    f.mf = L.m_pad;
    f.alpha = damping;
    f.s = s;
    f.scale = reinterpret_cast<double *>(buf + L.scale_off) + b * L.m_pad;
    f.A = reinterpret_cast<double *>(buf + L.a_off) + b * L.m_pad * L.m_pad;
    f.Dinv = reinterpret_cast<double *>(buf + L.dinv_off) + b * L.nblk * kNB * kNB;
    f.DinvT = f.Dinv + (int64_t)0;  // set by the caller (batch stride)
    f.tmp = reinterpret_cast<double *>(buf + L.tmp_off);
    f.part = reinterpret_cast<double *>(buf + L.part_off);
    f.split = reinterpret_cast<double *>(buf + L.split_off);
    f.dinfo = reinterpret_cast<int *>(buf + L.dinfo_off);
    f.err_word = reinterpret_cast<int *>(buf + L.err_off);
    return f;
}

}  // namespace

int eqs_factor_batch(const EqsWindow *wins, int64_t batch, int64_t n_max, int64_t m_max,
                     double damping, void *buffer, size_t buffer_bytes, int *status,
                     double *seconds, hipStream_t s) {
    if (batch <= 0) return HMX_OK;
    if (!(damping > 0.0)) return fail(HMX_ERR_ARG, "batched factorisation needs damping > 0");
    const BatchLayout L = batch_layout(n_max, m_max, batch);
    if (!buffer || buffer_bytes < L.total)
        return fail(HMX_ERR_ARG, "batch buffer too small: %zu < %zu bytes", buffer_bytes, L.total);
    char *buf = static_cast<char *>(buffer);
    PhaseEvents ev;
    ev.record(0, s);
    double *X = reinterpret_cast<double *>(buf + L.x_off);
    const int64_t bs_x = L.n_max * L.m_pad, bs_a = L.m_pad * L.m_pad;
    const int64_t bs_inv = L.nblk * kNB * kNB;
    hipLaunchKernelGGL(fill_zero_f64_kernel, dim3(blocks_for(batch * bs_x)), dim3(256), 0, s, X,
                       batch * bs_x);
    HMX_LAUNCHED("fill_zero_f64_kernel");
    // column scales and the scaled jacobian of each window (zero padding)
    std::vector<int64_t> dims(2 * (size_t)batch);
    for (int64_t b = 0; b < batch; ++b) {
        const EqsWindow &w = wins[b];
        if (w.nd <= 0 || w.m <= 0 || w.nd > n_max || w.m > m_max)
            return fail(HMX_ERR_ARG, "window %lld of the batch has a bad size", (long long)b);
        dims[2 * b] = w.nd;
        dims[2 * b + 1] = w.m;
        Fit f = batch_fit(w, damping, L, buf, b, s);
        int st = f.col_stats();
        if (st) return st;
        const unsigned gm = blocks_for(w.m), gy = (unsigned)(w.nd < 8192 ? w.nd : 8192);
        hipLaunchKernelGGL(scaled_jacobian_kernel<double>, dim3(gm, gy), dim3(256), 0, s, w.e,
                           w.n, w.u, w.w, w.nd, w.pe, w.pn, w.pu, f.scale, w.m, X + b * bs_x,
                           L.m_pad);
        HMX_LAUNCHED("scaled_jacobian_kernel");
    }
    // A_b = X_b^T X_b + damping I for the whole batch, then the factors;
    // each window's product over its own rows and sources only (the padding
    // is zero), the tiles past its sources left at zero plus the shift.
    auto *d_dims = reinterpret_cast<int64_t *>(buf + L.dims_off);
    HMX_HIP(hipMemcpyAsync(d_dims, dims.data(), dims.size() * sizeof(int64_t),
                           hipMemcpyHostToDevice, s));
    HMX_HIP(hipStreamSynchronize(s));  // dims is host memory of this frame
    double *A = reinterpret_cast<double *>(buf + L.a_off);
    int st = launch_syrk<double>(X, L.m_pad, L.n_max, L.m_pad, A, L.m_pad, 0.0, 1.0, damping, s,
                                 (int)batch, bs_x, bs_a, d_dims);
    if (st) return st;
    ev.record(1, s);
    Fit f = batch_fit(wins[0], damping, L, buf, 0, s);
    f.m = L.m_pad;
    f.DinvT = f.Dinv + batch * bs_inv;
    st = f.cholesky(status, (int)batch, bs_a, bs_inv, reinterpret_cast<int *>(buf + L.infos_off));
    if (st) return st;
    ev.record(2, s);
    HMX_HIP(hipStreamSynchronize(s));
    seconds[0] = ev.seconds(0, 1);
    seconds[1] = ev.seconds(1, 2);
    return HMX_OK;
}

int eqs_solve_factored(const EqsWindow &w, const double *data, double damping, void *buffer,
                       int64_t b, int64_t batch, int64_t n_max, int64_t m_max, int max_iter,
                       double rtol, double *coefs, double *info, hipStream_t s) {
    const BatchLayout L = batch_layout(n_max, m_max, batch);
    char *buf = static_cast<char *>(buffer);
    Fit f = batch_fit(w, damping, L, buf, b, s);
    f.DinvT = f.Dinv + batch * L.nblk * kNB * kNB;
    double *vec = reinterpret_cast<double *>(buf + L.vec_off);
    const int64_t mp = L.m_pad;
    double *bv = vec, *x = vec + mp, *r = x + mp, *z = r + mp, *pd = z + mp, *q = pd + mp,
           *wk = q + mp, *wk2 = wk + mp;
    PhaseEvents ev;
    ev.record(0, s);
    int st = f.jt(w.w, data, 0.0, nullptr, bv);  // b = S J^T W d
    if (st) return st;
    int iters = 0;
    double last_rel = 0.0;
    st = f.solve_refine(bv, x, r, z, pd, q, wk, wk2, max_iter, rtol, iters, last_rel);
    if (st) return st;
    hipLaunchKernelGGL(div_kernel, dim3(blocks_for(w.m)), dim3(256), 0, s, w.m, x, f.scale, coefs);
    HMX_LAUNCHED("div_kernel");
    ev.record(1, s);
    st = f.check_solves();
    if (st) return st;
    for (int k = 0; k < 11; ++k) info[k] = 0.0;
    info[0] = (double)iters;
    info[1] = last_rel;
    info[10] = ev.seconds(0, 1);
    return HMX_OK;
}

}  // namespace hmx

extern "C" {

int hmx_mfma_probe(int dtype, const double *a, const double *b, double *out, void *stream) {
    using namespace hmx;
    clear_error();
    if (dtype != HMX_F64 && dtype != HMX_F32 && dtype != HMX_BF16)
        return fail(HMX_ERR_ARG, "invalid dtype %d", dtype);
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
    if (dtype == HMX_F64)
        return launch_gemm<double>((const double *)A, lda, (const double *)B, ldb, K, M, N, C, ldc,
                                   beta, sign, shift, upper != 0, s);
    if (dtype == HMX_F32)
        return launch_gemm<float>((const float *)A, lda, (const float *)B, ldb, K, M, N, C, ldc,
                                  beta, sign, shift, upper != 0, s);
    return fail(HMX_ERR_ARG, "invalid dtype %d", dtype);
}

int hmx_cholesky_workspace_bytes(int64_t m, size_t *bytes) {
    using namespace hmx;
    if (!bytes) return fail(HMX_ERR_ARG, "bytes is NULL");
    if (m <= 0) return fail(HMX_ERR_ARG, "empty matrix");
    const size_t nblk = (size_t)((m + kNB - 1) / kNB);
    *bytes = 2 * nblk * kNB * kNB * sizeof(double) + (2 * (size_t)m + 256) * sizeof(double) + 256 +
             solve_err_bytes();
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
    f.err_word = f.dinfo + 64;
    st = f.cholesky(info);
    if (st) return st;
    if (*info) return HMX_OK;
    st = f.init_solves();
    if (st) return st;
    hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(m)), dim3(256), 0, f.s, m, x, wk);
    HMX_LAUNCHED("copy_kernel");
    st = f.solve(wk, wk2, x);
    if (st) return st;
    return f.check_solves();
}

int hmx_eqs_fit_workspace_bytes(int64_t n_data, int64_t n_src, int precision_mode,
                                size_t *bytes) {
    using namespace hmx;
    if (!bytes) return fail(HMX_ERR_ARG, "bytes is NULL");
    if (n_data <= 0 || n_src <= 0) return fail(HMX_ERR_ARG, "empty problem");
    const int mode = precision_mode & ~HMX_FIT_NO_DAMPING;
    if (mode != HMX_FIT_FP64 && mode != HMX_FIT_FP32_MFMA && mode != HMX_FIT_BF16_SPLIT &&
        mode != HMX_FIT_INT8_SPLIT)
        return fail(HMX_ERR_ARG, "invalid precision mode %d", precision_mode);
    const size_t m = (size_t)n_src, n = (size_t)n_data;
    const size_t nblk = (m + kNB - 1) / kNB;
    const size_t copies = mode == HMX_FIT_FP64 ? 1 : 2;
    *bytes = copies * m * m * sizeof(double)             // A, then its factor U (+ a copy of A)
             + fit_x_bytes(n_data, n_src, mode)        // the scaled jacobian (or its slices)
             + 2 * nblk * kNB * kNB * sizeof(double)  // inverses of diagonal blocks
             + (10 * m + n + 256) * sizeof(double)  // 10 m-vectors, n-vector, partials
             + split_doubles(m > n ? m : n) * sizeof(double) + 256  // split sums, info
             + solve_err_bytes();                                   // substitutions' error word
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
    // damping=None: LinearRegression's minimum-norm least squares
    const bool lstsq = (precision_mode & HMX_FIT_NO_DAMPING) != 0;
    precision_mode &= ~HMX_FIT_NO_DAMPING;
    if (lstsq && damping != 0.0)
        return fail(HMX_ERR_ARG, "damping must be 0 with HMX_FIT_NO_DAMPING");
    hipStream_t s = (hipStream_t)stream;
    const int64_t m = n_src, nd = n_data;
    const bool svd_ok = svd_route_fits(nd, m);
    // mixed: a low-precision normal matrix preconditions CG on the FP64 one
    const bool mixed = precision_mode != HMX_FIT_FP64;
    const bool int8 = precision_mode == HMX_FIT_INT8_SPLIT;
    const bool split = precision_mode == HMX_FIT_BF16_SPLIT || int8;
    const size_t nblk = (size_t)((m + kNB - 1) / kNB);
    // split modes: 16-byte slice units of 8 (bf16) or 16 (int8) rows, padded
    // to whole chunks of kSplitChunkUnits units
    const int64_t unit_rows = int8 ? 16 : 8;
    const int64_t n_groups = (nd + kSplitChunkUnits * unit_rows - 1) /
                             (kSplitChunkUnits * unit_rows) * kSplitChunkUnits;

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
    double *A_copy = nullptr;  // mixed mode: the unshifted normal matrix, for retries
    if (mixed) {
        A_copy = reinterpret_cast<double *>(ws);
        ws += (size_t)m * m * sizeof(double);
    }
    void *X = ws;
    double *fgroup = nullptr;  // split: (n_fold, m) power-of-2 column scales per row group
    if (split) fgroup = reinterpret_cast<double *>(ws + (size_t)kSlices * n_groups * m * 16);
    ws += fit_x_bytes(nd, m, precision_mode);
    f.Dinv = reinterpret_cast<double *>(ws);
    f.DinvT = f.Dinv + nblk * kNB * kNB;
    double *vec = f.DinvT + nblk * kNB * kNB;
    f.scale = vec;
    double *b = vec + m, *x = b + m, *r = x + m, *z = r + m, *pd = z + m, *q = pd + m, *wk = q + m,
           *wk2 = wk + m, *pow2 = wk2 + m;
    f.tmp = pow2 + m;  // n_data
    f.part = f.tmp + nd;
    f.split = f.part + 256;
    f.dinfo = reinterpret_cast<int *>(f.split + split_doubles(m > nd ? m : nd));
    f.err_word = f.dinfo + 64;  // after the 256 bytes of info
    const unsigned gm = blocks_for(m), gy = (unsigned)(nd < 8192 ? nd : 8192);

    // phases: 0 start, 1 factorisation start, 2 factorisation end, 3 end
    PhaseEvents ev;
    ev.record(0, s);

    // 1. column scales, right-hand side b = S J^T W d (FP64).
    st = f.col_stats();
    if (st) return st;
    st = f.jt(weights, data, 0.0, nullptr, b);
    if (st) return st;

    // 2. X = W^1/2 J S (or its slices) and the normal matrix in the mode's
    // precision.
    if (split) {
        hipLaunchKernelGGL(colscale_pow2_kernel, dim3(gm), dim3(256), 0, s, easting, northing,
                           upward, weights, nd, p_easting, p_northing, p_upward, f.scale, m,
                           fgroup);
        HMX_LAUNCHED("colscale_pow2_kernel");
        const unsigned gg = (unsigned)(n_groups < 4096 ? n_groups : 4096);
        if (int8) {
            hipLaunchKernelGGL(slice_i8_kernel, dim3(gm, gg), dim3(256), 0, s, easting, northing,
                               upward, weights, nd, p_easting, p_northing, p_upward, f.scale,
                               fgroup, m, n_groups, static_cast<int32x4_t *>(X));
            HMX_LAUNCHED("slice_i8_kernel");
        } else {
            hipLaunchKernelGGL(slice_kernel, dim3(gm, gg), dim3(256), 0, s, easting, northing,
                               upward, weights, nd, p_easting, p_northing, p_upward, f.scale,
                               fgroup, m, n_groups, static_cast<bf16x8_t *>(X));
            HMX_LAUNCHED("slice_kernel");
        }
    } else if (mixed) {
        hipLaunchKernelGGL(scaled_jacobian_kernel<float>, dim3(gm, gy), dim3(256), 0, s, easting,
                           northing, upward, weights, nd, p_easting, p_northing, p_upward, f.scale,
                           m, static_cast<float *>(X), m);
    } else {
        hipLaunchKernelGGL(scaled_jacobian_kernel<double>, dim3(gm, gy), dim3(256), 0, s, easting,
                           northing, upward, weights, nd, p_easting, p_northing, p_upward, f.scale,
                           m, static_cast<double *>(X), m);
    }
    if (!split) HMX_LAUNCHED("scaled_jacobian_kernel");

    // The FP32-MFMA normal matrix differs from the FP64 one by about 2^-24
    // of its entries (input rounding and FP32 sums within each MFMA), the
    // split-bf16 one by about 2^-31; with damping small against the largest
    // eigenvalue that can make it indefinite.  A failed factorisation is
    // retried with a shift tau (then tenfold larger), and the factor of the
    // (shifted) low-precision matrix serves as the preconditioner of CG on
    // the exact FP64 operator.
    double tau = 0.0, dmax = 0.0;
    auto diag_max = [&](double *out) -> int {
        if (split)  // the diagonal of the saved normal matrix
            return f.reduce(1, A_copy, nullptr, m, m + 1, out);
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
    // The SVD route: the reference's Ridge falls back to an SVD when its
    // Cholesky solve fails, and its damping=None fit is an SVD solve.
    bool use_svd = false;
    for (;;) {
        if (mixed && attempts > 0) {  // restore the normal matrix, add the shift
            HMX_HIP(hipMemcpyAsync(f.A, A_copy, (size_t)m * m * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(add_diag_kernel, dim3(gm), dim3(256), 0, s, f.A, m, tau);
            HMX_LAUNCHED("add_diag_kernel");
        } else if (mixed) {
            if (split) {
                const int64_t tiles = (m + kST - 1) / kST;
                const unsigned blocks = upper_tile_xcd_blocks(tiles);
                if (int8) {
                    const int32x4_t *Sl = static_cast<const int32x4_t *>(X);
                    hipLaunchKernelGGL((gram_split_kernel<SplitInt8, 0, 1>), dim3(blocks),
                                       dim3(512), 0, s, Sl, n_groups, m, fgroup, f.A, m, 0.0,
                                       damping, tiles);
                    HMX_LAUNCHED("gram_split_kernel<SplitInt8,0,1>");
                    hipLaunchKernelGGL((gram_split_kernel<SplitInt8, 2, 3>), dim3(blocks),
                                       dim3(512), 0, s, Sl, n_groups, m, fgroup, f.A, m, 1.0, 0.0,
                                       tiles);
                    HMX_LAUNCHED("gram_split_kernel<SplitInt8,2,3>");
                } else {
                    const bf16x8_t *Sl = static_cast<const bf16x8_t *>(X);
                    hipLaunchKernelGGL((gram_split_kernel<SplitBf16, 0, 1>), dim3(blocks),
                                       dim3(512), 0, s, Sl, n_groups, m, fgroup, f.A, m, 0.0,
                                       damping, tiles);
                    HMX_LAUNCHED("gram_split_kernel<SplitBf16,0,1>");
                    hipLaunchKernelGGL((gram_split_kernel<SplitBf16, 2, 3>), dim3(blocks),
                                       dim3(512), 0, s, Sl, n_groups, m, fgroup, f.A, m, 1.0, 0.0,
                                       tiles);
                    HMX_LAUNCHED("gram_split_kernel<SplitBf16,2,3>");
                }
            } else {
                st = launch_syrk<float>(static_cast<const float *>(X), m, nd, m, f.A, m, 0.0, 1.0,
                                        damping, s);
                if (st) return st;
            }
            HMX_HIP(hipMemcpyAsync(A_copy, f.A, (size_t)m * m * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
        } else {
            st = launch_syrk<double>(static_cast<const double *>(X), m, nd, m, f.A, m, 0.0, 1.0,
                                     damping + tau, s);
            if (st) return st;
        }
        if (attempts == 0) ev.record(1, s);
        st = f.cholesky(&pivot);
        if (st) return st;
        if (pivot == 0) break;
        if (svd_ok) {
            use_svd = true;
            break;
        }
        if (++attempts > 12)
            return fail(HMX_ERR_NUMERIC, "normal matrix not positive definite (column %d)", pivot);
        if (dmax == 0.0) {
            st = diag_max(&dmax);
            if (st) return st;
        }
        // First shift, as a fraction of the largest diagonal entry: 2^-26
        // for FP32 MFMA sums, 2^-30 for the bf16 slices (fixed-point
        // columns carry 2^-32 of their largest entry), 1e-14 for FP64; then
        // tenfold.  A shift just above the error of the normal matrix keeps
        // the conjugate-gradient refinement short.
        const double tau0 = split ? std::ldexp(dmax, -30) : mixed ? std::ldexp(dmax, -26)
                                                                  : 1e-14 * dmax;
        tau = tau > 0.0 ? 10.0 * tau : tau0;
    }

    ev.record(2, s);

    // An unregularised fit whose factor has a small pivot ratio may have
    // singular values below lstsq's cut-off: SVD route.
    if (lstsq && svd_ok && !use_svd) {
        double umin = 0.0, umax = 0.0;
        st = f.reduce(2, f.A, nullptr, m, m + 1, &umin);
        if (st) return st;
        st = f.reduce(1, f.A, nullptr, m, m + 1, &umax);
        if (st) return st;
        if (!(umin * umin >= kLstsqPivotRatio * umax * umax)) use_svd = true;
    }

    // 3-4. x = (factor)^-1 b, refined by preconditioned CG.
    int iters = 0;
    double last_rel = 0.0;
    // A shifted factorisation is only a preconditioner: refine even when
    // the caller asked for a direct solve.
    const int allowed = (tau > 0.0 && max_iter == 0) ? 200 : max_iter;
    if (!use_svd) {
        // Mixed modes: the factor (of the low-precision normal matrix) is a
        // preconditioner only, and the substitutions read an FP32 blocked
        // copy of it, held where the saved normal matrix was (2 m^2 bytes
        // at most, against its 8 m^2).
        if (mixed && allowed > 0 && f.upper_f32_bytes() <= (int64_t)(m * m * sizeof(double))) {
            st = f.use_upper_f32(reinterpret_cast<float *>(A_copy));
            if (st) return st;
        }
        st = f.solve_refine(b, x, r, z, pd, q, wk, wk2, allowed, rtol, iters, last_rel);
        if (st) return st;
        // an unregularised fit whose refinement did not converge
        if (allowed > 0 && lstsq && svd_ok && last_rel > rtol) use_svd = true;
    }

    // 4b. The SVD route on the scaled, weighted jacobian X = W^1/2 J S:
    // with n_data >= m its column-major copy (over the normal matrix's copy
    // and the jacobian region), else its row-major copy (X^T column-major,
    // the FP64 mode's own X); the rotations over the normal matrix;
    // x = V diag(f(s)) U^T W^1/2 d with the filter of the reference's
    // solver for this problem.
    int sweeps = 0, rank = 0;
    if (use_svd) {
        double *G = mixed ? A_copy : static_cast<double *>(X);
        const bool transposed = m > nd;
        const unsigned gyc = (unsigned)(m < 512 ? m : 512);
        if (!transposed) {
            hipLaunchKernelGGL(scaled_jacobian_cm_kernel, dim3(blocks_for(nd), gyc), dim3(256), 0,
                               s, easting, northing, upward, weights, nd, p_easting, p_northing,
                               p_upward, f.scale, m, G);
            HMX_LAUNCHED("scaled_jacobian_cm_kernel");
        } else if (mixed) {
            hipLaunchKernelGGL(scaled_jacobian_kernel<double>, dim3(gm, gy), dim3(256), 0, s,
                               easting, northing, upward, weights, nd, p_easting, p_northing,
                               p_upward, f.scale, m, G, m);
            HMX_LAUNCHED("scaled_jacobian_kernel");
        }
        hipLaunchKernelGGL(sqrtw_times_kernel, dim3(blocks_for(nd)), dim3(256), 0, s, nd, weights,
                           data, f.tmp);
        HMX_LAUNCHED("sqrtw_times_kernel");
        const double eps = std::ldexp(1.0, -52);
        int filter = HMX_SVD_RIDGE;
        double cutoff = 1e-15;  // sklearn's _solve_svd
        if (lstsq) {
            filter = HMX_SVD_LSTSQ;  // LinearRegression: lstsq(cond = max(n, m) eps)
            cutoff = (double)(nd > m ? nd : m) * eps;
        } else if (transposed) {
            filter = HMX_SVD_DUAL_LSTSQ;  // Ridge's dual form: lstsq(X X^T + alpha I)
            cutoff = eps;
        }
        st = svd_lsq(G, transposed ? m : nd, transposed ? nd : m, transposed, f.tmp,
                     lstsq ? 0.0 : damping, filter, cutoff, f.A, r, f.dinfo + 1, x,
                     kSvdMaxSweeps, &sweeps, &rank, s);
        if (st) return st;
        iters = 0;
        last_rel = 0.0;
    }

    // 5. c = S c'.
    hipLaunchKernelGGL(div_kernel, dim3(gm), dim3(256), 0, s, m, x, f.scale, coefs);
    HMX_LAUNCHED("div_kernel");
    ev.record(3, s);
    HMX_HIP(hipStreamSynchronize(s));
    st = f.check_solves();
    if (st) return st;
    info[0] = (double)iters;
    info[1] = last_rel;
    info[2] = tau;
    info[3] = (double)attempts;
    info[4] = dmax;
    info[5] = use_svd ? 1.0 : 0.0;
    info[6] = (double)sweeps;
    info[7] = (double)rank;
    info[8] = ev.seconds(0, 1);  // column statistics, jacobian and normal matrix
    info[9] = ev.seconds(1, 2);  // factorisation (with any retries)
    info[10] = ev.seconds(2, 3);  // solves, refinement or the SVD route
    return HMX_OK;
}

}  // extern "C"
