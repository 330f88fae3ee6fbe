// Gradient-boosted equivalent-source fit: the window loop of
// EquivalentSourcesGB._gradient_boosting
// (src/harmonica/_equivalent_sources/gradient_boosted.py:231-283) on the
// device.  The windows are sequential (each fits the residue that the
// previous ones left), so every step stays on the GPU and the host only
// walks the window list:
//   1. gather the window's data (coordinates, current residue, weights) and
//      sources into contiguous buffers (one kernel);
//   2. fit the window's coefficients with hmx_eqs_fit (the jacobian of the
//      window and verde's scaled ridge least squares);
//   3. evaluate the window's sources on every datum with hmx_eqs_predict and
//      subtract from the residue, with the residue's sum of squares reduced
//      in a fixed order (two kernels, deterministic);
//   4. add the window's coefficients to its sources' coefficients (the
//      sources of one window are distinct, so no atomics).
// In the FP64 mode with damping > 0 the windows' normal matrices, which do
// not depend on the residue, are formed and factorised `batch` windows at a
// time (eqs_factor_batch: every launch covers the batch); step 2 then
// solves the window from its stored factor (eqs_solve_factored: right-hand
// side, triangular solves and the same CG refinement).  A window whose
// factorisation fails takes hmx_eqs_fit's own route.
#include <vector>

#include "hmx_internal.h"

namespace hmx {
namespace {

constexpr int kGbThreads = 256;

__global__ void gb_gather_kernel(const double *__restrict__ e, const double *__restrict__ n,
                                 const double *__restrict__ u, const double *__restrict__ residue,
                                 const double *__restrict__ weights,
                                 int64_t n_data, const int64_t *__restrict__ didx, int64_t nd,
                                 const double *__restrict__ pe, const double *__restrict__ pn,
                                 const double *__restrict__ pu, int64_t n_src,
                                 const int64_t *__restrict__ sidx, int64_t ns,
                                 double *__restrict__ we, double *__restrict__ wn,
                                 double *__restrict__ wu, double *__restrict__ wr,
                                 double *__restrict__ ww, double *__restrict__ wpe,
                                 double *__restrict__ wpn, double *__restrict__ wpu) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // An index outside the arrays gives NaN instead of a read out of bounds.
    const double qnan = __builtin_nan("");
    if (k < nd) {
        const int64_t i = didx[k];
        const bool ok = i >= 0 && i < n_data;
        we[k] = ok ? e[i] : qnan;
        wn[k] = ok ? n[i] : qnan;
        wu[k] = ok ? u[i] : qnan;
        wr[k] = ok ? residue[i] : qnan;
        if (weights) ww[k] = ok ? weights[i] : qnan;
    }
    if (k < ns) {
        const int64_t j = sidx[k];
        const bool ok = j >= 0 && j < n_src;
        wpe[k] = ok ? pe[j] : qnan;
        wpn[k] = ok ? pn[j] : qnan;
        wpu[k] = ok ? pu[j] : qnan;
    }
}

// residue -= pred; part[b] = sum of the squares of block b's new residues
// (tree order within the block).
__global__ __launch_bounds__(kGbThreads) void gb_update_kernel(double *__restrict__ residue,
                                                               const double *__restrict__ pred,
                                                               int64_t n, double *__restrict__ part) {
    __shared__ double sq[kGbThreads];
    const int64_t i = (int64_t)blockIdx.x * kGbThreads + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const double r = residue[i] - pred[i];
        residue[i] = r;
        v = r * r;
    }
    sq[threadIdx.x] = v;
    __syncthreads();
    for (int s = kGbThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sq[threadIdx.x] += sq[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sq[0];
}

// Sum of squares of x (n values) into part[] (as gb_update_kernel, no update).
__global__ __launch_bounds__(kGbThreads) void gb_sumsq_kernel(const double *__restrict__ x,
                                                              int64_t n, double *__restrict__ part) {
    __shared__ double sq[kGbThreads];
    const int64_t i = (int64_t)blockIdx.x * kGbThreads + threadIdx.x;
    const double v = i < n ? x[i] * x[i] : 0.0;
    sq[threadIdx.x] = v;
    __syncthreads();
    for (int s = kGbThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sq[threadIdx.x] += sq[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sq[0];
}

// out = sqrt(sum(part[0, nb)) / n): one workgroup, each thread a strided
// partial sum, then a tree.
__global__ __launch_bounds__(kGbThreads) void gb_rms_kernel(const double *__restrict__ part,
                                                            int64_t nb, int64_t n,
                                                            double *__restrict__ out) {
    __shared__ double sq[kGbThreads];
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += kGbThreads) v += part[b];
    sq[threadIdx.x] = v;
    __syncthreads();
    for (int s = kGbThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sq[threadIdx.x] += sq[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sqrt(sq[0] / (double)n);
}

__global__ void gb_scatter_add_kernel(double *__restrict__ coefs, int64_t n_src,
                                      const int64_t *__restrict__ sidx,
                                      const double *__restrict__ wc, int64_t ns) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ns) {
        const int64_t j = sidx[k];
        if (j >= 0 && j < n_src) coefs[j] += wc[k];
    }
}

__global__ void gb_copy_kernel(const double *__restrict__ src, double *__restrict__ dst,
                               int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void gb_zero_kernel(double *__restrict__ x, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kGbThreads - 1) / kGbThreads); }

// Sub-buffers of the workspace, in order.
struct GbLayout {
    size_t wd_off[5];  // window data: easting, northing, upward, residue, weights
    size_t ws_off[4];  // window sources: easting, northing, upward, coefficients
    size_t pred_off, part_off, fit_off, fit_bytes, pred_ws_off, pred_ws_bytes;
    // batched factorisation: per slot the window's data (5 arrays) and
    // sources (3), then the batch buffers
    int64_t batch;
    size_t slot_off, slot_stride, batch_off, batch_bytes, total;
};

// Whether the windows are factorised in batches (FP64 mode with damping).
inline bool gb_batched(int mode, double damping, int64_t batch) {
    return batch > 1 && mode == HMX_FIT_FP64 && damping > 0.0;
}

int gb_layout(int64_t n_data, int64_t max_wd, int64_t max_ws, int mode, int64_t batch,
              double damping, GbLayout *L) {
    size_t off = 0;
    for (int k = 0; k < 5; ++k) {
        L->wd_off[k] = off;
        off += align256((size_t)max_wd * sizeof(double));
    }
    for (int k = 0; k < 4; ++k) {
        L->ws_off[k] = off;
        off += align256((size_t)max_ws * sizeof(double));
    }
    L->pred_off = off;
    off += align256((size_t)n_data * sizeof(double));
    L->part_off = off;
    off += align256((size_t)blocks_for(n_data) * sizeof(double));
    size_t fit_bytes = 0, pred_bytes = 0;
    int st = hmx_eqs_fit_workspace_bytes(max_wd, max_ws, mode, &fit_bytes);
    if (st) return st;
    st = hmx_eqs_predict_workspace_bytes(n_data, max_ws, &pred_bytes);
    if (st) return st;
    L->fit_off = off;
    L->fit_bytes = fit_bytes;
    off += align256(fit_bytes);
    L->pred_ws_off = off;
    L->pred_ws_bytes = pred_bytes;
    off += align256(pred_bytes);
    L->batch = gb_batched(mode, damping, batch) ? batch : 1;
    L->slot_off = off;
    L->slot_stride = 0;
    L->batch_off = off;
    L->batch_bytes = 0;
    if (L->batch > 1) {
        L->slot_stride = 5 * align256((size_t)max_wd * sizeof(double)) +
                         3 * align256((size_t)max_ws * sizeof(double));
        off += (size_t)L->batch * L->slot_stride;
        L->batch_off = off;
        L->batch_bytes = eqs_batch_bytes(max_wd, max_ws, L->batch);
        off += align256(L->batch_bytes);
    }
    L->total = off;
    return HMX_OK;
}

// residue of one window's data (as gb_gather_kernel's residue part)
__global__ void gb_gather_residue_kernel(const double *__restrict__ residue, int64_t n_data,
                                         const int64_t *__restrict__ didx, int64_t nd,
                                         double *__restrict__ wr) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nd) {
        const int64_t i = didx[k];
        wr[k] = (i >= 0 && i < n_data) ? residue[i] : __builtin_nan("");
    }
}

}  // namespace
}  // namespace hmx

extern "C" {

int hmx_eqs_gb_fit_workspace_bytes(int64_t n_data, int64_t max_window_data,
                                   int64_t max_window_sources, int precision_mode, double damping,
                                   int64_t batch, size_t *bytes) {
    hmx::clear_error();
    if (!bytes) return hmx::fail(HMX_ERR_ARG, "bytes is NULL");
    if (n_data <= 0 || max_window_data <= 0 || max_window_sources <= 0)
        return hmx::fail(HMX_ERR_ARG, "empty problem");
    hmx::GbLayout L;
    int st = hmx::gb_layout(n_data, max_window_data, max_window_sources, precision_mode, batch,
                            damping, &L);
    if (st) return st;
    *bytes = L.total;
    return HMX_OK;
}

int hmx_eqs_gb_fit(const double *easting, const double *northing, const double *upward,
                   int64_t n_data, const double *data, const double *weights,
                   const double *p_easting, const double *p_northing, const double *p_upward,
                   int64_t n_src, const int64_t *data_index, const int64_t *data_offset,
                   const int64_t *source_index, const int64_t *source_offset, int64_t n_windows,
                   double damping, int precision_mode, int64_t batch, int max_iter, double rtol,
                   double *coefs, double *residue, double *rmse, double *info, void *workspace,
                   size_t workspace_bytes, void *stream) {
    using namespace hmx;
    clear_error();
    if (n_data <= 0 || n_src <= 0) return fail(HMX_ERR_ARG, "empty problem");
    if (n_windows < 0) return fail(HMX_ERR_ARG, "negative number of windows");
    if (!easting || !northing || !upward || !data || !p_easting || !p_northing || !p_upward ||
        !coefs || !residue || !rmse || !info || !data_offset || !source_offset)
        return fail(HMX_ERR_ARG, "NULL array");
    if (n_windows > 0 && (!data_index || !source_index))
        return fail(HMX_ERR_ARG, "NULL window index array");
    // Window sizes from the host offsets; every index range must be
    // non-empty and increasing, and lie inside the index arrays the caller
    // sized by the last offset.
    int64_t max_wd = 1, max_ws = 1;
    if (data_offset[0] != 0 || source_offset[0] != 0)
        return fail(HMX_ERR_ARG, "window offsets must start at 0");
    for (int64_t w = 0; w < n_windows; ++w) {
        const int64_t nd = data_offset[w + 1] - data_offset[w];
        const int64_t ns = source_offset[w + 1] - source_offset[w];
        if (nd <= 0 || ns <= 0) return fail(HMX_ERR_ARG, "window %lld is empty", (long long)w);
        if (nd > n_data || ns > n_src)
            return fail(HMX_ERR_ARG, "window %lld is larger than the problem", (long long)w);
        max_wd = nd > max_wd ? nd : max_wd;
        max_ws = ns > max_ws ? ns : max_ws;
    }
    GbLayout L;
    int st = gb_layout(n_data, max_wd, max_ws, precision_mode, batch, damping, &L);
    if (st) return st;
    if (!workspace || workspace_bytes < L.total)
        return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, L.total);
    hipStream_t s = (hipStream_t)stream;
    char *ws = static_cast<char *>(workspace);
    double *wd[5], *wsrc[4];
    for (int k = 0; k < 5; ++k) wd[k] = reinterpret_cast<double *>(ws + L.wd_off[k]);
    for (int k = 0; k < 4; ++k) wsrc[k] = reinterpret_cast<double *>(ws + L.ws_off[k]);
    double *pred = reinterpret_cast<double *>(ws + L.pred_off);
    double *part = reinterpret_cast<double *>(ws + L.part_off);
    void *fit_ws = ws + L.fit_off;
    void *pred_ws = ws + L.pred_ws_off;
    const unsigned nb = blocks_for(n_data);

    // residue = data, coefs = 0, rmse[0] = RMS of the data
    hipLaunchKernelGGL(gb_copy_kernel, dim3(nb), dim3(kGbThreads), 0, s, data, residue, n_data);
    HMX_LAUNCHED("gb_copy_kernel");
    hipLaunchKernelGGL(gb_zero_kernel, dim3(blocks_for(n_src)), dim3(kGbThreads), 0, s, coefs,
                       n_src);
    HMX_LAUNCHED("gb_zero_kernel");
    hipLaunchKernelGGL(gb_sumsq_kernel, dim3(nb), dim3(kGbThreads), 0, s, data, n_data, part);
    HMX_LAUNCHED("gb_sumsq_kernel");
    hipLaunchKernelGGL(gb_rms_kernel, dim3(1), dim3(kGbThreads), 0, s, part, (int64_t)nb, n_data,
                       rmse);
    HMX_LAUNCHED("gb_rms_kernel");

    double total_iter = 0.0, max_shift = 0.0, total_retries = 0.0;
    double last_update = 0.0, max_diag = 0.0, svd_windows = 0.0, svd_sweeps = 0.0, svd_rank = 0.0;
    double phase[3] = {0.0, 0.0, 0.0};
    // batched factorisation: slot b holds the gathered data (easting,
    // northing, upward, residue, weights) and sources of window w0 + b
    const int64_t B = L.batch;
    auto slot_array = [&](int64_t b, int k) {
        const size_t dbytes = align256((size_t)max_wd * sizeof(double));
        const size_t sbytes = align256((size_t)max_ws * sizeof(double));
        const size_t off = L.slot_off + (size_t)b * L.slot_stride +
                           (k < 5 ? (size_t)k * dbytes : 5 * dbytes + (size_t)(k - 5) * sbytes);
        return reinterpret_cast<double *>(ws + off);
    };
    void *batch_buf = ws + L.batch_off;
    std::vector<EqsWindow> wins;
    std::vector<int> status;
    for (int64_t w0 = 0; w0 < n_windows; w0 += B) {
        const int64_t nbat = B > 1 ? (n_windows - w0 < B ? n_windows - w0 : B) : 0;
        if (nbat > 0) {
            wins.resize((size_t)nbat);
            status.assign((size_t)nbat, 0);
            for (int64_t b = 0; b < nbat; ++b) {
                const int64_t w = w0 + b;
                const int64_t d0 = data_offset[w], nd = data_offset[w + 1] - d0;
                const int64_t s0 = source_offset[w], ns = source_offset[w + 1] - s0;
                const int64_t nk = nd > ns ? nd : ns;
                hipLaunchKernelGGL(gb_gather_kernel, dim3(blocks_for(nk)), dim3(kGbThreads), 0, s,
                                   easting, northing, upward, residue, weights, n_data,
                                   data_index + d0, nd, p_easting, p_northing, p_upward, n_src,
                                   source_index + s0, ns, slot_array(b, 0), slot_array(b, 1),
                                   slot_array(b, 2), slot_array(b, 3), slot_array(b, 4),
                                   slot_array(b, 5), slot_array(b, 6), slot_array(b, 7));
                HMX_LAUNCHED("gb_gather_kernel");
                wins[(size_t)b] = EqsWindow{slot_array(b, 0), slot_array(b, 1), slot_array(b, 2),
                                            weights ? slot_array(b, 4) : nullptr, nd,
                                            slot_array(b, 5), slot_array(b, 6), slot_array(b, 7),
                                            ns};
            }
            double secs[2] = {0.0, 0.0};
            st = eqs_factor_batch(wins.data(), nbat, max_wd, max_ws, damping, batch_buf,
                                  L.batch_bytes, status.data(), secs, s);
            if (st) return st;
            phase[0] += secs[0];
            phase[1] += secs[1];
        }
        const int64_t wend = B > 1 ? w0 + nbat : w0 + 1;
        for (int64_t w = w0; w < wend; ++w) {
            const int64_t d0 = data_offset[w], nd = data_offset[w + 1] - d0;
            const int64_t s0 = source_offset[w], ns = source_offset[w + 1] - s0;
            const int64_t nk = nd > ns ? nd : ns;
            double winfo[11] = {};
            const double *spe = wsrc[0], *spn = wsrc[1], *spu = wsrc[2];
            if (B > 1 && status[(size_t)(w - w0)] == 0) {
                const int64_t b = w - w0;
                hipLaunchKernelGGL(gb_gather_residue_kernel, dim3(blocks_for(nd)),
                                   dim3(kGbThreads), 0, s, residue, n_data, data_index + d0, nd,
                                   slot_array(b, 3));
                HMX_LAUNCHED("gb_gather_residue_kernel");
                st = eqs_solve_factored(wins[(size_t)b], slot_array(b, 3), damping, batch_buf, b,
                                        nbat, max_wd, max_ws, max_iter, rtol, wsrc[3], winfo, s);
                if (st) return st;
                spe = slot_array(b, 5);
                spn = slot_array(b, 6);
                spu = slot_array(b, 7);
            } else {
                hipLaunchKernelGGL(gb_gather_kernel, dim3(blocks_for(nk)), dim3(kGbThreads), 0, s,
                                   easting, northing, upward, residue, weights, n_data,
                                   data_index + d0, nd, p_easting, p_northing, p_upward, n_src,
                                   source_index + s0, ns, wd[0], wd[1], wd[2], wd[3], wd[4],
                                   wsrc[0], wsrc[1], wsrc[2]);
                HMX_LAUNCHED("gb_gather_kernel");
                st = hmx_eqs_fit(wd[0], wd[1], wd[2], nd, wd[3], weights ? wd[4] : nullptr,
                                 wsrc[0], wsrc[1], wsrc[2], ns, damping, precision_mode, max_iter,
                                 rtol, wsrc[3], winfo, fit_ws, L.fit_bytes, stream);
                if (st) return st;
            }
            total_iter += winfo[0];
            last_update = winfo[1];
            max_shift = winfo[2] > max_shift ? winfo[2] : max_shift;
            total_retries += winfo[3];
            max_diag = winfo[4] > max_diag ? winfo[4] : max_diag;
            svd_windows += winfo[5];
            svd_sweeps += winfo[6] < 0 ? -winfo[6] : winfo[6];
            svd_rank += winfo[7];
            for (int k = 0; k < 3; ++k) phase[k] += winfo[8 + k];
            st = hmx_eqs_predict(easting, northing, upward, n_data, spe, spn, spu, wsrc[3], ns,
                                 HMX_F64, pred, pred_ws, L.pred_ws_bytes, stream);
            if (st) return st;
            hipLaunchKernelGGL(gb_update_kernel, dim3(nb), dim3(kGbThreads), 0, s, residue, pred,
                               n_data, part);
            HMX_LAUNCHED("gb_update_kernel");
            hipLaunchKernelGGL(gb_rms_kernel, dim3(1), dim3(kGbThreads), 0, s, part, (int64_t)nb,
                               n_data, rmse + w + 1);
            HMX_LAUNCHED("gb_rms_kernel");
            hipLaunchKernelGGL(gb_scatter_add_kernel, dim3(blocks_for(ns)), dim3(kGbThreads), 0, s,
                               coefs, n_src, source_index + s0, wsrc[3], ns);
            HMX_LAUNCHED("gb_scatter_add_kernel");
        }
    }
    info[0] = total_iter;
    info[1] = last_update;
    info[2] = max_shift;
    info[3] = total_retries;
    info[4] = max_diag;
    info[5] = svd_windows;
    info[6] = svd_sweeps;
    info[7] = svd_rank;
    for (int k = 0; k < 3; ++k) info[8 + k] = phase[k];
    HMX_HIP(hipStreamSynchronize(s));
    return HMX_OK;
}

}  // extern "C"
