// Derivatives of the prism g_z forward model with respect to prism
// densities and prism top heights (no counterpart in the reference, whose
// path is forward-only).
//
// With S(i, p) = sum over the 8 vertices of (-1)^(i+j+k) K_u (the g_u vertex
// sum of prism p at point i, without G rho) and
// F(i, p) = sum over the 4 top vertices of (-1)^(i+j) A_z (x, y, T - u),
//   d g_u(i) / d rho_p = G S(i, p),   d g_u(i) / d T_p = G rho_p F(i, p):
// only the top vertices depend on T, and summed over them with their signs
// d K_u / d z reduces to A_z (the other terms cancel pairwise across the
// east/west and north/south bounds).  F reuses the top-vertex arctangents
// of S, so one pass yields both.
//
//   gz_jvp: out_i = sum_p a_p S(i, p) + b_p F(i, p)     (point per lane,
//           prisms staged in LDS, chunk partials reduced in order)
//   gz_vjp: ga_p = sum_i v_i S(i, p),  gb_p = sum_i v_i F(i, p)
//           (prism per lane, points staged in LDS, partials per point chunk
//           reduced in order)
// Outputs are the raw sums; the caller applies G and harmonica's -1e5
// (g_z in mGal points down).
#include "hmx_internal.h"
#include "prism_device.h"

namespace hmx {
namespace {

// S and F of one prism at one point.
__device__ __forceinline__ void gz_sums(double e, double n, double u, double W, double E, double S,
                                        double N, double B, double T, const dm::Tables &tab,
                                        double &s_out, double &f_out) {
    s_out = dev::gu_sum_paired<true>(E - e, W - e, N - n, S - n, T - u, B - u, tab, &f_out);
}

struct alignas(16) StagedGradPrism {
    double w, e, s, n, b, t, a, c;  // bounds, weights (a: S, c: F)
};

// Bounds of prism p and whether it is skipped (layer rules minus the zero
// density rule: the derivative with respect to a zero density is not zero).
struct ExplicitGeom {
    const double *prisms;
    __device__ bool get(int64_t p, double *bnd) const {
        for (int c = 0; c < 6; ++c) bnd[c] = prisms[6 * p + c];
        return false;
    }
    // index of prism p in the per-prism arrays
    __device__ int64_t index(int64_t p) const { return p; }
};

struct LayerGeom {
    const double *ce, *cn, *bottom, *top, *density;
    int n_east, n_north;
    double threshold;
    __device__ bool get(int64_t p, double *bnd) const {
        const int j = (int)(p / n_north), k = (int)(p - (int64_t)j * n_north);
        const int64_t idx = (int64_t)k * n_east + j;
        const double half_e = (ce[1] - ce[0]) / 2, half_n = (cn[1] - cn[0]) / 2;
        bnd[0] = ce[j] - half_e;
        bnd[1] = ce[j] + half_e;
        bnd[2] = cn[k] - half_n;
        bnd[3] = cn[k] + half_n;
        bnd[4] = bottom[idx];
        bnd[5] = top[idx];
        const double d = density[idx];
        return d != d || (bnd[5] - bnd[4] < threshold) || bnd[5] != bnd[5] || bnd[4] != bnd[4];
    }
    // layer arrays are (n_north, n_east) row-major
    __device__ int64_t index(int64_t p) const {
        const int j = (int)(p / n_north), k = (int)(p - (int64_t)j * n_north);
        return (int64_t)k * n_east + j;
    }
};

template <class Geom>
__global__ __launch_bounds__(kObsPerBlock) void gz_jvp_kernel(
    const double *__restrict__ oe, const double *__restrict__ on, const double *__restrict__ ou,
    int n_obs, Geom geom, const double *__restrict__ wa, const double *__restrict__ wb,
    int64_t n_src, int64_t chunk_len, double *__restrict__ partial) {
    __shared__ StagedGradPrism tile[kTile];
    __shared__ dm::Tables tables;
    dm::load_tables(&tables);
    const int i = blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = i < n_obs;
    const int ii = valid ? i : n_obs - 1;
    const double e = oe[ii], n = on[ii], u = ou[ii];
    double acc = 0.0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_src ? c0 + chunk_len : n_src;
    for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
        const int cnt = (int)(c1 - t0 < kTile ? c1 - t0 : kTile);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            double bnd[6];
            const bool skip = geom.get(t0 + q, bnd);
            const int64_t idx = geom.index(t0 + q);
            StagedGradPrism sp;
            sp.w = bnd[0];
            sp.e = bnd[1];
            sp.s = bnd[2];
            sp.n = bnd[3];
            sp.b = bnd[4];
            sp.t = bnd[5];
            sp.a = skip ? 0.0 : wa[idx];
            sp.c = skip ? 0.0 : (wb ? wb[idx] : 0.0);
            tile[q] = sp;
        }
        __syncthreads();
        for (int p = 0; p < cnt; ++p) {
            const StagedGradPrism sp = tile[p];
            if (sp.a == 0.0 && sp.c == 0.0) continue;
            double S, F;
            gz_sums(e, n, u, sp.w, sp.e, sp.s, sp.n, sp.b, sp.t, tables, S, F);
            acc = fma(sp.a, S, fma(sp.c, F, acc));
        }
    }
    if (valid) partial[(int64_t)blockIdx.y * n_obs + i] = acc;
}

template <class Geom>
__global__ __launch_bounds__(kObsPerBlock) void gz_vjp_kernel(
    const double *__restrict__ oe, const double *__restrict__ on, const double *__restrict__ ou,
    const double *__restrict__ v, int64_t n_obs, Geom geom, int64_t n_src, int64_t chunk_len,
    double *__restrict__ partial_a, double *__restrict__ partial_b) {
    __shared__ double4 tile[kTile];
    __shared__ dm::Tables tables;
    dm::load_tables(&tables);
    const int64_t p = (int64_t)blockIdx.x * kObsPerBlock + threadIdx.x;
    const bool valid = p < n_src;
    double bnd[6];
    const bool skip = geom.get(valid ? p : n_src - 1, bnd);
    double ga = 0.0, gb = 0.0;
    const int64_t c0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t c1 = c0 + chunk_len < n_obs ? c0 + chunk_len : n_obs;
    for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
        const int cnt = (int)(c1 - t0 < kTile ? c1 - t0 : kTile);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += blockDim.x)
            tile[q] = make_double4(oe[t0 + q], on[t0 + q], ou[t0 + q], v[t0 + q]);
        __syncthreads();
        if (valid && !skip) {
            for (int q = 0; q < cnt; ++q) {
                const double4 o = tile[q];
                double S, F;
                gz_sums(o.x, o.y, o.z, bnd[0], bnd[1], bnd[2], bnd[3], bnd[4], bnd[5], tables, S, F);
                ga = fma(o.w, S, ga);
                gb = fma(o.w, F, gb);
            }
        }
    }
    if (valid) {
        const int64_t idx = geom.index(p);
        partial_a[(int64_t)blockIdx.y * n_src + idx] = ga;
        partial_b[(int64_t)blockIdx.y * n_src + idx] = gb;
    }
}

__global__ void reduce_scaled_kernel(const double *__restrict__ partial, int64_t chunks, int64_t n,
                                     double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sum = 0.0;
    for (int64_t c = 0; c < chunks; ++c) sum += partial[c * n + i];
    out[i] = sum;
}

// Chunks of the point dimension for the VJP: enough workgroups over the
// prism dimension times point chunks to fill the device.
inline int64_t vjp_chunks(int64_t n_src, int64_t n_obs, int n_cu) {
    return forward_chunks(n_src, n_obs, n_cu);
}

template <class Geom>
int run_jvp(const double *e, const double *n, const double *u, int64_t n_obs, const Geom &g,
            const double *wa, const double *wb, int64_t n_src, double *out, void *ws,
            size_t ws_bytes, hipStream_t s) {
    if (n_obs <= 0) return HMX_OK;
    const int cu = device_cu_count();
    if (cu < 0) return cu;
    const int64_t chunks = n_src > 0 ? forward_chunks(n_obs, n_src, cu) : 0;
    const size_t need = (size_t)chunks * (size_t)n_obs * sizeof(double);
    if (chunks > 0) {
        if (!ws || ws_bytes < need)
            return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", ws_bytes, need);
        const int64_t chunk_len = (n_src + chunks - 1) / chunks;
        dim3 grid((unsigned)((n_obs + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
        hipLaunchKernelGGL(gz_jvp_kernel<Geom>, grid, dim3(kObsPerBlock), 0, s, e, n, u, (int)n_obs,
                           g, wa, wb, n_src, chunk_len, static_cast<double *>(ws));
        HMX_LAUNCHED("gz_jvp_kernel");
    }
    hipLaunchKernelGGL(reduce_scaled_kernel, dim3((unsigned)((n_obs + 255) / 256)), dim3(256), 0, s,
                       static_cast<const double *>(ws), chunks, n_obs, out);
    HMX_LAUNCHED("reduce_scaled_kernel");
    return HMX_OK;
}

template <class Geom>
int run_vjp(const double *e, const double *n, const double *u, const double *v, int64_t n_obs,
            const Geom &g, int64_t n_src, double *ga, double *gb, void *ws, size_t ws_bytes,
            hipStream_t s) {
    if (n_src <= 0) return HMX_OK;
    const int cu = device_cu_count();
    if (cu < 0) return cu;
    const int64_t chunks = n_obs > 0 ? vjp_chunks(n_src, n_obs, cu) : 0;
    const size_t need = 2 * (size_t)chunks * (size_t)n_src * sizeof(double);
    double *pa = static_cast<double *>(ws);
    double *pb = pa + chunks * n_src;
    if (chunks > 0) {
        if (!ws || ws_bytes < need)
            return fail(HMX_ERR_ARG, "workspace too small: %zu < %zu bytes", ws_bytes, need);
        const int64_t chunk_len = (n_obs + chunks - 1) / chunks;
        dim3 grid((unsigned)((n_src + kObsPerBlock - 1) / kObsPerBlock), (unsigned)chunks);
        hipLaunchKernelGGL(gz_vjp_kernel<Geom>, grid, dim3(kObsPerBlock), 0, s, e, n, u, v, n_obs, g,
                           n_src, chunk_len, pa, pb);
        HMX_LAUNCHED("gz_vjp_kernel");
    }
    const unsigned rb = (unsigned)((n_src + 255) / 256);
    hipLaunchKernelGGL(reduce_scaled_kernel, dim3(rb), dim3(256), 0, s, pa, chunks, n_src, ga);
    HMX_LAUNCHED("reduce_scaled_kernel");
    hipLaunchKernelGGL(reduce_scaled_kernel, dim3(rb), dim3(256), 0, s, pb, chunks, n_src, gb);
    HMX_LAUNCHED("reduce_scaled_kernel");
    return HMX_OK;
}

int check_sizes(int64_t n_obs, int64_t n_src) {
    if (n_obs < 0 || n_src < 0) return fail(HMX_ERR_ARG, "negative size");
    if (n_obs > INT32_MAX || n_src > INT32_MAX)
        return fail(HMX_ERR_ARG, "more than 2^31-1 observation points or prisms");
    return HMX_OK;
}

}  // namespace
}  // namespace hmx

extern "C" {

int hmx_gz_grad_workspace_bytes(int64_t n_obs, int64_t n_prisms, size_t *bytes) {
    using namespace hmx;
    if (!bytes) return fail(HMX_ERR_ARG, "bytes is NULL");
    int st = check_sizes(n_obs, n_prisms);
    if (st) return st;
    const int cu = device_cu_count();
    if (cu < 0) return cu;
    const size_t jvp = (size_t)forward_chunks(n_obs, n_prisms, cu) * (size_t)n_obs;
    const size_t vjp = 2 * (size_t)vjp_chunks(n_prisms, n_obs, cu) * (size_t)n_prisms;
    *bytes = (jvp > vjp ? jvp : vjp) * sizeof(double);
    return HMX_OK;
}

int hmx_prism_gz_jvp(const double *easting, const double *northing, const double *upward,
                     int64_t n_obs, const double *prisms, const double *weight_s,
                     const double *weight_f, int64_t n_prisms, double *out, void *workspace,
                     size_t workspace_bytes, void *stream) {
    using namespace hmx;
    clear_error();
    int st = check_sizes(n_obs, n_prisms);
    if (st) return st;
    if (n_obs > 0 && (!easting || !northing || !upward || !out))
        return fail(HMX_ERR_ARG, "NULL array");
    if (n_prisms > 0 && (!prisms || !weight_s)) return fail(HMX_ERR_ARG, "NULL prism array");
    ExplicitGeom g{prisms};
    return run_jvp(easting, northing, upward, n_obs, g, weight_s, weight_f, n_prisms, out,
                   workspace, workspace_bytes, (hipStream_t)stream);
}

int hmx_prism_gz_vjp(const double *easting, const double *northing, const double *upward,
                     const double *v, int64_t n_obs, const double *prisms, int64_t n_prisms,
                     double *grad_s, double *grad_f, void *workspace, size_t workspace_bytes,
                     void *stream) {
    using namespace hmx;
    clear_error();
    int st = check_sizes(n_obs, n_prisms);
    if (st) return st;
    if (n_obs > 0 && (!easting || !northing || !upward || !v)) return fail(HMX_ERR_ARG, "NULL array");
    if (n_prisms > 0 && (!prisms || !grad_s || !grad_f)) return fail(HMX_ERR_ARG, "NULL prism array");
    ExplicitGeom g{prisms};
    return run_vjp(easting, northing, upward, v, n_obs, g, n_prisms, grad_s, grad_f, workspace,
                   workspace_bytes, (hipStream_t)stream);
}

int hmx_prism_layer_gz_jvp(const double *easting, const double *northing, const double *upward,
                           int64_t n_obs, const double *centre_easting, int64_t n_east,
                           const double *centre_northing, int64_t n_north, const double *bottom,
                           const double *top, const double *density, double thickness_threshold,
                           const double *weight_s, const double *weight_f, double *out,
                           void *workspace, size_t workspace_bytes, void *stream) {
    using namespace hmx;
    clear_error();
    if (n_east < 2 || n_north < 2)
        return fail(HMX_ERR_ARG, "a prism layer needs at least 2 centres per direction");
    int st = check_sizes(n_obs, n_east * n_north);
    if (st) return st;
    if (n_obs > 0 && (!easting || !northing || !upward || !out))
        return fail(HMX_ERR_ARG, "NULL array");
    if (!centre_easting || !centre_northing || !bottom || !top || !density || !weight_s)
        return fail(HMX_ERR_ARG, "NULL layer array");
    LayerGeom g{centre_easting, centre_northing, bottom, top, density,
                (int)n_east, (int)n_north, thickness_threshold};
    return run_jvp(easting, northing, upward, n_obs, g, weight_s, weight_f, n_east * n_north, out,
                   workspace, workspace_bytes, (hipStream_t)stream);
}

int hmx_prism_layer_gz_vjp(const double *easting, const double *northing, const double *upward,
                           const double *v, int64_t n_obs, const double *centre_easting,
                           int64_t n_east, const double *centre_northing, int64_t n_north,
                           const double *bottom, const double *top, const double *density,
                           double thickness_threshold, double *grad_s, double *grad_f,
                           void *workspace, size_t workspace_bytes, void *stream) {
    using namespace hmx;
    clear_error();
    if (n_east < 2 || n_north < 2)
        return fail(HMX_ERR_ARG, "a prism layer needs at least 2 centres per direction");
    int st = check_sizes(n_obs, n_east * n_north);
    if (st) return st;
    if (n_obs > 0 && (!easting || !northing || !upward || !v)) return fail(HMX_ERR_ARG, "NULL array");
    if (!centre_easting || !centre_northing || !bottom || !top || !density || !grad_s || !grad_f)
        return fail(HMX_ERR_ARG, "NULL layer array");
    LayerGeom g{centre_easting, centre_northing, bottom, top, density,
                (int)n_east, (int)n_north, thickness_threshold};
    return run_vjp(easting, northing, upward, v, n_obs, g, n_east * n_north, grad_s, grad_f,
                   workspace, workspace_bytes, (hipStream_t)stream);
}

}  // extern "C"
