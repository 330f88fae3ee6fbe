/*
 * hmx.h -- C-ABI of the MI355X (gfx950) pairwise gravity forward-modelling
 * library (libhmx.so).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status (HMX_OK = 0, negative on error; the message is in
 * hmx_last_error(), thread-local).  Array arguments are DEVICE pointers
 * (HBM) unless a comment says otherwise; `stream` is a hipStream_t (NULL =
 * the default stream) and every call is stream-ordered and asynchronous.
 * Scratch memory ("workspace") is provided by the caller, sized with the
 * matching *_workspace_bytes query, so no entry point allocates device
 * memory.
 *
 * The reference (fatiando/harmonica) has no FFI: its operator interface is
 * a Numba-jitted loop called with numpy arrays plus a jitted per-pair
 * function (a "FIELDS" table entry).  Each entry point below replaces one
 * such loop; the comment names the reference code it replaces.  The Python
 * host layer (harmonica_amd) keeps the reference's Python API and performs
 * its validation before calling here.
 */
#ifndef HMX_H
#define HMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define HMX_OK 0
#define HMX_ERR_ARG (-1)   /* invalid argument (size, pointer, mask, workspace) */
#define HMX_ERR_HIP (-2)   /* a HIP runtime call failed */
#define HMX_ERR_NODEV (-3) /* no usable gfx950 device */
#define HMX_ERR_NUMERIC (-4) /* factorisation failed (matrix not positive definite) */

/* Field bits, in the order of the reference's FIELDS table
 * (src/harmonica/_forward/prisms/gravity.py:37-48).  Outputs hold one row
 * of n_obs values per selected field, in this order, in harmonica's units
 * and signs: potential in J/kg, g_e/g_n/g_z in mGal, tensor terms in
 * Eotvos, g_z/g_ez/g_nz sign-flipped to the downward convention
 * (gravity.py:228-235; point.py:253-260). */
#define HMX_POTENTIAL (1u << 0)
#define HMX_G_E (1u << 1)
#define HMX_G_N (1u << 2)
#define HMX_G_Z (1u << 3)
#define HMX_G_EE (1u << 4)
#define HMX_G_NN (1u << 5)
#define HMX_G_ZZ (1u << 6)
#define HMX_G_EN (1u << 7)
#define HMX_G_EZ (1u << 8)
#define HMX_G_NZ (1u << 9)
#define HMX_ALL_FIELDS 0x3ffu

/* Components of hmx_prism_magnetic, in the order of the reference's
 * field "b" tuple (magnetic.py:195-199). */
#define HMX_B_E (1u << 0)
#define HMX_B_N (1u << 1)
#define HMX_B_U (1u << 2)

/* Coordinate systems of hmx_point_gravity (point.py:231-233). */
#define HMX_CARTESIAN 0
#define HMX_SPHERICAL 1

/* Element types of the equivalent-source jacobian/predict
 * (EquivalentSources(dtype=...), cartesian.py:165-191). */
#define HMX_F64 0
#define HMX_F32 1
#define HMX_BF16 2 /* hmx_mfma_probe only */

/* Precision modes of hmx_eqs_fit. */
#define HMX_FIT_FP64 0      /* J^T W J on FP64 MFMA, FP64 Cholesky, direct solve refined
                               by max_iter CG iterations (0: direct solve only) */
#define HMX_FIT_FP32_MFMA 1 /* J^T W J on FP32 MFMA, FP64 Cholesky, FP64-residual CG */
#define HMX_FIT_BF16_SPLIT 2 /* J^T W J from exact bf16 slice products on MFMA, FP64
                                Cholesky, FP64-residual CG */
#define HMX_FIT_INT8_SPLIT 3 /* J^T W J from exact int8 slice products on int8 MFMA,
                                FP64 Cholesky, FP64-residual CG */
/* Flag OR'ed into the precision mode of hmx_eqs_fit / hmx_eqs_gb_fit:
 * damping=None, i.e. sklearn LinearRegression (scipy.linalg.lstsq with the
 * cut-off max(n, m) * 2^-52) instead of Ridge(alpha=damping); damping must
 * then be 0. */
#define HMX_FIT_NO_DAMPING 0x100

/* Filters of the singular values in hmx_lsq_svd (scikit-learn / SciPy
 * semantics, see hmx_lsq_svd). */
#define HMX_SVD_RIDGE 0       /* s > cutoff (absolute): s / (s^2 + alpha) */
#define HMX_SVD_LSTSQ 1       /* s > cutoff * s_max: 1 / s */
#define HMX_SVD_DUAL_LSTSQ 2  /* s^2 + alpha > cutoff * (s_max^2 + alpha): s / (s^2 + alpha) */

/* Last error message of the calling thread ("" if none). */
const char *hmx_last_error(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int hmx_version(void);

/* Number of visible gfx950 devices (0 if none; status HMX_OK either way). */
int hmx_device_count(int *count);

/* Bytes of workspace needed by hmx_prism_gravity, hmx_prism_layer_gravity
 * and hmx_point_gravity for n_obs observation points, n_sources prisms or
 * point masses and the given field mask, on the current device. */
int hmx_forward_workspace_bytes(int64_t n_obs, int64_t n_sources, uint32_t field_mask,
                                size_t *bytes);

/* Replaces jit_prism_gravity (src/harmonica/_forward/prisms/gravity.py:489-545)
 * called with the choclo per-pair functions of FIELDS (gravity.py:37-48),
 * plus the unit/sign post-processing of prism_gravity (gravity.py:227-236).
 *   easting/northing/upward : (n_obs) observation coordinates [m]
 *   prisms  : (n_prisms, 6) row-major [west, east, south, north, bottom, top]
 *   density : (n_prisms) [kg/m^3]
 *   out     : (popcount(field_mask), n_obs), overwritten
 *   singular_flags : optional uint32 (may be NULL); bit f is OR-ed in when an
 *             observation point is a singular point of tensor field f of some
 *             prism (gravity.py:239-269); that value is NaN.
 * Null prisms are NOT discarded here (the Python layer does it as
 * _discard_null_prisms, gravity.py:452-486, does).
 * Calls with at least 2048 points and 2^30 (point, prism) pairs evaluate the
 * points in a wavefront-coherent strip order sorted into the workspace
 * (also hmx_prism_layer_gravity and hmx_prism_magnetic); out is in the
 * caller's order either way. */
int hmx_prism_gravity(const double *easting, const double *northing, const double *upward,
                      int64_t n_obs, const double *prisms, const double *density,
                      int64_t n_prisms, uint32_t field_mask, double *out,
                      uint32_t *singular_flags, void *workspace, size_t workspace_bytes,
                      void *stream);

/* No counterpart in the reference: hmx_prism_gravity for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  Same arguments and
 * results, without the workspace: the chunk partial sums live in a
 * file-scope device array of 32 MiB (2^22 doubles) per GPU, held until the
 * process ends.  n_obs times the fields of one launch (all of field_mask
 * for the masks with their own kernel: one field, g_e/g_n/g_z, the six
 * tensor terms, all ten; else 1) must not exceed 2^22, else HMX_ERR_ARG;
 * larger sums over prisms use fewer source chunks to fit.  Allocates
 * nothing.  Every call on one GPU shares its array: calls that overlap on
 * one GPU (several host threads or streams) race. */
int hmx_prism_gravity_static(const double *easting, const double *northing,
                             const double *upward, int64_t n_obs, const double *prisms,
                             const double *density, int64_t n_prisms, uint32_t field_mask,
                             double *out, uint32_t *singular_flags, void *stream);

/* Replaces the singular-point scans _any_singular_point_g_{ee,nn,zz,en,ez,nz}
 * of _check_singular_points (src/harmonica/_forward/prisms/gravity.py:239-449):
 * ORs bit f into *flags (device uint32, caller-zeroed) when some observation
 * point lies on a singular point (vertex or edge, gravity.py:141-152) of
 * tensor field f of some prism.  Only tensor bits of field_mask are set. */
int hmx_prism_singular_scan(const double *easting, const double *northing, const double *upward,
                            int64_t n_obs, const double *prisms, int64_t n_prisms,
                            uint32_t field_mask, uint32_t *flags, void *stream);

/* Replaces _forward_gravity_prism_layer (src/harmonica/_forward/prisms/layer.py:522-633)
 * and the post-processing of DatasetAccessorPrismLayer.gravity (layer.py:425-433).
 *   centre_easting (n_east), centre_northing (n_north): prism centres;
 *   bottom, top, density: (n_north, n_east) row-major.
 * Prisms with zero or NaN density, NaN top or bottom, or thickness below
 * thickness_threshold are skipped, as in the reference.  n_east, n_north >= 2. */
int hmx_prism_layer_gravity(const double *easting, const double *northing, const double *upward,
                            int64_t n_obs, const double *centre_easting, int64_t n_east,
                            const double *centre_northing, int64_t n_north,
                            const double *bottom, const double *top, const double *density,
                            double thickness_threshold, uint32_t field_mask, double *out,
                            uint32_t *singular_flags, void *workspace, size_t workspace_bytes,
                            void *stream);

/* Bytes of workspace needed by hmx_prism_magnetic (as
 * hmx_forward_workspace_bytes). */
int hmx_prism_magnetic_workspace_bytes(int64_t n_obs, int64_t n_prisms, uint32_t component_mask,
                                       size_t *bytes);

/* Replaces _jit_prism_magnetic_field (src/harmonica/_forward/prisms/magnetic.py:275-338,
 * field "b") and _jit_prism_magnetic_component (magnetic.py:341-400, one
 * component) with choclo's per-pair prism magnetic functions, plus the nT
 * conversion (magnetic.py:195-197, :271).
 *   prisms : (n_prisms, 6) row-major [west, east, south, north, bottom, top]
 *   magnetization_e/_n/_u : (n_prisms) magnetisation components [A/m]
 *   component_mask : HMX_B_E | HMX_B_N | HMX_B_U (any non-empty subset)
 *   out    : (popcount(component_mask), n_obs) [nT], overwritten
 * A value is NaN where the point lies on an edge or vertex of a prism.
 * Null prisms are NOT discarded here (the Python layer does it as
 * _discard_null_prisms, magnetic.py:403-440, does). */
int hmx_prism_magnetic(const double *easting, const double *northing, const double *upward,
                       int64_t n_obs, const double *prisms, const double *magnetization_e,
                       const double *magnetization_n, const double *magnetization_u,
                       int64_t n_prisms, uint32_t component_mask, double *out, void *workspace,
                       size_t workspace_bytes, void *stream);

/* Bytes of workspace needed by hmx_tesseroid_gravity. */
int hmx_tesseroid_workspace_bytes(int64_t n_obs, int64_t n_tesseroids, size_t *bytes);

/* Replaces jit_tesseroid_gravity and jit_tesseroid_gravity_variable_density
 * (src/harmonica/_forward/tesseroid_gravity.py:236-452) with the adaptive
 * discretisation and quadrature of _tesseroid_utils.py, plus the sign flip
 * and mGal conversion of g_z (tesseroid_gravity.py:207-211).
 *   longitude, latitude [deg], radius [m]: (n_obs) observation points;
 *   tesseroids: (n_tesseroids, 6) row-major [west, east, south, north,
 *     bottom, top], evaluated as given (the Python layer applies the
 *     reference's checks and, with them, the longitude continuity);
 *   density_nodes: (n_tesseroids, 2) density at the lower and upper radial
 *     Gauss-Legendre node (both the tesseroid's density for a constant one);
 *   density_cheb: NULL for constant densities, or (n_tesseroids, 32) the
 *     Chebyshev coefficients of each tesseroid's density over its radial
 *     extent [bottom, top] (sum_k c_k T_k((2 r - bottom - top) / (top -
 *     bottom))), which give the densities at the radial nodes of the parts
 *     that radial_discretization creates (the reference evaluates its
 *     density function there, _tesseroid_variable_density.py);
 *   field: HMX_POTENTIAL or HMX_G_Z; out: (n_obs), overwritten;
 *   overflow: device int, set to 1 (stack of 100 exceeded) or 2 (more than
 *     100000 small tesseroids for a pair) as the reference's OverflowError;
 *     the caller clears it. */
int hmx_tesseroid_gravity(const double *longitude, const double *latitude, const double *radius,
                          int64_t n_obs, const double *tesseroids, const double *density_nodes,
                          const double *density_cheb, int64_t n_tesseroids, uint32_t field,
                          int radial_discretization, double *out, int *overflow, void *workspace,
                          size_t workspace_bytes, void *stream);

/* Replaces point_mass_cartesian / point_mass_spherical
 * (src/harmonica/_forward/point.py:357-454) with the choclo.point kernels
 * (point.py:281-316) or the spherical kernels (point.py:324-354), plus the
 * post-processing of point_gravity (point.py:252-260).
 * Spherical: coordinates are (longitude [deg], latitude [deg], radius [m]) and
 * only HMX_POTENTIAL and HMX_G_Z are valid (point.py:306-315). */
int hmx_point_gravity(const double *c0, const double *c1, const double *c2, int64_t n_obs,
                      const double *p0, const double *p1, const double *p2,
                      const double *masses, int64_t n_points, int coordinate_system,
                      uint32_t field_mask, double *out, void *workspace,
                      size_t workspace_bytes, void *stream);

/* Replaces jacobian_numba_{serial,parallel} (src/harmonica/_equivalent_sources/utils.py:54-74)
 * with greens_func_cartesian (cartesian.py:634-644):
 *   jac[i * ld + j] = 1 / |x_i - p_j|, i < n_data, j < n_src.
 * dtype HMX_F64: all pointers double; HMX_F32: coordinates and jac float. */
int hmx_eqs_jacobian(const void *easting, const void *northing, const void *upward,
                     int64_t n_data, const void *p_easting, const void *p_northing,
                     const void *p_upward, int64_t n_src, int dtype, void *jac, int64_t ld,
                     void *stream);

/* Derivatives of g_z with respect to prism densities and top heights (no
 * counterpart in the reference, whose forward models are not
 * differentiable; these serve the autograd path and Gauss-Newton
 * inversion).  With S(i, p) the g_u vertex sum of prism p at point i
 * (gravity.py closed form without G rho) and F(i, p) the signed sum over
 * the 4 top vertices of atan(x y / (z r)):
 *   d g_u(i) / d rho_p = G S(i, p),  d g_u(i) / d top_p = G rho_p F(i, p).
 * jvp: out_i = sum_p weight_s[p] S(i, p) + weight_f[p] F(i, p)
 *      (weight_f may be NULL);
 * vjp: grad_s[p] = sum_i v_i S(i, p), grad_f[p] = sum_i v_i F(i, p).
 * Raw sums: the caller applies G and harmonica's -1e5 for g_z in mGal.
 * Layer variants take the layer arrays of hmx_prism_layer_gravity; the
 * per-prism weights and gradients are (n_north, n_east) arrays; prisms
 * skipped by the layer rules (NaN density, top or bottom, thickness below
 * the threshold) contribute nothing, zero-density prisms are kept (the
 * derivative with respect to a zero density is not zero). */
int hmx_gz_grad_workspace_bytes(int64_t n_obs, int64_t n_prisms, size_t *bytes);
int hmx_prism_gz_jvp(const double *easting, const double *northing, const double *upward,
                     int64_t n_obs, const double *prisms, const double *weight_s,
                     const double *weight_f, int64_t n_prisms, double *out, void *workspace,
                     size_t workspace_bytes, void *stream);
int hmx_prism_gz_vjp(const double *easting, const double *northing, const double *upward,
                     const double *v, int64_t n_obs, const double *prisms, int64_t n_prisms,
                     double *grad_s, double *grad_f, void *workspace, size_t workspace_bytes,
                     void *stream);
int hmx_prism_layer_gz_jvp(const double *easting, const double *northing, const double *upward,
                           int64_t n_obs, const double *centre_easting, int64_t n_east,
                           const double *centre_northing, int64_t n_north, const double *bottom,
                           const double *top, const double *density, double thickness_threshold,
                           const double *weight_s, const double *weight_f, double *out,
                           void *workspace, size_t workspace_bytes, void *stream);
int hmx_prism_layer_gz_vjp(const double *easting, const double *northing, const double *upward,
                           const double *v, int64_t n_obs, const double *centre_easting,
                           int64_t n_east, const double *centre_northing, int64_t n_north,
                           const double *bottom, const double *top, const double *density,
                           double thickness_threshold, double *grad_s, double *grad_f,
                           void *workspace, size_t workspace_bytes, void *stream);

/* Many independent g_z prism models in one persistent launch (no
 * counterpart in the reference, where each model is one call of
 * jit_prism_gravity, gravity.py:489-545).  Model b owns points
 * [obs_offsets[b], obs_offsets[b+1]) of the concatenated point arrays and
 * prisms [prism_offsets[b], prism_offsets[b+1]) of the concatenated
 * (n, 6) prism array.
 * hmx_batched_plan_size / hmx_batched_plan (HOST arrays): cut every model
 *   into items (model, 256-point tile, chunk of <= prisms_per_item prisms,
 *   slot), 5 int64 each, in decreasing-cost order; item_base[b] is the
 *   first slot of model b and chunks[b] its chunk count.
 * hmx_batched_prism_gz (DEVICE arrays): runs the items with a persistent
 *   grid of blocks_per_cu workgroups per compute unit pulling from an
 *   atomic counter, then writes out (n_obs_total) = g_z in mGal.  Only
 *   kernel launches (graph-capturable). */
int hmx_batched_plan_size(const int64_t *obs_offsets, const int64_t *prism_offsets,
                          int64_t n_models, int64_t prisms_per_item, int64_t *n_items);
int hmx_batched_plan(const int64_t *obs_offsets, const int64_t *prism_offsets, int64_t n_models,
                     int64_t prisms_per_item, int64_t *items, int64_t capacity, int64_t *n_items,
                     int64_t *item_base, int64_t *chunks);
int hmx_batched_workspace_bytes(int64_t n_items, size_t *bytes);
int hmx_batched_prism_gz(const double *easting, const double *northing, const double *upward,
                         const int64_t *obs_offsets, const double *prisms, const double *density,
                         const int64_t *items, int64_t n_items, const int64_t *item_base,
                         const int64_t *chunks, int64_t n_models, int64_t n_obs_total,
                         int blocks_per_cu, double *out, void *workspace, size_t workspace_bytes,
                         void *stream);

/* Spherical equivalent sources (EquivalentSourcesSph,
 * src/harmonica/_equivalent_sources/spherical.py): their Green's function
 * greens_func_spherical (spherical.py:412-424) is 1 / l with l the distance
 * of _forward/utils.py:163-201 between (longitude, latitude, radius)
 * points, which equals 1 / |X - X'| for the geocentric Cartesian points
 *   x = r cos(lat) cos(lon), y = r cos(lat) sin(lon), z = r sin(lat).
 * This writes those points (x, y, z: (n) device outputs [m]; longitude,
 * latitude in degrees); hmx_eqs_jacobian, hmx_eqs_predict and hmx_eqs_fit
 * called on them replace jacobian_numba_* / predict_numba_* with the
 * spherical Green's function and the fit of spherical.py:176-217. */
int hmx_geocentric(const double *longitude, const double *latitude, const double *radius,
                   int64_t n, double *x, double *y, double *z, void *stream);

/* Bytes of workspace needed by hmx_eqs_predict on the current device. */
int hmx_eqs_predict_workspace_bytes(int64_t n_obs, int64_t n_src, size_t *bytes);

/* Replaces predict_numba_{serial,parallel} (src/harmonica/_equivalent_sources/utils.py:77-97):
 *   out[i] = sum_j coefs[j] / |x_i - p_j|  (out overwritten).
 * dtype HMX_F64: all double.  HMX_F32: coordinates, coefs, out float; the
 * pair terms are evaluated in FP32 and accumulated in FP64. */
int hmx_eqs_predict(const void *easting, const void *northing, const void *upward,
                    int64_t n_obs, const void *p_easting, const void *p_northing,
                    const void *p_upward, const void *coefs, int64_t n_src, int dtype,
                    void *out, void *workspace, size_t workspace_bytes, void *stream);

/* Bytes of workspace needed by hmx_eqs_fit (the n_src x n_src normal matrix
 * and its factor, the n_data x n_src scaled Jacobian in the mode's
 * precision, inverses of the diagonal Cholesky blocks and vectors). */
int hmx_eqs_fit_workspace_bytes(int64_t n_data, int64_t n_src, int precision_mode,
                                size_t *bytes);

/* Replaces the fit path of EquivalentSources.fit
 * (src/harmonica/_equivalent_sources/cartesian.py:279-280): the jacobian
 * (utils.py:54-74) followed by verde.base.least_squares (third-party, verde
 * >= 1.8.1): column scaling by the population standard deviation of J
 * (sklearn StandardScaler(with_mean=False)), the weighted least-squares
 * solve (sklearn Ridge(alpha=damping, fit_intercept=False), or with
 * HMX_FIT_NO_DAMPING in precision_mode LinearRegression(fit_intercept=False)
 * for damping=None), unscaling of the coefficients.  The solve follows
 * scikit-learn's routes: the Cholesky factorisation of the normal
 * equations; when it fails, and for damping=None when the factor's pivots
 * or the refinement show an ill-conditioned jacobian, the SVD solve of
 * hmx_lsq_svd on the scaled, weighted jacobian (Ridge's SVD fallback with
 * its 1e-15 cut-off, or its dual-form fallback when n_src > n_data;
 * LinearRegression's lstsq cut-off max(n, m) 2^-52).  The SVD route
 * serves fits of at most 8192 sources and 2^27 jacobian entries; larger
 * fits whose factorisation fails are shifted and refined by CG instead.
 *   data, weights (may be NULL): (n_data) device arrays;
 *   precision_mode: HMX_FIT_FP64 (FP64 MFMA normal matrix, FP64 Cholesky,
 *     direct solve), HMX_FIT_FP32_MFMA (FP32 Jacobian and FP32-MFMA normal
 *     matrix), HMX_FIT_BF16_SPLIT (the scaled Jacobian cut into four
 *     8-bit integer slices per column and group of 1024 rows, normal matrix
 *     from the exact bf16 MFMA products of the slice pairs up to order 3,
 *     about 2^-31 of its entries) or HMX_FIT_INT8_SPLIT (the same with
 *     balanced int8 digits on int8 MFMA, also about 2^-31); in the low-precision
 *     modes the normal matrix, shifted if
 *     needed, is factorised in FP64 and used as the preconditioner of
 *     conjugate gradients on the exact FP64 normal equations;
 *   max_iter, rtol: CG limit and stop rule (relative change of the
 *     data-space prediction per iteration);
 *   coefs: (n_src) device output;
 *   info: HOST array of 11 doubles: [CG iterations, last relative update,
 *     shift tau, factorisation retries, max diagonal of the normal matrix
 *     (set when a shift was needed, else 0),
 *     1 if the SVD route solved the fit, its Jacobi sweeps (negative if
 *     the sweep limit stopped them), its kept singular values, seconds of
 *     the normal matrix (column statistics and jacobian included), of the
 *     factorisation and of the solves and refinement (stream events)].
 * Synchronous: returns after the coefficients are written. */
int hmx_eqs_fit(const double *easting, const double *northing, const double *upward,
                int64_t n_data, const double *data, const double *weights,
                const double *p_easting, const double *p_northing, const double *p_upward,
                int64_t n_src, double damping, int precision_mode, int max_iter, double rtol,
                double *coefs, double *info, void *workspace, size_t workspace_bytes,
                void *stream);

/* Bytes of workspace needed by hmx_lsq_svd for an n x m matrix (the k x k
 * rotation matrix, k = min(n, m), the transpose of G when m > n, three
 * k-vectors and a few scalars). */
int hmx_lsq_svd_workspace_bytes(int64_t n, int64_t m, size_t *bytes);

/* Replaces the SVD solves behind verde.base.least_squares
 * (cartesian.py:280; third-party: scikit-learn 1.7
 * sklearn/linear_model/_ridge.py _solve_svd, the fallback of Ridge when its
 * Cholesky solve fails, and scipy.linalg.lstsq, used by LinearRegression
 * and by Ridge's dual-form fallback):
 *   coef = V diag(f(s)) U^T y,  G = U diag(s) V^T,
 * with the filter f of HMX_SVD_RIDGE (s / (s^2 + damping) for s > cutoff,
 * sklearn's cut-off 1e-15), HMX_SVD_LSTSQ (1 / s for s > cutoff * s_max) or
 * HMX_SVD_DUAL_LSTSQ (s / (s^2 + damping) for s^2 + damping > cutoff *
 * (s_max^2 + damping)).  One-sided Jacobi SVD on the GPU (of G, or of its
 * transpose when m > n).
 *   G: (n x m) column-major device matrix, OVERWRITTEN when n >= m;
 *   y: (n) device vector; coef: (m) device output;
 *   sigma: (m) device output of the min(n, m) singular values (unsorted),
 *     then zeros, or NULL;
 *   info: HOST array of 2 doubles: [Jacobi sweeps (negative if the sweep
 *     limit stopped it), kept components].
 * Synchronous. */
int hmx_lsq_svd(double *G, int64_t n, int64_t m, const double *y, double damping, int filter,
                double cutoff, double *coef, double *sigma, double *info, void *workspace,
                size_t workspace_bytes, void *stream);

/* Bytes of workspace needed by hmx_eqs_gb_fit for n_data data and windows
 * of at most max_window_data data and max_window_sources sources, with the
 * same precision_mode, damping and batch as the fit (batch > 1 adds the
 * buffers of `batch` windows' scaled jacobians, normal matrices and
 * factors, about 8 (max_window_data + max_window_sources) x
 * max_window_sources bytes per window). */
int hmx_eqs_gb_fit_workspace_bytes(int64_t n_data, int64_t max_window_data,
                                   int64_t max_window_sources, int precision_mode, double damping,
                                   int64_t batch, size_t *bytes);

/* Replaces EquivalentSourcesGB._gradient_boosting
 * (src/harmonica/_equivalent_sources/gradient_boosted.py:231-283): for each
 * window in order, the coefficients of its sources by hmx_eqs_fit on the
 * current residue of its data (jacobian + verde least squares, as
 * gradient_boosted.py:257-264), then residue -= the prediction of those
 * sources on every datum (predict_numba_parallel, :266-274) and
 * coefs[window sources] += the window's coefficients (:280).
 *   data_index / source_index: device int64 arrays, the windows' data and
 *     source indices back to back (a window's sources must be distinct);
 *   data_offset / source_offset: HOST int64 arrays of n_windows + 1
 *     offsets into them (every window non-empty);
 *   coefs (n_src), residue (n_data): device outputs, overwritten (coefs
 *     start from zero, residue from the data);
 *   rmse: device output of n_windows + 1 values, the RMS of the residue
 *     before the first window and after each (rmse_per_iteration_, :250,
 *     :278);
 *   info: HOST array of 11 doubles: [total CG iterations, last relative
 *     update, largest shift, total factorisation retries, largest diagonal
 *     (of the windows that needed a shift),
 *     windows solved by the SVD route, their total Jacobi sweeps, their
 *     total kept singular values, and the phase seconds of hmx_eqs_fit
 *     summed over the windows].
 *   batch: in the FP64 mode with damping > 0 and batch > 1, the windows'
 *     normal matrices (which do not depend on the residue) are formed and
 *     factorised `batch` windows at a time, every launch covering the batch,
 *     and each window is then solved from its factor (right-hand side,
 *     triangular solves, CG refinement as hmx_eqs_fit); a window whose
 *     factorisation fails goes through hmx_eqs_fit.  Otherwise each window
 *     is one hmx_eqs_fit call.
 * Indices outside [0, n_data) or [0, n_src) give NaN, never an access out of
 * bounds.  Synchronous. */
int hmx_eqs_gb_fit(const double *easting, const double *northing, const double *upward,
                   int64_t n_data, const double *data, const double *weights,
                   const double *p_easting, const double *p_northing, const double *p_upward,
                   int64_t n_src, const int64_t *data_index, const int64_t *data_offset,
                   const int64_t *source_index, const int64_t *source_offset, int64_t n_windows,
                   double damping, int precision_mode, int64_t batch, int max_iter, double rtol,
                   double *coefs, double *residue, double *rmse, double *info, void *workspace,
                   size_t workspace_bytes, void *stream);

/* Building blocks of the fit, exposed for testing and reuse (no
 * counterpart in the reference):
 * hmx_mfma_probe: one MFMA of a single wavefront; out[4 l + v] is element
 *   v of lane l's accumulator (documents the operand layout).  HMX_F64 and
 *   HMX_F32: 16x16x4 with lane operands a[l], b[l] (64 values each);
 *   HMX_BF16: 16x16x32 with lane operands a[8 l .. 8 l + 7] and
 *   b[8 l .. 8 l + 7] (512 values each, exactly representable in bf16).
 * hmx_gemm_tn: C = beta C + sign A^T B + shift I on MFMA, A (K x M, lda),
 *   B (K x N, ldb) row-major in dtype, C (M x N, ldc) FP64; upper != 0
 *   computes only the upper triangle (M == N).
 * hmx_cholesky_solve: factor the SPD matrix A (m x m, row-major, upper
 *   triangle read) in place as U^T U (FP64, blocked) and overwrite x with
 *   A^-1 x; *info (host) = 0 or the 1-based column of a failed pivot. */
int hmx_mfma_probe(int dtype, const double *a, const double *b, double *out, void *stream);
int hmx_gemm_tn(int dtype, const void *A, int64_t lda, const void *B, int64_t ldb, int64_t K,
                int64_t M, int64_t N, double *C, int64_t ldc, double beta, double sign,
                double shift, int upper, void *stream);
int hmx_cholesky_workspace_bytes(int64_t m, size_t *bytes);
int hmx_cholesky_solve(double *A, int64_t m, double *x, int *info, void *workspace,
                       size_t workspace_bytes, void *stream);

/* Device math routines exposed for accuracy testing (no counterpart in the
 * reference; these are the FP64 sqrt, division, logarithm and
 * principal-value arctangent used inside the pairwise kernels):
 *   out[i] = f(a[i], b[i]),  b unused for HMX_DM_SQRT and HMX_DM_LOG. */
#define HMX_DM_SQRT 0       /* sqrt(a), a >= 0 */
#define HMX_DM_DIV 1        /* a / b */
#define HMX_DM_LOG 2        /* ln(a), a > 0 */
#define HMX_DM_ATAN_RATIO 3 /* atan(a / b), +-pi/2 when b == 0 */
#define HMX_DM_ATAN2 4      /* atan2(a, b) in (-pi, pi], a and b not both 0 */
#define HMX_DM_RSQRT 5      /* 1 / sqrt(a), a > 0 */
int hmx_devmath_eval(int fn, const double *a, const double *b, double *out, int64_t n,
                     void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HMX_H */
