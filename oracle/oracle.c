/*
 * oracle.c -- CPU oracle for the pairwise gravity forward-modelling path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, not the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The harmonica_amd package never calls it.
 *
 * It restates, in plain C99 + OpenMP, the loops of the reference
 * (fatiando/harmonica) and the per-pair closed forms those loops call from
 * the third-party package `choclo` (pinned in the reference as
 * `choclo >= 0.1`, pyproject.toml:41; not vendored, not installed here).
 * The closed forms are restated from the published papers the reference
 * cites (Nagy et al. 2000, 2002; Fukushima 2020; gravity.py:160-164),
 * with the reference's cancellation-avoiding log branch and principal-value
 * arctangent.  Loop structure, loop order and accumulation order follow the
 * reference's Numba loops so that this file is also the "reference CPU
 * path" timed beside the GPU:
 *
 *   or_prism_gravity      src/harmonica/_forward/prisms/gravity.py:489-545
 *   or_prism_any_singular src/harmonica/_forward/prisms/gravity.py:239-449
 *   or_prism_layer        src/harmonica/_forward/prisms/layer.py:522-633
 *   or_point_cartesian    src/harmonica/_forward/point.py:357-398
 *   or_point_spherical    src/harmonica/_forward/point.py:324-354,401-454
 *                         src/harmonica/_forward/utils.py:163-201
 *   or_eqs_jacobian       src/harmonica/_equivalent_sources/utils.py:54-74
 *   or_eqs_predict        src/harmonica/_equivalent_sources/utils.py:77-97
 *                         (Green's function cartesian.py:634-644)
 *   or_eqs_sph_jacobian   src/harmonica/_equivalent_sources/utils.py:54-74
 *   or_eqs_sph_predict    src/harmonica/_equivalent_sources/utils.py:77-97
 *                         (Green's function spherical.py:412-424)
 *   or_tesseroid_gravity  src/harmonica/_forward/tesseroid_gravity.py:236-336,
 *                         _tesseroid_utils.py (discretisation, quadrature)
 *   or_prism_magnetic     src/harmonica/_forward/prisms/magnetic.py:275-400
 *                         (per pair: choclo prism.magnetic_field / _e / _n / _u)
 *
 * Pinning (what checks each function against something other than itself)
 * lives in tests/test_oracle_*.py: reference doctest values, the reference's
 * sample_point_gravity.csv fixture, Gauss-Legendre volume quadrature of the
 * point-mass kernel, Laplace/Poisson, finite differences, the Bouguer slab
 * limit, symmetry and brute force.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* harmonica/constants.py:12 (choclo.constants uses the same CODATA value) */
#define OR_G 6.6743e-11
#define OR_PI_2 1.57079632679489661923

/* Field ids in the reference's FIELDS order (gravity.py:37-48). */
enum {
    OR_POT = 0, OR_E = 1, OR_N = 2, OR_U = 3,
    OR_EE = 4, OR_NN = 5, OR_UU = 6, OR_EN = 7, OR_EU = 8, OR_NU = 9
};

/* Principal value of atan(a / b), defined on b == 0 as +-pi/2 or 0
 * (Fukushima 2020 convention: the "outside" limit on faces). */
static double or_safe_atan(double a, double b)
{
    if (b != 0.0) return atan(a / b);
    if (a > 0.0) return OR_PI_2;
    if (a < 0.0) return -OR_PI_2;
    return 0.0;
}

/* ln(x + r) with r = sqrt(x^2 + y^2 + z^2).  For x < 0 the identity
 * x + r = (y^2 + z^2) / (r - x) avoids the cancellation in x + r.
 * Returns 0 where the argument is zero (y = z = 0 and x <= 0). */
static double or_safe_log(double x, double y, double z, double r)
{
    if (x < 0.0) {
        double s = y * y + z * z;
        if (s == 0.0) return 0.0;
        return log(s / (r - x));
    }
    double a = x + r;
    if (a == 0.0) return 0.0;
    return log(a);
}

/* One vertex term K_field(x, y, z, r); x, y, z are (prism bound - obs). */
static double or_kernel(int field, double x, double y, double z, double r)
{
    switch (field) {
    case OR_POT: {
        double res = 0.0;
        if (x != 0.0 && y != 0.0) res += x * y * or_safe_log(z, x, y, r);
        if (y != 0.0 && z != 0.0) res += y * z * or_safe_log(x, y, z, r);
        if (z != 0.0 && x != 0.0) res += z * x * or_safe_log(y, z, x, r);
        if (x != 0.0) res -= 0.5 * x * x * or_safe_atan(y * z, x * r);
        if (y != 0.0) res -= 0.5 * y * y * or_safe_atan(z * x, y * r);
        if (z != 0.0) res -= 0.5 * z * z * or_safe_atan(x * y, z * r);
        return res;
    }
    case OR_E:
        return -(y * or_safe_log(z, x, y, r) + z * or_safe_log(y, z, x, r)
                 - x * or_safe_atan(y * z, x * r));
    case OR_N:
        return -(z * or_safe_log(x, y, z, r) + x * or_safe_log(z, x, y, r)
                 - y * or_safe_atan(z * x, y * r));
    case OR_U:
        return -(x * or_safe_log(y, z, x, r) + y * or_safe_log(x, y, z, r)
                 - z * or_safe_atan(x * y, z * r));
    case OR_EE: return -or_safe_atan(y * z, x * r);
    case OR_NN: return -or_safe_atan(z * x, y * r);
    case OR_UU: return -or_safe_atan(x * y, z * r);
    case OR_EN: return or_safe_log(z, x, y, r);
    case OR_EU: return or_safe_log(y, z, x, r);
    case OR_NU: return or_safe_log(x, y, z, r);
    }
    return 0.0;
}

/* Singular-point predicates (exact float comparisons). */
static int or_on_easting_edge(double e, double n, double u, const double *p)
{
    return (p[0] <= e && e <= p[1]) && (n == p[2] || n == p[3]) &&
           (u == p[4] || u == p[5]);
}
static int or_on_northing_edge(double e, double n, double u, const double *p)
{
    return (p[2] <= n && n <= p[3]) && (e == p[0] || e == p[1]) &&
           (u == p[4] || u == p[5]);
}
static int or_on_upward_edge(double e, double n, double u, const double *p)
{
    return (p[4] <= u && u <= p[5]) && (e == p[0] || e == p[1]) &&
           (n == p[2] || n == p[3]);
}

/* Is (e, n, u) a singular point of `field` for prism p? */
static int or_is_singular(int field, double e, double n, double u, const double *p)
{
    switch (field) {
    case OR_EE: return or_on_northing_edge(e, n, u, p) || or_on_upward_edge(e, n, u, p);
    case OR_NN: return or_on_easting_edge(e, n, u, p) || or_on_upward_edge(e, n, u, p);
    case OR_UU: return or_on_easting_edge(e, n, u, p) || or_on_northing_edge(e, n, u, p);
    case OR_EN: return or_on_upward_edge(e, n, u, p);
    case OR_EU: return or_on_northing_edge(e, n, u, p);
    case OR_NU: return or_on_easting_edge(e, n, u, p);
    }
    return 0;
}

/* Field of one prism p = [W, E, S, N, B, T] with density rho, in SI units,
 * before harmonica's sign flip and unit conversion.  The vertex loop runs
 * east bound first, north bound first, top bound first, with sign
 * (-1)^(i+j+k). */
double or_prism_pair(int field, double e, double n, double u, const double *p,
                     double rho)
{
    if (or_is_singular(field, e, n, u, p)) return NAN;
    double sum = 0.0;
    for (int i = 0; i < 2; ++i) {
        double x = (i == 0 ? p[1] : p[0]) - e;
        double x2 = x * x;
        for (int j = 0; j < 2; ++j) {
            double y = (j == 0 ? p[3] : p[2]) - n;
            double y2 = y * y;
            for (int k = 0; k < 2; ++k) {
                double z = (k == 0 ? p[5] : p[4]) - u;
                double z2 = z * z;
                double r = sqrt(x2 + y2 + z2);
                double t = or_kernel(field, x, y, z, r);
                sum += ((i + j + k) & 1) ? -t : t;
            }
        }
    }
    return OR_G * rho * sum;
}

/* mu_0 / (4 pi) with mu_0 = 4 pi 1e-7 H/m, evaluated in the order
 * mu_0 / 4 / pi (the magnetic closed forms' constant c_m; choclo's value of
 * mu_0 is not in this image -- see or_prism_magnetic_pair). */
#define OR_PI 3.14159265358979311600
static double or_cm(void)
{
    volatile double mu0 = 4.0 * OR_PI * 1e-7;
    return mu0 / 4.0 / OR_PI;
}

/* Magnetic field (b_e, b_n, b_u), in T, of one prism p = [W, E, S, N, B, T]
 * with uniform magnetisation (me, mn, mu) in A/m, at (e, n, u): for each
 * vertex, with the sign (-1)^(i+j+k) of the gravity sums,
 *   b_e += me K_ee + mn K_en + mu K_eu
 *   b_n += me K_en + mn K_nn + mu K_nu
 *   b_u += me K_eu + mn K_nu + mu K_uu
 * and b *= c_m at the end: B = (mu_0 / 4 pi) grad (M . grad) int 1/r dV
 * with the second-derivative kernels of or_kernel in the east-north-up
 * frame (Poisson's relation; Blakely 1995, ch. 5).  This restates choclo's
 * prism.magnetic_field / magnetic_e / _n / _u, called per pair at
 * src/harmonica/_forward/prisms/magnetic.py:319-331,384-397.  On any edge of
 * the prism (vertices included) all three are NaN. */
void or_prism_magnetic_pair(double e, double n, double u, const double *p,
                            double me, double mn, double mu, double *b)
{
    if (or_on_easting_edge(e, n, u, p) || or_on_northing_edge(e, n, u, p) ||
        or_on_upward_edge(e, n, u, p)) {
        b[0] = b[1] = b[2] = NAN;
        return;
    }
    double be = 0.0, bn = 0.0, bu = 0.0;
    for (int i = 0; i < 2; ++i) {
        double x = (i == 0 ? p[1] : p[0]) - e;
        double x2 = x * x;
        for (int j = 0; j < 2; ++j) {
            double y = (j == 0 ? p[3] : p[2]) - n;
            double y2 = y * y;
            for (int k = 0; k < 2; ++k) {
                double z = (k == 0 ? p[5] : p[4]) - u;
                double r = sqrt(x2 + y2 + z * z);
                double kee = or_kernel(OR_EE, x, y, z, r);
                double knn = or_kernel(OR_NN, x, y, z, r);
                double kuu = or_kernel(OR_UU, x, y, z, r);
                double ken = or_kernel(OR_EN, x, y, z, r);
                double keu = or_kernel(OR_EU, x, y, z, r);
                double knu = or_kernel(OR_NU, x, y, z, r);
                double s = ((i + j + k) & 1) ? -1.0 : 1.0;
                be += s * (me * kee + mn * ken + mu * keu);
                bn += s * (me * ken + mn * knn + mu * knu);
                bu += s * (me * keu + mn * knu + mu * kuu);
            }
        }
    }
    double cm = or_cm();
    b[0] = cm * be;
    b[1] = cm * bn;
    b[2] = cm * bu;
}

static void or_set_threads(int n_threads)
{
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
}

/* gravity.py:489-545: out[i] += f(obs_i, prism_j, rho_j), prisms in index
 * order, parallel over observations.  prisms is (m, 6) row-major.  out is
 * accumulated into (not cleared) exactly as the reference does. */
void or_prism_gravity(const double *e, const double *n, const double *u,
                      int64_t n_obs, const double *prisms, const double *rho,
                      int64_t n_prisms, int field, double *out, int n_threads)
{
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_obs; ++i) {
        double acc = out[i];
        for (int64_t j = 0; j < n_prisms; ++j)
            acc += or_prism_pair(field, e[i], n[i], u[i], prisms + 6 * j, rho[j]);
        out[i] = acc;
    }
}

/* magnetic.py:275-338 (field "b") and :341-400 (one component): out[c][i]
 * += component c of or_prism_magnetic_pair over the prisms in index order,
 * parallel over observations.  comp_mask bit c (0 = b_e, 1 = b_n, 2 = b_u)
 * selects the rows of out, (popcount(comp_mask), n_obs) row-major.  mag is
 * (3, m): the east, north and up components.  Results in T (the caller
 * converts to nT as magnetic.py:195-197,271 do). */
void or_prism_magnetic(const double *e, const double *n, const double *u,
                       int64_t n_obs, const double *prisms, const double *mag,
                       int64_t n_prisms, int comp_mask, double *out, int n_threads)
{
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_obs; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int64_t j = 0; j < n_prisms; ++j) {
            double b[3];
            or_prism_magnetic_pair(e[i], n[i], u[i], prisms + 6 * j, mag[j],
                                   mag[n_prisms + j], mag[2 * n_prisms + j], b);
            for (int c = 0; c < 3; ++c) acc[c] += b[c];
        }
        int row = 0;
        for (int c = 0; c < 3; ++c)
            if (comp_mask & (1 << c)) out[(row++) * n_obs + i] += acc[c];
    }
}

/* gravity.py:239-449 (_any_singular_point_g_*): 1 if any observation point
 * is a singular point of `field` for any prism. */
int or_prism_any_singular(const double *e, const double *n, const double *u,
                          int64_t n_obs, const double *prisms, int64_t n_prisms,
                          int field)
{
    for (int64_t i = 0; i < n_obs; ++i)
        for (int64_t j = 0; j < n_prisms; ++j)
            if (or_is_singular(field, e[i], n[i], u[i], prisms + 6 * j)) return 1;
    return 0;
}

/* layer.py:522-633: prism bounds built on the fly from the centres, loop
 * easting-major then northing, skipping zero/NaN density, thin prisms and
 * NaN tops/bottoms.  bottom/top/rho are (n_north, n_east) row-major. */
void or_prism_layer(const double *e, const double *n, const double *u,
                    int64_t n_obs, const double *c_east, int64_t n_east,
                    const double *c_north, int64_t n_north,
                    const double *bottom, const double *top, const double *rho,
                    double thickness_threshold, int field, double *out,
                    int n_threads)
{
    double half_e = (c_east[1] - c_east[0]) / 2;
    double half_n = (c_north[1] - c_north[0]) / 2;
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_obs; ++i) {
        double acc = out[i];
        for (int64_t jj = 0; jj < n_east; ++jj) {
            double p[6];
            p[0] = c_east[jj] - half_e;
            p[1] = c_east[jj] + half_e;
            for (int64_t k = 0; k < n_north; ++k) {
                int64_t idx = k * n_east + jj;
                double d = rho[idx];
                if (d == 0.0 || isnan(d)) continue;
                double b = bottom[idx], t = top[idx];
                if (t - b < thickness_threshold) continue;
                if (isnan(t) || isnan(b)) continue;
                p[2] = c_north[k] - half_n;
                p[3] = c_north[k] + half_n;
                p[4] = b;
                p[5] = t;
                acc += or_prism_pair(field, e[i], n[i], u[i], p, d);
            }
        }
        out[i] = acc;
    }
}

/* Point-mass kernels in Cartesian coordinates (restated from the point-mass
 * potential G m / l and its derivatives; point.py:124-203 docstring),
 * before harmonica's sign flip and units. */
double or_point_pair(int field, double e, double n, double u, double ep,
                     double np_, double up, double m)
{
    double de = e - ep, dn = n - np_, du = u - up;
    double d = sqrt(de * de + dn * dn + du * du);
    double d3 = d * d * d;
    double d5 = d3 * d * d;
    switch (field) {
    case OR_POT: return OR_G * m / d;
    case OR_E: return -OR_G * m * de / d3;
    case OR_N: return -OR_G * m * dn / d3;
    case OR_U: return -OR_G * m * du / d3;
    case OR_EE: return OR_G * m * (3 * de * de - d * d) / d5;
    case OR_NN: return OR_G * m * (3 * dn * dn - d * d) / d5;
    case OR_UU: return OR_G * m * (3 * du * du - d * d) / d5;
    case OR_EN: return OR_G * m * 3 * de * dn / d5;
    case OR_EU: return OR_G * m * 3 * de * du / d5;
    case OR_NU: return OR_G * m * 3 * dn * du / d5;
    }
    return 0.0;
}

/* The distance of _forward/utils.py:121-201 between two (longitude,
 * latitude, radius) points, angles in degrees: the Green's function of the
 * spherical equivalent sources is its inverse (spherical.py:412-424). */
static double or_distance_spherical(double lon, double lat, double r, double plon,
                                    double plat, double pr)
{
    const double deg = 3.14159265358979323846 / 180.0;
    double la = lat * deg, pla = plat * deg;
    double cosphi_p = cos(pla), sinphi_p = sin(pla);
    double cosphi = cos(la), sinphi = sin(la);
    double coslambda = cos(plon * deg - lon * deg);
    double cospsi = sinphi_p * sinphi + cosphi_p * cosphi * coslambda;
    double dr = r - pr;
    return sqrt(dr * dr + 2.0 * r * pr * (1.0 - cospsi));
}

/* utils.py:54-74 with greens_func_spherical (spherical.py:412-424):
 * jac[i * m + j] = 1 / l(point_i, source_j). */
void or_eqs_sph_jacobian(const double *lon, const double *lat, const double *rad,
                         int64_t n, const double *plon, const double *plat,
                         const double *prad, int64_t m, double *jac, int n_threads)
{
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < m; ++j)
            jac[i * m + j] = 1.0 / or_distance_spherical(lon[i], lat[i], rad[i], plon[j],
                                                         plat[j], prad[j]);
}

/* utils.py:77-97 with greens_func_spherical: out[i] += sum_j c_j / l. */
void or_eqs_sph_predict(const double *lon, const double *lat, const double *rad,
                        int64_t n, const double *plon, const double *plat,
                        const double *prad, const double *coefs, int64_t m,
                        double *out, int n_threads)
{
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double acc = out[i];
        for (int64_t j = 0; j < m; ++j)
            acc += coefs[j] / or_distance_spherical(lon[i], lat[i], rad[i], plon[j], plat[j],
                                                    prad[j]);
        out[i] = acc;
    }
}

/* ---------------------------------------------------------------- tesseroids
 * src/harmonica/_forward/tesseroid_gravity.py:236-336 and
 * _tesseroid_utils.py: per point and tesseroid, adaptive discretisation with
 * a stack (STACK_SIZE 100, at most MAX_DISCRETIZATIONS 100000 small
 * tesseroids), each small tesseroid integrated by 2 x 2 x 2 Gauss-Legendre
 * quadrature as point masses with the spherical point-mass kernels. */
#define OR_TESS_STACK 100
#define OR_TESS_MAX 100000
#define OR_RAD(x) ((x) * (3.14159265358979323846 / 180.0))

/* _tesseroid_utils.py:_tesseroid_dimensions */
static void or_tess_dimensions(const double *t, double *l_lon, double *l_lat, double *l_rad)
{
    double w = OR_RAD(t[0]), e = OR_RAD(t[1]), s = OR_RAD(t[2]), n = OR_RAD(t[3]);
    double top = t[5], bottom = t[4];
    double lat_c = (n + s) / 2;
    *l_lat = top * acos(sin(n) * sin(s) + cos(n) * cos(s));
    double sc = sin(lat_c), cc = cos(lat_c);
    *l_lon = top * acos(sc * sc + cc * cc * cos(e - w));
    *l_rad = top - bottom;
}

/* _tesseroid_utils.py:_distance_tesseroid_point (distance_spherical to the
 * tesseroid's centre) */
static double or_tess_distance(double lon, double lat, double r, const double *t)
{
    return or_distance_spherical(lon, lat, r, (t[0] + t[1]) / 2, (t[2] + t[3]) / 2,
                                 (t[4] + t[5]) / 2);
}

/* _tesseroid_utils.py:gauss_legendre_quadrature, degree 2 in each direction
 * (nodes -+1/sqrt(3), weights 1), with the density at the two radial nodes
 * (rho[0] for the lower node, rho[1] for the upper: the variable-density
 * quadrature of _tesseroid_variable_density.py evaluates the density there). */
static double or_tess_glq(double lon_rad, double cosphi, double sinphi, double r,
                          const double *t, const double *rho, int field)
{
    /* numpy.polynomial.legendre.leggauss(2): nodes -+0x1.279a74590331cp-1, weights 1 */
    static const double node[2] = {-0x1.279a74590331cp-1, 0x1.279a74590331cp-1};
    double w = t[0], e = t[1], s = t[2], n = t[3], bottom = t[4], top = t[5];
    double a_factor = 1.0 / 8.0 * OR_RAD(e - w) * OR_RAD(n - s) * (top - bottom);
    double result = 0.0;
    for (int j = 0; j < 2; ++j) {
        double lat_p = OR_RAD(0.5 * (n - s) * node[j] + 0.5 * (n + s));
        double cosphi_p = cos(lat_p), sinphi_p = sin(lat_p);
        for (int k = 0; k < 2; ++k) {
            double radius_p = 0.5 * (top - bottom) * node[k] + 0.5 * (top + bottom);
            double kappa = radius_p * radius_p * cosphi_p;
            for (int i = 0; i < 2; ++i) {
                double lon_p = OR_RAD(0.5 * (e - w) * node[i] + 0.5 * (e + w));
                double mass = rho[k] * a_factor * kappa * 1.0 * 1.0 * 1.0;
                double coslambda = cos(lon_p - lon_rad);
                double cospsi = sinphi_p * sinphi + cosphi_p * cosphi * coslambda;
                double dr = r - radius_p;
                double dist = sqrt(dr * dr + 2.0 * r * radius_p * (1.0 - cospsi));
                double k_val;
                if (field == OR_POT) {
                    k_val = 1.0 / dist * OR_G;
                } else {
                    double dz = r - radius_p * cospsi;
                    k_val = -OR_G * dz / (dist * dist * dist);
                }
                result += mass * k_val;
            }
        }
    }
    return result;
}

/* The contribution of tesseroid t at the point (adaptive discretisation of
 * _tesseroid_utils.py:_adaptive_discretization then quadrature in the order
 * the small tesseroids are produced).  *overflow gets 1 (stack) or 2 (too
 * many small tesseroids), as the reference's OverflowError. */
static double or_tess_pair(double lon, double lat, double r, double lon_rad, double cosphi,
                           double sinphi, const double *t0, const double *rho, int field,
                           double ratio, int radial, int *overflow)
{
    double stack[OR_TESS_STACK][6];
    int top = 0;
    long n_small = 0;
    double result = 0.0;
    for (int c = 0; c < 6; ++c) stack[0][c] = t0[c];
    while (top >= 0) {
        double t[6];
        for (int c = 0; c < 6; ++c) t[c] = stack[top][c];
        --top;
        double l_lon, l_lat, l_rad;
        or_tess_dimensions(t, &l_lon, &l_lat, &l_rad);
        double d = or_tess_distance(lon, lat, r, t);
        int n_lon = 1, n_lat = 1, n_rad = 1;
        if (d / l_lon < ratio) n_lon = 2;
        if (d / l_lat < ratio) n_lat = 2;
        if (d / l_rad < ratio && radial) n_rad = 2;
        if (n_lon * n_lat * n_rad > 1) {
            if ((top + 1) + n_lon * n_lat * n_rad > OR_TESS_STACK) {
                *overflow = 1;
                return NAN;
            }
            /* _split_tesseroid */
            double d_lon = (t[1] - t[0]) / n_lon, d_lat = (t[3] - t[2]) / n_lat;
            double d_rad = (t[5] - t[4]) / n_rad;
            for (int i = 0; i < n_lon; ++i)
                for (int j = 0; j < n_lat; ++j)
                    for (int k = 0; k < n_rad; ++k) {
                        ++top;
                        stack[top][0] = t[0] + d_lon * i;
                        stack[top][1] = t[0] + d_lon * (i + 1);
                        stack[top][2] = t[2] + d_lat * j;
                        stack[top][3] = t[2] + d_lat * (j + 1);
                        stack[top][4] = t[4] + d_rad * k;
                        stack[top][5] = t[4] + d_rad * (k + 1);
                    }
        } else {
            if (n_small + 1 > OR_TESS_MAX) {
                *overflow = 2;
                return NAN;
            }
            ++n_small;
            result += or_tess_glq(lon_rad, cosphi, sinphi, r, t, rho, field);
        }
    }
    return result;
}

/* tesseroid_gravity.py:236-336: out[i] += sum_j (tesseroid j at point i),
 * field OR_POT or OR_U (SI, before the reference's sign flip and mGal).
 * tess: (m, 6); rho: (m, 2) densities at the lower and upper radial nodes.
 * Returns 0, or 1 / 2 on a stack / discretisation overflow. */
int or_tesseroid_gravity(const double *lon, const double *lat, const double *rad,
                         int64_t n_obs, const double *tess, const double *rho, int64_t m,
                         int field, double ratio, int radial, double *out, int n_threads)
{
    int overflow = 0;
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n_obs; ++i) {
        double lon_rad = OR_RAD(lon[i]);
        double cosphi = cos(OR_RAD(lat[i])), sinphi = sin(OR_RAD(lat[i]));
        double acc = out[i];
        int ov = 0;
        for (int64_t j = 0; j < m && !ov; ++j)
            acc += or_tess_pair(lon[i], lat[i], rad[i], lon_rad, cosphi, sinphi, tess + 6 * j,
                                rho + 2 * j, field, ratio, radial, &ov);
        out[i] = acc;
        if (ov) {
            #pragma omp atomic write
            overflow = ov;
        }
    }
    return overflow;
}

/* The small tesseroids of one (point, tesseroid) pair in production order
 * (for tests of the discretisation); returns their count, -1 / -2 on
 * overflow, at most max_out rows written to out (n, 6). */
long or_tesseroid_split_list(double lon, double lat, double r, const double *t0, double ratio,
                             int radial, int stack_size, double *out, long max_out)
{
    double stack[OR_TESS_STACK][6];
    if (stack_size > OR_TESS_STACK) stack_size = OR_TESS_STACK;
    int top = 0;
    long n_small = 0;
    for (int c = 0; c < 6; ++c) stack[0][c] = t0[c];
    while (top >= 0) {
        double t[6];
        for (int c = 0; c < 6; ++c) t[c] = stack[top][c];
        --top;
        double l_lon, l_lat, l_rad;
        or_tess_dimensions(t, &l_lon, &l_lat, &l_rad);
        double d = or_tess_distance(lon, lat, r, t);
        int n_lon = 1, n_lat = 1, n_rad = 1;
        if (d / l_lon < ratio) n_lon = 2;
        if (d / l_lat < ratio) n_lat = 2;
        if (d / l_rad < ratio && radial) n_rad = 2;
        if (n_lon * n_lat * n_rad > 1) {
            if ((top + 1) + n_lon * n_lat * n_rad > stack_size) return -1;
            double d_lon = (t[1] - t[0]) / n_lon, d_lat = (t[3] - t[2]) / n_lat;
            double d_rad = (t[5] - t[4]) / n_rad;
            for (int i = 0; i < n_lon; ++i)
                for (int j = 0; j < n_lat; ++j)
                    for (int k = 0; k < n_rad; ++k) {
                        ++top;
                        stack[top][0] = t[0] + d_lon * i;
                        stack[top][1] = t[0] + d_lon * (i + 1);
                        stack[top][2] = t[2] + d_lat * j;
                        stack[top][3] = t[2] + d_lat * (j + 1);
                        stack[top][4] = t[4] + d_rad * k;
                        stack[top][5] = t[4] + d_rad * (k + 1);
                    }
        } else {
            if (n_small + 1 > max_out) return -2;
            for (int c = 0; c < 6; ++c) out[6 * n_small + c] = t[c];
            ++n_small;
        }
    }
    return n_small;
}

/* point.py:357-398 */
void or_point_cartesian(const double *e, const double *n, const double *u,
                        int64_t n_obs, const double *pe, const double *pn,
                        const double *pu, const double *mass, int64_t n_pts,
                        int field, double *out, int n_threads)
{
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_obs; ++i) {
        double acc = out[i];
        for (int64_t j = 0; j < n_pts; ++j)
            acc += or_point_pair(field, e[i], n[i], u[i], pe[j], pn[j], pu[j], mass[j]);
        out[i] = acc;
    }
}

/* point.py:401-454 with the kernels of point.py:324-354 and the distance of
 * _forward/utils.py:163-201.  Angles in degrees; field is OR_POT or OR_U. */
void or_point_spherical(const double *lon, const double *lat, const double *rad,
                        int64_t n_obs, const double *plon, const double *plat,
                        const double *prad, const double *mass, int64_t n_pts,
                        int field, double *out, int n_threads)
{
    const double deg = 3.14159265358979323846 / 180.0;
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_obs; ++i) {
        double lo = lon[i] * deg, la = lat[i] * deg;
        double cp = cos(la), sp = sin(la), r = rad[i];
        double acc = out[i];
        for (int64_t j = 0; j < n_pts; ++j) {
            double lop = plon[j] * deg, lap = plat[j] * deg;
            double cpp = cos(lap), spp = sin(lap), rp = prad[j];
            double coslambda = cos(lop - lo);
            double cospsi = spp * sp + cpp * cp * coslambda;
            double dist = sqrt((r - rp) * (r - rp) + 2 * r * rp * (1 - cospsi));
            double k;
            if (field == OR_POT) {
                k = 1 / dist * OR_G;
            } else {
                double dz = r - rp * cospsi;
                k = -OR_G * dz / (dist * dist * dist);
            }
            acc += mass[j] * k;
        }
        out[i] = acc;
    }
}

/* _equivalent_sources/utils.py:54-74 with greens_func_cartesian
 * (cartesian.py:634-644): jac[i, j] = 1 / |x_i - p_j|, (n_data, n_src). */
void or_eqs_jacobian(const double *e, const double *n, const double *u,
                     int64_t n_data, const double *pe, const double *pn,
                     const double *pu, int64_t n_src, double *jac, int n_threads)
{
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_data; ++i)
        for (int64_t j = 0; j < n_src; ++j) {
            double de = e[i] - pe[j], dn = n[i] - pn[j], du = u[i] - pu[j];
            jac[i * n_src + j] = 1 / sqrt(de * de + dn * dn + du * du);
        }
}

/* _equivalent_sources/utils.py:77-97: out[i] += c_j / |x_i - p_j|. */
void or_eqs_predict(const double *e, const double *n, const double *u,
                    int64_t n_obs, const double *pe, const double *pn,
                    const double *pu, const double *coefs, int64_t n_src,
                    double *out, int n_threads)
{
    or_set_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_obs; ++i) {
        double acc = out[i];
        for (int64_t j = 0; j < n_src; ++j) {
            double de = e[i] - pe[j], dn = n[i] - pn[j], du = u[i] - pu[j];
            acc += coefs[j] * (1 / sqrt(de * de + dn * dn + du * du));
        }
        out[i] = acc;
    }
}

/* d g_z / d top for one prism (SI, before harmonica's sign flip): the
 * derivative of the vertex sum with respect to the top bound T.  Only the
 * four top vertices depend on T.  Summed over those four vertices with
 * their signs, d K_u / d z reduces to A_z (the other terms of the
 * derivative cancel pairwise across the east/west and north/south bounds):
 *   d/dT [G rho sum] = G rho sum_{i,j} (-1)^(i+j) safe_atan(x y, z r).
 * Used to check the backward kernel; pinned in the tests by central finite
 * differences of or_prism_pair. */
double or_prism_dgu_dtop(double e, double n, double u, const double *p, double rho)
{
    double sum = 0.0;
    double z = p[5] - u;
    for (int i = 0; i < 2; ++i) {
        double x = (i == 0 ? p[1] : p[0]) - e;
        for (int j = 0; j < 2; ++j) {
            double y = (j == 0 ? p[3] : p[2]) - n;
            double r = sqrt(x * x + y * y + z * z);
            double t = or_safe_atan(x * y, z * r);
            sum += ((i + j) & 1) ? -t : t;
        }
    }
    return OR_G * rho * sum;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
