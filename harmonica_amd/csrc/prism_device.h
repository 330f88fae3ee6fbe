// Device-side closed forms of the gravity fields of a right rectangular
// prism (FP64), shared by the forward, layer, batched and backward kernels.
//
// For an observation point (e, n, u) and a prism [W, E, S, N, B, T] the
// field is G * rho * sum over the 8 vertices of (-1)^(i+j+k) K(x, y, z, r),
// with x = {E, W}[i] - e, y = {N, S}[j] - n, z = {T, B}[k] - u and
// r = sqrt(x^2 + y^2 + z^2); the kernels K (Nagy et al. 2000, 2002;
// Fukushima 2020, as cited at src/harmonica/_forward/prisms/gravity.py:160-164)
// are written with
//   L_x = ln(x + r), evaluated as ln((y^2 + z^2) / (r - x)) for x < 0 to avoid
//         the cancellation of x + r, and 0 where its argument is 0;
//   A_x = atan(y z / (x r)) as a principal value, +-pi/2 or 0 where x r = 0;
// and L_y, L_z, A_y, A_z by cyclic permutation (below, "safe_log" and
// "safe_atan" name these conventions).  Field slots follow the reference's
// FIELDS table (gravity.py:37-48):
//   potential  x y L_z + y z L_x + z x L_y - (x^2 A_x + y^2 A_y + z^2 A_z) / 2
//   g_e  -(y L_z + z L_y - x A_x)    g_n  -(z L_x + x L_z - y A_y)
//   g_u  -(x L_y + y L_x - z A_z)    g_ee -A_x   g_nn -A_y   g_uu -A_z
//   g_en  L_z    g_eu  L_y    g_nu  L_x
// Tensor terms are NaN on their singular points (gravity.py:141-158).
//
// The kernels here do not evaluate the 8 vertices one by one: the terms are
// grouped by the coordinates their coefficients depend on, so that each
// group of logarithms becomes the logarithm of one quotient and each group
// of arctangents one two-argument arctangent (the signed solid angle of a
// face).  The sums are the same; the grouped quantities are formed without
// the cancellation between vertex terms of distant prisms, and they need
// 12 logarithms and 6 arctangents for all ten fields (4 and 2 for g_z)
// instead of 24 and 24 (16 and 8).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "devmath.h"

namespace hmx {
namespace dev {

constexpr double kG = 6.6743e-11;  // harmonica/constants.py:12

// Vertex distance r = sqrt((x^2 + y^2) + zt) with zt = z^2 + 2^-1000
// (z_tiny below, formed once per face).  The 2^-1000 keeps the argument
// positive (sqrt_pos needs s > 0) and changes nothing above 1e-284 m^2; r
// is then ~1e-150 instead of 0 only when the point is the vertex, where
// every grouped term that uses it has a zero coefficient or takes the
// face-plane branch.
__device__ __forceinline__ double z_tiny(double zs) { return zs + 0x1p-1000; }
__device__ __forceinline__ double vertex_r(double xys, double zt) { return dm::sqrt_pos(xys + zt); }

// Compile-time description of a set of fields (bit f = FIELDS[f]).
template <unsigned M>
struct FieldSet {
    static constexpr bool pot = M & 1u, ge = M & 2u, gn = M & 4u, gu = M & 8u;
    static constexpr bool ee = M & 16u, nn = M & 32u, uu = M & 64u;
    static constexpr bool en = M & 128u, eu = M & 256u, nu = M & 512u;
    static constexpr bool Lx = pot || gn || gu || nu;
    static constexpr bool Ly = pot || ge || gu || eu;
    static constexpr bool Lz = pot || ge || gn || en;
    static constexpr bool Ax = pot || ge || ee;
    static constexpr bool Ay = pot || gn || nn;
    static constexpr bool Az = pot || gu || uu;
    static constexpr bool tensor = (M & 0x3f0u) != 0;
    static constexpr int count = __builtin_popcount(M & 0x3ffu);
    // Position of field f among the selected fields.
    static constexpr int slot(int f) { return __builtin_popcount(M & ((1u << f) - 1u)); }
};

// sum_k (-1)^k ln(P(0,k) / P(1,k)) for four safe_log arguments
// P(j,k) = a_j + r_jk if a_j >= 0, else w2_k / (r_jk - a_j), where a is the
// logarithm's own coordinate (a_0 >= a_1) and w2_k = cs + z_k^2 the sum of
// squares of the other two (passed as cs and zt_k = z_k^2 + 2^-1000).  With
// t_jk = r_jk + |a_j| (no cancellation), P = t if a >= 0 and w2 / t if
// a < 0, so the four logarithms reduce to one:
//   a_1 >= 0:         ln(t00 t11 / (t10 t01))
//   a_0 < 0:         -ln(t00 t11 / (t10 t01))
//   a_0 >= 0 > a_1:   ln(t00 t10 w2_1 / (t01 t11 w2_0))   (kGeneral only:
//                     without it the caller guarantees this case is absent).
// Every t is at least 2^-500 and w2 (with its 2^-1000) at least 2^-1000,
// so for a prism of nonzero extent the quotient is a positive normal number
// and the logarithm finite even where the true argument vanishes -- which
// happens only where the logarithm's coefficient is zero, so that term is
// exactly 0 without a test.
// `flip` must be dm::sign_mask(a0 < 0) (shared by the quotients of one axis).
template <bool kGeneral>
__device__ __forceinline__ double log_quad(double a0, double a1, double r00, double r01,
                                           double r10, double r11, double cs, double zt0,
                                           double zt1, unsigned flip, const dm::Tables &tab) {
    const bool neg0 = a0 < 0.0;
    const double t00 = r00 + fabs(a0), t01 = r01 + fabs(a0);
    const double t10 = r10 + fabs(a1), t11 = r11 + fabs(a1);
    double num = t00 * t11, den = t10 * t01;
    if constexpr (kGeneral) {
        const bool mixed = !neg0 && a1 < 0.0;  // the point lies between the two a-planes
        num = mixed ? t00 * (t10 * (cs + zt1)) : num;
        den = mixed ? t01 * (t11 * (cs + zt0)) : den;
    }
    return dm::flip_sign(dm::log_pos(dm::div(num, den), tab), flip);
}

// sum_{i,j} (-1)^(i+j) A_z(i, j) on the face at height z (r_ij the vertex
// distances): the signed solid angle of the face rectangle.  With
// a_i = x_i y / (z r_i) for one y, tan(A_0 - A_1) = (a_0 - a_1) / (1 + a_0 a_1)
// and 1 + a_0 a_1 has the sign of cos(A_0 - A_1) (cos A_i > 0), so each
// difference over i is one two-argument angle (N_j, D_j) (both scaled by
// z^2 r_0 r_1 > 0), and the difference of the two over j is again one angle
// (N_0 D_1 - N_1 D_0, D_0 D_1 + N_0 N_1), correct modulo 2 pi.  The solid
// angle lies in (-pi, pi) unless the point projects inside the rectangle,
// where it has the sign of z and magnitude up to 2 pi.  With kWrap, `inside`
// (the point projects into the rectangle or onto its boundary) selects the
// branch with the sign of z; a point on the boundary never needs it unless
// rounding crossed pi.  Without kWrap the caller guarantees the point does
// not project into the rectangle.  For z = 0 the result is 0 (the angle
// pair is (0, D >= 0)).
// kZeroSafe: z, x_i or y_j may be 0 (then the angle pair may be (0, 0) and
// the result is 0); without it none of them is.
// The angle pair (N, D) of face_angle below: the face's signed solid angle
// is atan2(N, D) modulo 2 pi.
__device__ __forceinline__ void face_nd(double x0, double x1, double y0, double y1, double z,
                                        double r00, double r01, double r10, double r11,
                                        double &N, double &D) {
    const double n0 = y0 * z * fma(x0, r10, -x1 * r00);
    const double d0 = fma(z * z, r00 * r10, x0 * x1 * (y0 * y0));
    const double n1 = y1 * z * fma(x0, r11, -x1 * r01);
    const double d1 = fma(z * z, r01 * r11, x0 * x1 * (y1 * y1));
    N = fma(n0, d1, -n1 * d0);
    D = fma(d0, d1, n0 * n1);
}

template <bool kWrap, bool kZeroSafe = true>
__device__ __forceinline__ double face_angle(double x0, double x1, double y0, double y1, double z,
                                             double r00, double r01, double r10, double r11,
                                             bool inside, const dm::Tables &tab) {
    double N, D;
    face_nd(x0, x1, y0, y1, z, r00, r01, r10, r11, N, D);
    const double a = dm::atan2_full<kZeroSafe>(N, D, tab);
    if constexpr (kWrap) {
        constexpr double kTwoPi = 6.283185307179586232;
        const bool wrap = inside && a * z < 0.0;
        return wrap ? a + copysign(kTwoPi, z) : a;
    }
    return a;
}

// ln(P_0 / P_1) for two safe_log arguments that differ only in the
// logarithm's own coordinate a (a_0 >= a_1), P = a + r (a >= 0) or
// w2 / (r - a) (a < 0), w2 the sum of squares of the other two coordinates.
// With t = r + |a|: P = t for a >= 0 and w2 / t for a < 0, so
//   a_1 >= 0: ln(t_0 / t_1);  a_0 < 0: -ln(t_0 / t_1);  else ln(t_0 t_1 / w2).
// On the line of an edge (w2 = 0) safe_log's arguments are 2a (a > 0) or 0
// (a <= 0, logarithm 0): ln(t_0 / t_1) if a_1 > 0, ln(t_0) if only a_0 > 0,
// else 0.  The result is finite for every input.  Without kGeneral the
// caller guarantees w2 != 0 and that the point is not between the two
// a-planes (a_0 >= 0 > a_1).
// `flip` must be dm::sign_mask(a0 < 0) (shared by the quotients of one axis).
template <bool kGeneral>
__device__ __forceinline__ double log_diff(double a0, double a1, double r0, double r1, double w2,
                                           unsigned flip, const dm::Tables &tab) {
    const double t0 = r0 + fabs(a0), t1 = r1 + fabs(a1);
    double num = t0, den = t1;
    if constexpr (kGeneral) {
        const bool line = w2 == 0.0, mixed = !(a0 < 0.0) && a1 < 0.0;
        num = line ? (a0 > 0.0 ? t0 : 1.0) : (mixed ? t0 * t1 : t0);
        den = line ? (a1 > 0.0 ? t1 : 1.0) : (mixed ? w2 : t1);
        flip = line ? 0u : flip;
    }
    return dm::flip_sign(dm::log_pos(dm::div(num, den), tab), flip);
}

// Signed solid angle sum_{i,j} (-1)^(i+j) atan(a_i b_j / (c r_ij)) of the
// face at height c, including c = 0 where safe_atan gives +-pi/2 by sign.
// Without kGeneral the caller guarantees c != 0 and that the point does
// not project into the face (a_0 >= 0 > a_1 and b_0 >= 0 > b_1 not both).
template <bool kGeneral>
__device__ __forceinline__ double face_sum(double a0, double a1, double b0, double b1, double c,
                                           double r00, double r01, double r10, double r11,
                                           const dm::Tables &tab) {
    if constexpr (!kGeneral) {
        return face_angle<false, false>(a0, a1, b0, b1, c, r00, r01, r10, r11, false, tab);
    } else {
        const bool inside = a0 >= 0.0 && a1 < 0.0 && b0 >= 0.0 && b1 < 0.0;
        const double w = face_angle<true>(a0, a1, b0, b1, c, r00, r01, r10, r11, inside, tab);
        const double sa = (double)((a0 > 0.0) - (a0 < 0.0) - (a1 > 0.0) + (a1 < 0.0));
        const double sb = (double)((b0 > 0.0) - (b0 < 0.0) - (b1 > 0.0) + (b1 < 0.0));
        return c == 0.0 ? 1.5707963267948966 * sa * sb : w;
    }
}

// True on every lane of the wave if any lane has `need` (the lean
// instantiations below are valid only where `need` is false on every lane;
// choosing per wave keeps the branch uniform, so only one path executes).
__device__ __forceinline__ bool any_lane(bool need) { return __any(need) != 0; }

// Sum over the 8 vertices of (-1)^(i+j+k) K_u, K_u = -(x L_y + y L_x - z A_z),
// with the vertices grouped:
//   sum x L_y = sum_i (-1)^i x_i sum_k (-1)^k ln(P(i,0,k) / P(i,1,k))
//   sum y L_x = sum_j (-1)^j y_j sum_k (-1)^k ln(P(0,j,k) / P(1,j,k))
//   sum z A_z = sum_k (-1)^k z_k sum_{i,j} (-1)^(i+j) A_z(i,j,k),
// i.e. 4 logarithms and 2 arctangents per prism instead of 16 and 8.  The
// grouped terms are formed without the cancellation between vertex terms
// of far prisms, so they are also more accurate.
// Terms whose coefficient (x_i, y_j or z_k) is zero are zero, as in the
// per-vertex form (their logarithms and angles are finite, see log_quad and
// face_angle).  With kTop, *ftop receives the top-face sum
// sum_{i,j} (-1)^(i+j) A_z(i, j, 0) (the derivative with respect to the top).
// Without kGeneral the caller guarantees the point lies between neither the
// two x-planes nor the two y-planes (no mixed quotient, no 2 pi branch);
// kGenX lifts the first condition (the logarithms along x take their general
// form; the face angles need the 2 pi branch only between both pairs).
template <bool kTop, bool kGeneral, bool kShare = false, bool kGenX = kGeneral>
__device__ __forceinline__ double gu_sum(double x0, double x1, double y0, double y1, double z0,
                                         double z1, const dm::Tables &tab, double *ftop,
                                         int share = 0, double (*rb)[2] = nullptr) {
    const double ys0 = y0 * y0, ys1 = y1 * y1;
    const double zs0 = z0 * z0, zs1 = z1 * z1;
    const double zt0 = z_tiny(zs0), zt1 = z_tiny(zs1);
    // x^2 + y^2 as one fma (x^2 itself is needed only by the general form).
    const double X[2] = {x0, x1}, YS[2] = {ys0, ys1};
    double r[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const double xy = fma(X[i], X[i], YS[j]);
            r[i][j][0] = vertex_r(xy, zt0);
            if constexpr (!kShare) r[i][j][1] = vertex_r(xy, zt1);
        }
    if constexpr (kShare) {
        // Bottom vertices (k = 1).  share 1: this prism's west edge is the
        // previous prism's east edge (same S, N, B), so r[1][j][1] equals the
        // previous r[0][j][1]; share 2: its south edge is the previous north
        // edge (same W, E, B), so r[i][1][1] equals the previous r[i][0][1].
        // rb holds the previous prism's bottom distances and receives these.
        r[0][0][1] = vertex_r(fma(x0, x0, ys0), zt1);
        if (share == 1) {
            r[1][0][1] = rb[0][0];
            r[1][1][1] = rb[0][1];
            r[0][1][1] = vertex_r(fma(x0, x0, ys1), zt1);
        } else if (share == 2) {
            r[0][1][1] = rb[0][0];
            r[1][1][1] = rb[1][0];
            r[1][0][1] = vertex_r(fma(x1, x1, ys0), zt1);
        } else {
            r[0][1][1] = vertex_r(fma(x0, x0, ys1), zt1);
            r[1][0][1] = vertex_r(fma(x1, x1, ys0), zt1);
            r[1][1][1] = vertex_r(fma(x1, x1, ys1), zt1);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) rb[i][j] = r[i][j][1];
    }
    double sum = 0.0, top = 0.0;
    const unsigned flip_y = dm::sign_mask(y0 < 0.0), flip_x = dm::sign_mask(x0 < 0.0);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double c = i ? x1 : x0, cs = c * c;  // dead unless kGeneral
        const double l = log_quad<kGeneral>(y0, y1, r[i][0][0], r[i][0][1], r[i][1][0],
                                            r[i][1][1], cs, zt0, zt1, flip_y, tab);
        sum = fma(i ? -c : c, l, sum);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double c = j ? y1 : y0, cs = j ? ys1 : ys0;
        const double l = log_quad<kGeneral || kGenX>(x0, x1, r[0][j][0], r[0][j][1], r[1][j][0],
                                            r[1][j][1], cs, zt0, zt1, flip_x, tab);
        sum = fma(j ? -c : c, l, sum);
    }
    // The point projects into the face or onto its boundary (see face_angle).
    const bool inside = x0 >= 0.0 && x1 < 0.0 && y0 >= 0.0 && y1 < 0.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double z = k ? z1 : z0;
        const double w = face_angle<kGeneral>(x0, x1, y0, y1, z, r[0][0][k], r[0][1][k],
                                              r[1][0][k], r[1][1][k], inside, tab);
        sum = fma(k ? z : -z, w, sum);  // w = 0 for z = 0
        if constexpr (kTop) {
            if (k == 0) {
                // z_T = 0: safe_atan(x y, 0) = +-pi/2 by the sign of x y, or 0.
                const double sx = (x0 > 0.0) - (x0 < 0.0) - ((x1 > 0.0) - (x1 < 0.0));
                const double sy = (y0 > 0.0) - (y0 < 0.0) - ((y1 > 0.0) - (y1 < 0.0));
                top = z == 0.0 ? 1.5707963267948966 * sx * sy : w;
            }
        }
    }
    if constexpr (kTop) *ftop = top;
    return -sum;
}

// gu_sum with the lean instantiation unless a lane of the wave lies between
// the x-planes or the y-planes of the prism.
template <bool kTop>
__device__ __forceinline__ double gu_sum_paired(double x0, double x1, double y0, double y1,
                                                double z0, double z1, const dm::Tables &tab,
                                                double *ftop) {
    if (any_lane((x0 >= 0.0 && x1 < 0.0) || (y0 >= 0.0 && y1 < 0.0)))
        return gu_sum<kTop, true>(x0, x1, y0, y1, z0, z1, tab, ftop);
    return gu_sum<kTop, false>(x0, x1, y0, y1, z0, z1, tab, ftop);
}

// The fields of FieldSet<M> of one prism (relative bounds x0 = E - e,
// x1 = W - e, y0 = N - n, y1 = S - n, z0 = T - u, z1 = B - u) into vs[],
// singular-point bits into sflags.  Without kGeneral the caller guarantees
// that the point is on no face plane of the prism and between the two
// planes of no axis; kGenX lifts the last condition for the x-planes (the
// logarithms along x take their general form; the face angles need theirs
// only between two pairs of planes, the singular points lie on face planes).
template <unsigned M, bool kGeneral, bool kGenX = kGeneral>
__device__ __forceinline__ void prism_fields(double x0, double x1, double y0, double y1,
                                             double z0, double z1, double (&vs)[FieldSet<M>::count],
                                             uint32_t &sflags, const dm::Tables &tab) {
    using F = FieldSet<M>;
    const double xs[2] = {x0 * x0, x1 * x1}, ys[2] = {y0 * y0, y1 * y1};
    const double zs[2] = {z0 * z0, z1 * z1};
    const double X[2] = {x0, x1}, Y[2] = {y0, y1}, Z[2] = {z0, z1};
    double r[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) r[i][j][k] = vertex_r(xs[i] + ys[j], z_tiny(zs[k]));
    // Vertex sums grouped by the coordinates each term depends on:
    //   sum (-1)^(i+j+k) f(y_j, z_k) L_x = sum_{j,k} (-1)^(j+k) f(y_j, z_k) DLx[j][k]
    //   with DLx[j][k] = L_x(0,j,k) - L_x(1,j,k) (one logarithm), and likewise
    //   DLy[i][k], DLz[i][j];
    //   sum (-1)^(i+j+k) f(x_i) A_x = sum_i (-1)^i f(x_i) Ox[i], Ox[i] the
    //   face solid angle (one arctangent), and likewise Oy[j], Oz[k].
    double DLx[2][2] = {}, DLy[2][2] = {}, DLz[2][2] = {}, Ox[2] = {}, Oy[2] = {}, Oz[2] = {};
    const unsigned fx = dm::sign_mask(x0 < 0.0), fy = dm::sign_mask(y0 < 0.0);
    const unsigned fz = dm::sign_mask(z0 < 0.0);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            if constexpr (F::Lx)
                DLx[a][b] = log_diff<kGeneral || kGenX>(x0, x1, r[0][a][b], r[1][a][b],
                                                        ys[a] + zs[b], fx,
                                               tab);
            if constexpr (F::Ly)
                DLy[a][b] = log_diff<kGeneral>(y0, y1, r[a][0][b], r[a][1][b], xs[a] + zs[b], fy,
                                               tab);
            if constexpr (F::Lz)
                DLz[a][b] = log_diff<kGeneral>(z0, z1, r[a][b][0], r[a][b][1], xs[a] + ys[b], fz,
                                               tab);
        }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if constexpr (F::Ax)
            Ox[a] = face_sum<kGeneral>(y0, y1, z0, z1, X[a], r[a][0][0], r[a][0][1], r[a][1][0],
                                       r[a][1][1], tab);
        if constexpr (F::Ay)
            Oy[a] = face_sum<kGeneral>(z0, z1, x0, x1, Y[a], r[0][a][0], r[1][a][0], r[0][a][1],
                                       r[1][a][1], tab);
        if constexpr (F::Az)
            Oz[a] = face_sum<kGeneral>(x0, x1, y0, y1, Z[a], r[0][0][a], r[0][1][a], r[1][0][a],
                                       r[1][1][a], tab);
    }
    // Signed sums with one coefficient (or none) per grouped quantity.
    auto s2 = [](const double (&d)[2][2], double c0, double c1, bool by_first) {
        // sum_{a,b} (-1)^(a+b) c_(a or b) d[a][b]
        return by_first ? c0 * (d[0][0] - d[0][1]) - c1 * (d[1][0] - d[1][1])
                        : c0 * (d[0][0] - d[1][0]) - c1 * (d[0][1] - d[1][1]);
    };
    auto s1 = [](const double (&o)[2], double c0, double c1) { return c0 * o[0] - c1 * o[1]; };
    if constexpr (F::pot) {
        double k = x0 * (y0 * DLz[0][0] - y1 * DLz[0][1]) - x1 * (y0 * DLz[1][0] - y1 * DLz[1][1]);
        k += y0 * (z0 * DLx[0][0] - z1 * DLx[0][1]) - y1 * (z0 * DLx[1][0] - z1 * DLx[1][1]);
        k += x0 * (z0 * DLy[0][0] - z1 * DLy[0][1]) - x1 * (z0 * DLy[1][0] - z1 * DLy[1][1]);
        k -= 0.5 * (s1(Ox, xs[0], xs[1]) + s1(Oy, ys[0], ys[1]) + s1(Oz, zs[0], zs[1]));
        vs[F::slot(0)] = k;
    }
    if constexpr (F::ge)
        vs[F::slot(1)] = -(s2(DLz, y0, y1, false) + s2(DLy, z0, z1, false) - s1(Ox, x0, x1));
    if constexpr (F::gn)
        vs[F::slot(2)] = -(s2(DLx, z0, z1, false) + s2(DLz, x0, x1, true) - s1(Oy, y0, y1));
    if constexpr (F::gu)
        vs[F::slot(3)] = -(s2(DLy, x0, x1, true) + s2(DLx, y0, y1, true) - s1(Oz, z0, z1));
    if constexpr (F::ee) vs[F::slot(4)] = -(Ox[0] - Ox[1]);
    if constexpr (F::nn) vs[F::slot(5)] = -(Oy[0] - Oy[1]);
    if constexpr (F::uu) vs[F::slot(6)] = -(Oz[0] - Oz[1]);
    if constexpr (F::en) vs[F::slot(7)] = s2(DLz, 1.0, 1.0, true);
    if constexpr (F::eu) vs[F::slot(8)] = s2(DLy, 1.0, 1.0, true);
    if constexpr (F::nu) vs[F::slot(9)] = s2(DLx, 1.0, 1.0, true);
    if constexpr (F::tensor && kGeneral) {
        // Singular points (gravity.py:141-158): on an edge of the prism,
        // in x = E - e etc. (x0 = 0 is e == E, x1 <= 0 <= x0 is W <= e <= E).
        const bool in_e = x1 <= 0.0 && 0.0 <= x0, in_n = y1 <= 0.0 && 0.0 <= y0;
        const bool in_u = z1 <= 0.0 && 0.0 <= z0;
        const bool at_e = x0 == 0.0 || x1 == 0.0, at_n = y0 == 0.0 || y1 == 0.0;
        const bool at_u = z0 == 0.0 || z1 == 0.0;
        const bool edge_e = in_e && at_n && at_u;  // on an easting-parallel edge
        const bool edge_n = in_n && at_e && at_u;  // on a northing-parallel edge
        const bool edge_u = in_u && at_e && at_n;  // on an upward edge
        const double qnan = __builtin_nan("");
        auto sing = [&](int f, bool s) {
            if (s) {
                vs[F::slot(f)] = qnan;
                sflags |= 1u << f;
            }
        };
        if constexpr (F::ee) sing(4, edge_n || edge_u);
        if constexpr (F::nn) sing(5, edge_e || edge_u);
        if constexpr (F::uu) sing(6, edge_e || edge_n);
        if constexpr (F::en) sing(7, edge_u);
        if constexpr (F::eu) sing(8, edge_n);
        if constexpr (F::nu) sing(9, edge_e);
    }
}

// The second-derivative vertex sums K[f - 4] (f = 4..9: ee, nn, uu, en,
// eu, nu) of the gravity field bits in TM, for a point that lies between the
// two planes of no axis and on no face plane of the prism (relative bounds
// as prism_fields).  There every logarithm of one axis takes the same branch
// of log_diff and no face is seen from inside, so the sums regroup further:
//  * K_en = sum (-1)^(i+j+k) L_z = +-ln(prod_even t / prod_odd t), t = r + |z|
//    at the vertices of even and odd i + j + k, the sign that of log_diff for
//    the z-axis: one logarithm instead of four.  K_eu (along y) and K_nu
//    (along x) likewise.
//  * K_ee = -(O_x(x0) - O_x(x1)): the two face angles have the sign of
//    x0 (= that of x1) and magnitude below pi, so their difference lies in
//    (-pi, pi) and is one two-argument angle formed from the faces' (N, D)
//    pairs: one arctangent instead of two.  K_nn, K_uu likewise; with all
//    three, K_uu = -(K_ee + K_nn) (Laplace's equation outside the prism).
template <unsigned TM>
__device__ __forceinline__ void tensor_lean(double x0, double x1, double y0, double y1, double z0,
                                            double z1, double (&K)[6], const dm::Tables &tab) {
    using F = FieldSet<TM>;
    const double xs[2] = {x0 * x0, x1 * x1}, ys[2] = {y0 * y0, y1 * y1};
    const double zs[2] = {z0 * z0, z1 * z1};
    const double X[2] = {x0, x1}, Y[2] = {y0, y1}, Z[2] = {z0, z1};
    double r[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) r[i][j][k] = vertex_r(xs[i] + ys[j], z_tiny(zs[k]));
    // +-ln(prod over even vertices of (r + |a|) / prod over odd ones)
    auto off = [&](auto coord, double a0) {
        double num = 1.0, den = 1.0;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double t = r[i][j][k] + fabs(coord(i, j, k));
                    if ((i + j + k) & 1)
                        den *= t;
                    else
                        num *= t;
                }
        return dm::flip_sign(dm::log_pos(dm::div(num, den), tab), dm::sign_mask(a0 < 0.0));
    };
    if constexpr (F::en) K[3] = off([&](int, int, int k) { return Z[k]; }, z0);
    if constexpr (F::eu) K[4] = off([&](int, int j, int) { return Y[j]; }, y0);
    if constexpr (F::nu) K[5] = off([&](int i, int, int) { return X[i]; }, x0);
    // -(A(face 0) - A(face 1)) from the two faces' angle pairs
    auto diag = [&](double N0, double D0, double N1, double D1) {
        return dm::atan2_full<false>(fma(N1, D0, -N0 * D1), fma(D0, D1, N0 * N1), tab);
    };
    constexpr bool kLaplace = F::ee && F::nn && F::uu;
    if constexpr (F::ee) {
        double N0, D0, N1, D1;
        face_nd(y0, y1, z0, z1, x0, r[0][0][0], r[0][0][1], r[0][1][0], r[0][1][1], N0, D0);
        face_nd(y0, y1, z0, z1, x1, r[1][0][0], r[1][0][1], r[1][1][0], r[1][1][1], N1, D1);
        K[0] = diag(N0, D0, N1, D1);
    }
    if constexpr (F::nn) {
        double N0, D0, N1, D1;
        face_nd(z0, z1, x0, x1, y0, r[0][0][0], r[1][0][0], r[0][0][1], r[1][0][1], N0, D0);
        face_nd(z0, z1, x0, x1, y1, r[0][1][0], r[1][1][0], r[0][1][1], r[1][1][1], N1, D1);
        K[1] = diag(N0, D0, N1, D1);
    }
    if constexpr (F::uu) {
        if constexpr (kLaplace) {
            K[2] = -(K[0] + K[1]);
        } else {
            double N0, D0, N1, D1;
            face_nd(x0, x1, y0, y1, z0, r[0][0][0], r[0][1][0], r[1][0][0], r[1][1][0], N0, D0);
            face_nd(x0, x1, y0, y1, z1, r[0][0][1], r[0][1][1], r[1][0][1], r[1][1][1], N1, D1);
            K[2] = diag(N0, D0, N1, D1);
        }
    }
    (void)X;
    (void)Y;
    (void)Z;
}

// tensor_lean for a point between the two x-planes of the prism but between
// no other pair and on no face plane: the logarithms along x change branch
// across the point, so K_nu is the sum of the four grouped x-logarithms of
// prism_fields (general form), and the two x-face angles have opposite
// signs, so K_ee is their difference taken angle by angle (neither face
// is seen from inside: that needs the point between the y- and z-planes
// too).  The other kernels keep the lean regrouping, which needs only their
// own axis: K_en, K_eu one logarithm each, K_nn (and K_uu, or Laplace's
// equation outside the prism) one arctangent each.
template <unsigned TM>
__device__ __forceinline__ void tensor_xonly(double x0, double x1, double y0, double y1,
                                             double z0, double z1, double (&K)[6],
                                             const dm::Tables &tab) {
    using F = FieldSet<TM>;
    const double xs[2] = {x0 * x0, x1 * x1}, ys[2] = {y0 * y0, y1 * y1};
    const double zs[2] = {z0 * z0, z1 * z1};
    const double Y[2] = {y0, y1}, Z[2] = {z0, z1};
    double r[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) r[i][j][k] = vertex_r(xs[i] + ys[j], z_tiny(zs[k]));
    auto off = [&](auto coord, double a0) {
        double num = 1.0, den = 1.0;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double t = r[i][j][k] + fabs(coord(i, j, k));
                    if ((i + j + k) & 1)
                        den *= t;
                    else
                        num *= t;
                }
        return dm::flip_sign(dm::log_pos(dm::div(num, den), tab), dm::sign_mask(a0 < 0.0));
    };
    if constexpr (F::en) K[3] = off([&](int, int, int k) { return Z[k]; }, z0);
    if constexpr (F::eu) K[4] = off([&](int, int j, int) { return Y[j]; }, y0);
    if constexpr (F::nu) {
        const unsigned fx = dm::sign_mask(x0 < 0.0);
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const double d =
                    log_diff<true>(x0, x1, r[0][j][k], r[1][j][k], ys[j] + zs[k], fx, tab);
                s = (j + k) & 1 ? s - d : s + d;
            }
        K[5] = s;
    }
    auto diag = [&](double N0, double D0, double N1, double D1) {
        return dm::atan2_full<false>(fma(N1, D0, -N0 * D1), fma(D0, D1, N0 * N1), tab);
    };
    constexpr bool kLaplace = F::ee && F::nn && F::uu;
    if constexpr (F::ee) {
        const double o0 = face_sum<false>(y0, y1, z0, z1, x0, r[0][0][0], r[0][0][1], r[0][1][0],
                                          r[0][1][1], tab);
        const double o1 = face_sum<false>(y0, y1, z0, z1, x1, r[1][0][0], r[1][0][1], r[1][1][0],
                                          r[1][1][1], tab);
        K[0] = -(o0 - o1);
    }
    if constexpr (F::nn) {
        double N0, D0, N1, D1;
        face_nd(z0, z1, x0, x1, y0, r[0][0][0], r[1][0][0], r[0][0][1], r[1][0][1], N0, D0);
        face_nd(z0, z1, x0, x1, y1, r[0][1][0], r[1][1][0], r[0][1][1], r[1][1][1], N1, D1);
        K[1] = diag(N0, D0, N1, D1);
    }
    if constexpr (F::uu) {
        if constexpr (kLaplace) {
            K[2] = -(K[0] + K[1]);
        } else {
            double N0, D0, N1, D1;
            face_nd(x0, x1, y0, y1, z0, r[0][0][0], r[0][1][0], r[1][0][0], r[1][1][0], N0, D0);
            face_nd(x0, x1, y0, y1, z1, r[0][0][1], r[0][1][1], r[1][0][1], r[1][1][1], N1, D1);
            K[2] = diag(N0, D0, N1, D1);
        }
    }
}

// Whether the point (e, n, u) needs the general instantiation for this
// prism: for g_u, g_e, g_n alone when it lies between the two planes of one
// of the two axes the logarithms run along; for other field sets when it
// lies between the two planes of any axis or on a face plane (a product of
// the two differences that is 0, which includes underflow: the general
// path is exact everywhere).
template <unsigned M>
__device__ __forceinline__ bool needs_general(double e, double n, double u, double W, double E,
                                              double S, double N, double B, double T) {
    const double x0 = E - e, x1 = W - e, y0 = N - n, y1 = S - n, z0 = T - u, z1 = B - u;
    const bool bx = x0 >= 0.0 && x1 < 0.0, by = y0 >= 0.0 && y1 < 0.0;
    const bool bz = z0 >= 0.0 && z1 < 0.0;
    if constexpr (M == 0x8u) return bx || by;
    if constexpr (M == 0x2u) return by || bz;
    if constexpr (M == 0x4u) return bz || bx;
    return bx || by || bz || x0 * x1 == 0.0 || y0 * y1 == 0.0 || z0 * z1 == 0.0;
}

// Field sets with an instantiation for points between the x-planes only (on
// no face plane): g_u alone, those evaluated by prism_fields and the
// tensor-only sets (tensor_xonly).
template <unsigned M>
__host__ __device__ constexpr bool has_x_only() {
    return M != 0x2u && M != 0x4u;
}

// Whether the point needs the general instantiation for a reason other than
// lying between the x-planes (has_x_only field sets): between the y-planes
// (g_u alone), or between the y- or z-planes or on a face plane.
template <unsigned M>
__device__ __forceinline__ bool needs_full_general(double e, double n, double u, double W,
                                                   double E, double S, double N, double B,
                                                   double T) {
    const double x0 = E - e, x1 = W - e, y0 = N - n, y1 = S - n, z0 = T - u, z1 = B - u;
    const bool by = y0 >= 0.0 && y1 < 0.0, bz = z0 >= 0.0 && z1 < 0.0;
    if constexpr (M == 0x8u) return by;
    return by || bz || x0 * x1 == 0.0 || y0 * y1 == 0.0 || z0 * z1 == 0.0;
}

// Add the fields of one prism (bounds W, E, S, N, B, T; grho = G * density)
// at (e, n, u) to acc[]; OR singular-point bits into sflags.  Without
// kGeneral, needs_general<M> must be false for the point, or with kGenX
// (has_x_only field sets) needs_full_general<M>.
template <unsigned M, bool kGeneral, bool kGenX = kGeneral>
__device__ __forceinline__ void prism_pair_path(double e, double n, double u, double W, double E,
                                                double S, double N, double B, double T,
                                                double grho, double (&acc)[FieldSet<M>::count],
                                                uint32_t &sflags, const dm::Tables &tab) {
    using F = FieldSet<M>;
    if constexpr (M == 0x8u) {  // g_u alone
        acc[0] += grho * gu_sum<false, kGeneral, false, kGenX>(E - e, W - e, N - n, S - n, T - u,
                                                               B - u, tab, nullptr);
    } else if constexpr (M == 0x2u) {  // g_e alone: K_e is K_u with (x, y, z) -> (y, z, x)
        acc[0] += grho * gu_sum<false, kGeneral>(N - n, S - n, T - u, B - u, E - e, W - e, tab,
                                                 nullptr);
    } else if constexpr (M == 0x4u) {  // g_n alone: K_n is K_u with (x, y, z) -> (z, x, y)
        acc[0] += grho * gu_sum<false, kGeneral>(T - u, B - u, E - e, W - e, N - n, S - n, tab,
                                                 nullptr);
    } else if constexpr ((M & 0xfu) == 0 && !kGeneral) {  // tensor terms only
        double K[6];
        if constexpr (kGenX)
            tensor_xonly<M>(E - e, W - e, N - n, S - n, T - u, B - u, K, tab);
        else
            tensor_lean<M>(E - e, W - e, N - n, S - n, T - u, B - u, K, tab);
#pragma unroll
        for (int f = 4; f < 10; ++f)
            if ((M >> f) & 1u) acc[F::slot(f)] += grho * K[f - 4];
    } else {
        double vs[F::count];
        prism_fields<M, kGeneral, kGenX>(E - e, W - e, N - n, S - n, T - u, B - u, vs, sflags,
                                         tab);
#pragma unroll
        for (int s = 0; s < F::count; ++s) acc[s] += grho * vs[s];
    }
}

// prism_pair_path with the instantiation chosen per prism and wave.
template <unsigned M>
__device__ __forceinline__ void prism_pair(double e, double n, double u, double W, double E,
                                           double S, double N, double B, double T, double grho,
                                           double (&acc)[FieldSet<M>::count], uint32_t &sflags,
                                           const dm::Tables &tab) {
    if (any_lane(needs_general<M>(e, n, u, W, E, S, N, B, T)))
        prism_pair_path<M, true>(e, n, u, W, E, S, N, B, T, grho, acc, sflags, tab);
    else
        prism_pair_path<M, false>(e, n, u, W, E, S, N, B, T, grho, acc, sflags, tab);
}

// Accumulate the prisms tile[0, cnt) (LDS, read as wave-uniform broadcasts;
// P has w, e, s, n, b, t, grho, int share and, with kMaySkip, int skip) at
// (e, n, u).
// Pass 1 evaluates the lean form for every prism that no lane of the wave
// needs the general form for and marks the others in wave-uniform 128-bit
// masks: those for which the x-planes alone made the lean form unusable
// (has_x_only field sets; for the 64 points of a wave along a grid row,
// the prisms of the 64 columns the points fall in) and the rest.  Pass 2
// evaluates the first with the x-only form, pass 3 the rest with the
// general form.
// Separate loops keep the forms' register allocations apart.  The order of
// the additions is fixed by the prisms and the points, so results are
// deterministic.
template <unsigned M, bool kMaySkip, class P>
__device__ __forceinline__ void tile_pairs(const P *tile, int cnt, double e, double n, double u,
                                           double (&acc)[FieldSet<M>::count], uint32_t &sflags,
                                           const dm::Tables &tab) {
    uint64_t general_lo = 0, general_hi = 0, xonly_lo = 0, xonly_hi = 0;
    // g_u alone: bottom vertex distances of the previous prism evaluated in
    // this pass, reused where P::share says an edge is shared.
    double rb[2][2];
    bool prev_lean = false;
    for (int p = 0; p < cnt; ++p) {
        const P sp = tile[p];
        if constexpr (kMaySkip) {
            if (sp.skip) {
                prev_lean = false;
                continue;
            }
        }
        if (any_lane(needs_general<M>(e, n, u, sp.w, sp.e, sp.s, sp.n, sp.b, sp.t))) {
            bool full = true;
            if constexpr (has_x_only<M>())
                full = any_lane(needs_full_general<M>(e, n, u, sp.w, sp.e, sp.s, sp.n, sp.b, sp.t));
            uint64_t &lo = full ? general_lo : xonly_lo, &hi = full ? general_hi : xonly_hi;
            if (p < 64)
                lo |= 1ull << p;
            else
                hi |= 1ull << (p - 64);
            prev_lean = false;
            continue;
        }
        if constexpr (M == 0x8u) {
            const int share = prev_lean ? __builtin_amdgcn_readfirstlane(sp.share) : 0;
            acc[0] += sp.grho * gu_sum<false, false, true>(sp.e - e, sp.w - e, sp.n - n, sp.s - n,
                                                          sp.t - u, sp.b - u, tab, nullptr,
                                                          share, rb);
            prev_lean = true;
        } else {
            prism_pair_path<M, false>(e, n, u, sp.w, sp.e, sp.s, sp.n, sp.b, sp.t, sp.grho, acc,
                                      sflags, tab);
        }
    }
    if constexpr (has_x_only<M>()) {
        for (int half = 0; half < 2; ++half) {
            uint64_t m = half ? xonly_hi : xonly_lo;
            while (m) {
                const int p = 64 * half + __builtin_ctzll(m);
                m &= m - 1;
                const P sp = tile[p];
                prism_pair_path<M, false, true>(e, n, u, sp.w, sp.e, sp.s, sp.n, sp.b, sp.t,
                                                sp.grho, acc, sflags, tab);
            }
        }
    }
    for (int half = 0; half < 2; ++half) {
        uint64_t m = half ? general_hi : general_lo;
        while (m) {
            const int p = 64 * half + __builtin_ctzll(m);
            m &= m - 1;
            const P sp = tile[p];
            prism_pair_path<M, true>(e, n, u, sp.w, sp.e, sp.s, sp.n, sp.b, sp.t, sp.grho, acc,
                                     sflags, tab);
        }
    }
}

// Edge sharing between tile[q - 1] and tile[q] for the g_u bottom-vertex
// reuse: 1 if q's west edge is q - 1's east edge (same S, N, B), 2 if q's
// south edge is q - 1's north edge (same W, E, B), else 0.
template <class P>
__device__ __forceinline__ int shared_edge(const P &prev, const P &cur) {
    if (cur.w == prev.e && cur.s == prev.s && cur.n == prev.n && cur.b == prev.b) return 1;
    if (cur.s == prev.n && cur.w == prev.w && cur.e == prev.e && cur.b == prev.b) return 2;
    return 0;
}

// Harmonica's sign flip and unit conversion of field f (gravity.py:228-235).
__device__ __forceinline__ double to_harmonica_units(int f, double v) {
    if (f == 3 || f == 8 || f == 9) v = -v;
    if (f >= 1 && f <= 3) v *= 1e5;
    if (f >= 4) v *= 1e9;
    return v;
}

}  // namespace dev
}  // namespace hmx
