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
// and L_y, L_z, A_y, A_z by cyclic permutation.  Field slots follow the
// reference's FIELDS table (gravity.py:37-48):
//   potential  x y L_z + y z L_x + z x L_y - (x^2 A_x + y^2 A_y + z^2 A_z) / 2
//   g_e  -(y L_z + z L_y - x A_x)    g_n  -(z L_x + x L_z - y A_y)
//   g_u  -(x L_y + y L_x - z A_z)    g_ee -A_x   g_nn -A_y   g_uu -A_z
//   g_en  L_z    g_eu  L_y    g_nu  L_x
// Tensor terms are NaN on their singular points (gravity.py:141-158).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "devmath.h"

namespace hmx {
namespace dev {

constexpr double kG = 6.6743e-11;  // harmonica/constants.py:12

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

// Principal value of atan(a / b); b == 0 gives +-pi/2 by the sign of a, or 0.
__device__ __forceinline__ double safe_atan(double a, double b, const dm::Tables &tab) {
    return dm::atan_ratio(a, b, tab);
}

// ln(x + r); the x < 0 branch uses (y^2 + z^2) / (r - x); 0 where the
// argument is 0.
__device__ __forceinline__ double safe_log(double x, double y, double z, double r,
                                           const dm::Tables &tab) {
    double arg;
    if (x < 0.0) {
        arg = dm::div(fma(y, y, z * z), r - x);
    } else {
        arg = x + r;
    }
    const double l = dm::log_pos(arg, tab);
    return arg == 0.0 ? 0.0 : l;
}

// Accumulate into vs[] the kernel values of one vertex times its sign sg
// (+1 or -1).
template <unsigned M>
__device__ __forceinline__ void vertex(double x, double y, double z, double x2, double y2,
                                       double z2, double sg, double (&vs)[FieldSet<M>::count],
                                       const dm::Tables &tab) {
    using F = FieldSet<M>;
    const double r = dm::sqrt_nn(x2 + y2 + z2);
    double Lx = 0.0, Ly = 0.0, Lz = 0.0, Ax = 0.0, Ay = 0.0, Az = 0.0;
    if constexpr (F::Lx) Lx = safe_log(x, y, z, r, tab);
    if constexpr (F::Ly) Ly = safe_log(y, z, x, r, tab);
    if constexpr (F::Lz) Lz = safe_log(z, x, y, r, tab);
    if constexpr (F::Ax) Ax = safe_atan(y * z, x * r, tab);
    if constexpr (F::Ay) Ay = safe_atan(z * x, y * r, tab);
    if constexpr (F::Az) Az = safe_atan(x * y, z * r, tab);
    auto add = [&](int f, double k) { vs[F::slot(f)] = fma(sg, k, vs[F::slot(f)]); };
    if constexpr (F::pot) {
        double k = 0.0;
        k += (x != 0.0 && y != 0.0) ? x * y * Lz : 0.0;
        k += (y != 0.0 && z != 0.0) ? y * z * Lx : 0.0;
        k += (z != 0.0 && x != 0.0) ? z * x * Ly : 0.0;
        k -= (x != 0.0) ? 0.5 * x2 * Ax : 0.0;
        k -= (y != 0.0) ? 0.5 * y2 * Ay : 0.0;
        k -= (z != 0.0) ? 0.5 * z2 * Az : 0.0;
        add(0, k);
    }
    if constexpr (F::ge) add(1, -(y * Lz + z * Ly - x * Ax));
    if constexpr (F::gn) add(2, -(z * Lx + x * Lz - y * Ay));
    if constexpr (F::gu) add(3, -(x * Ly + y * Lx - z * Az));
    if constexpr (F::ee) add(4, -Ax);
    if constexpr (F::nn) add(5, -Ay);
    if constexpr (F::uu) add(6, -Az);
    if constexpr (F::en) add(7, Lz);
    if constexpr (F::eu) add(8, Ly);
    if constexpr (F::nu) add(9, Lx);
}

// Vertices evaluated per loop iteration: 4 for one field (register use
// allows 3 waves per SIMD), 2 for several fields (more live accumulators).
template <unsigned M>
constexpr int kVertexUnroll = FieldSet<M>::count == 1 ? 4 : 2;

// Add the fields of one prism (bounds p[0..5] = W, E, S, N, B, T; grho =
// G * density) at (e, n, u) to acc[]; OR singular-point bits into *sflags.
template <unsigned M>
__device__ __forceinline__ void prism_pair(double e, double n, double u, double W, double E,
                                           double S, double N, double B, double T, double grho,
                                           double (&acc)[FieldSet<M>::count], uint32_t &sflags,
                                           const dm::Tables &tab) {
    using F = FieldSet<M>;
    double vs[F::count];
#pragma unroll
    for (int s = 0; s < F::count; ++s) vs[s] = 0.0;
    const double x0 = E - e, x1 = W - e;
    const double y0 = N - n, y1 = S - n;
    const double z0 = T - u, z1 = B - u;
    const double x0s = x0 * x0, x1s = x1 * x1, y0s = y0 * y0, y1s = y1 * y1;
    const double z0s = z0 * z0, z1s = z1 * z1;
    // Vertex order of the reference's loop: east, north, top first; sign
    // (-1)^(i + j + k).  A partially unrolled loop bounds register pressure
    // (occupancy) while leaving the scheduler independent vertices.
#pragma unroll kVertexUnroll<M>
    for (int v = 0; v < 8; ++v) {
        const bool bi = v & 4, bj = v & 2, bk = v & 1;
        const double x = bi ? x1 : x0, xs = bi ? x1s : x0s;
        const double y = bj ? y1 : y0, ys = bj ? y1s : y0s;
        const double z = bk ? z1 : z0, zs = bk ? z1s : z0s;
        const double sg = (__builtin_popcount(v) & 1) ? -1.0 : 1.0;
        vertex<M>(x, y, z, xs, ys, zs, sg, vs, tab);
    }
    if constexpr (F::tensor) {
        const bool in_e = W <= e && e <= E, in_n = S <= n && n <= N, in_u = B <= u && u <= T;
        const bool at_e = e == W || e == E, at_n = n == S || n == N, at_u = u == B || u == T;
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
#pragma unroll
    for (int s = 0; s < F::count; ++s) acc[s] += grho * vs[s];
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
