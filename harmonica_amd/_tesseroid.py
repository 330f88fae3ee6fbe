"""
Drop-in ``tesseroid_gravity``: potential and g_z of tesseroids with the
reference's signature, checks, messages and units, computed by the HIP
kernel of libhmx (``hmx_tesseroid_gravity``: adaptive discretisation and
Gauss-Legendre quadrature per (point, tesseroid) pair on the GPU).

Reference: src/harmonica/_forward/tesseroid_gravity.py:35-213,
_tesseroid_utils.py (checks, longitude continuity, null tesseroids) and
_tesseroid_variable_density.py (density-based radial discretisation, run
on the host as in the reference).
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._eqs import spatial_order
from ._forward import _prepare_coordinates, progressbar_context, to_device
from ._lib import check, lib, ptr, stream_handle

#: Gauss-Legendre node of degree 2 (numpy.polynomial.legendre.leggauss(2)).
GLQ_NODE = float.fromhex("0x1.279a74590331cp-1")
#: Accuracy parameter of the density-based discretisation
#: (_tesseroid_variable_density.py DELTA_RATIO).
DELTA_RATIO = 0.1


def _check_tesseroids(tesseroids):
    """Boundary checks and longitude continuity (_tesseroid_utils.py)."""
    west, east, south, north, bottom, top = tuple(tesseroids[:, i] for i in range(6))
    err_msg = "Invalid tesseroid or tesseroids. "
    invalid = np.logical_or(np.logical_or(south < -90, south > 90),
                            np.logical_or(north < -90, north > 90))
    if invalid.any():
        err_msg += ("The latitudinal boundaries must be inside the [-90, 90] "
                    + "degrees interval.\n")
        for tess in tesseroids[invalid]:
            err_msg += f"\tInvalid tesseroid: {tess}\n"
        raise ValueError(err_msg)
    invalid = south > north
    if invalid.any():
        err_msg += "The south boundary can't be greater than the north one.\n"
        for tess in tesseroids[invalid]:
            err_msg += f"\tInvalid tesseroid: {tess}\n"
        raise ValueError(err_msg)
    invalid = np.logical_or(bottom < 0, top < 0)
    if invalid.any():
        err_msg += "The bottom and top radii should be positive or zero.\n"
        for tess in tesseroids[invalid]:
            err_msg += f"\tInvalid tesseroid: {tess}\n"
        raise ValueError(err_msg)
    invalid = bottom > top
    if invalid.any():
        err_msg += "The bottom radius boundary can't be greater than the top one.\n"
        for tess in tesseroids[invalid]:
            err_msg += f"\tInvalid tesseroid: {tess}\n"
        raise ValueError(err_msg)
    invalid = np.logical_or(np.logical_or(west < -180, west > 360),
                            np.logical_or(east < -180, east > 360))
    if invalid.any():
        err_msg += ("The longitudinal boundaries must be inside the [-180, 360] "
                    + "degrees interval.\n")
        for tess in tesseroids[invalid]:
            err_msg += f"\tInvalid tesseroid: {tess}\n"
        raise ValueError(err_msg)
    if (west > east).any():
        tesseroids = _longitude_continuity(tesseroids)
        west, east = tesseroids[:, 0], tesseroids[:, 1]
    invalid = west > east
    if invalid.any():
        err_msg += "The west boundary can't be greater than the east one.\n"
        for tess in tesseroids[invalid]:
            err_msg += f"\tInvalid tesseroid: {tess}\n"
        raise ValueError(err_msg)
    invalid = east - west > 360
    if invalid.any():
        err_msg += ("The difference between east and west boundaries cannot be greater than "
                    + "one turn around the globe.\n")
        for tess in tesseroids[invalid]:
            err_msg += f"\tInvalid tesseroid: {tess}\n"
        raise ValueError(err_msg)
    return tesseroids


def _longitude_continuity(tesseroids):
    """Move west > east tesseroids to [-180, 180) (_tesseroid_utils.py)."""
    tesseroids = tesseroids.copy()
    west, east = tesseroids[:, 0], tesseroids[:, 1]
    change = west > east
    east[change] = ((east[change] + 180) % 360) - 180
    west[change] = ((west[change] + 180) % 360) - 180
    return tesseroids


def check_points_outside_tesseroids(coordinates, tesseroids, chunk=2048):
    """ValueError listing the (point, tesseroid) pairs with the point inside
    (_tesseroid_utils.py check_points_outside_tesseroids), point-major."""
    longitude, latitude, radius = coordinates
    w, e, s, n, b, t = (tesseroids[:, i][None, :] for i in range(6))
    conflicting = []
    for i0 in range(0, longitude.size, chunk):
        lon = longitude[i0:i0 + chunk, None]
        lat = latitude[i0:i0 + chunk, None]
        rad = radius[i0:i0 + chunk, None]
        lon_360 = lon % 360
        lon_180 = ((lon + 180) % 360) - 180
        inside = (((w < lon_180) & (lon_180 < e)) | ((w < lon_360) & (lon_360 < e))) \
            & (s < lat) & (lat < n) & (b < rad) & (rad < t)
        ii, jj = np.nonzero(inside)
        conflicting.extend(zip((ii + i0).tolist(), jj.tolist()))
    if conflicting:
        err_msg = ("Found computation point(s) inside tesseroid(s). "
                   "Computation points must be outside of tesseroids.\n")
        for i, j in conflicting:
            west, east, south, north, bottom, top = tesseroids[j, :]
            err_msg += (
                f" - Computation point '({longitude[i]}, {latitude[i]}, {radius[i]})' "
                "inside tesseroid "
                f"'({west}, {east}, {south}, {north}, {bottom}, {top})'.\n"
            )
        raise ValueError(err_msg)


def _discard_null_tesseroids(tesseroids, density):
    """Drop zero-volume and zero-density tesseroids (_tesseroid_utils.py)."""
    west, east, south, north, bottom, top = tuple(tesseroids[:, i] for i in range(6))
    null = (west == east) | (south == north) | (bottom == top)
    null[density == 0] = True
    keep = np.logical_not(null)
    return tesseroids[keep, :], density[keep]


# --------------------------------------------- density-based discretisation
def _density_minmax(density, bottom, top):
    from scipy.optimize import minimize_scalar

    lo, hi = np.sort([density(bottom), density(top)])
    kwargs = {"bounds": [bottom, top], "method": "bounded"}
    minimum = np.min((minimize_scalar(density, **kwargs).fun, lo))
    maximum = np.max((-minimize_scalar(lambda r: -density(r), **kwargs).fun, hi))
    return minimum, maximum


def _maximum_absolute_diff(normalized_density, bottom, top):
    from scipy.optimize import minimize_scalar

    d_bottom, d_top = normalized_density(bottom), normalized_density(top)
    slope = (d_top - d_bottom) / (top - bottom)

    def neg_abs_diff(radius):
        return -np.abs(normalized_density(radius) - (slope * (radius - bottom) + d_bottom))

    result = minimize_scalar(neg_abs_diff, bounds=[bottom, top], method="bounded")
    return result.x, -result.fun


def _density_based_discretization(tesseroid, density):
    """Radial splits of one tesseroid until the density is close to linear
    in each piece (_tesseroid_variable_density.py)."""
    w, e, s, n, bottom, top = tesseroid[:]
    density_min, density_max = _density_minmax(density, bottom, top)
    if np.isclose(density_min, density_max):
        return [list(tesseroid)]

    def normalized_density(radius):
        return (density(radius) - density_min) / (density_max - density_min)

    size_original = top - bottom
    pending, pieces = [list(tesseroid)], []
    while pending:
        tess = pending.pop(0)
        bottom, top = tess[-2:]
        radius_split, max_diff = _maximum_absolute_diff(normalized_density, bottom, top)
        if max_diff * (top - bottom) / size_original > DELTA_RATIO:
            pending.append([w, e, s, n, radius_split, top])
            pending.append([w, e, s, n, bottom, radius_split])
        else:
            pieces.append([w, e, s, n, bottom, top])
    return pieces


def density_based_discretization(tesseroids, density):
    pieces = []
    for tesseroid in tesseroids:
        pieces.extend(_density_based_discretization(tesseroid, density))
    return np.atleast_2d(pieces)


def _node_densities(tesseroids, density):
    """Density at the two radial Gauss-Legendre nodes of each tesseroid."""
    bottom, top = tesseroids[:, 4], tesseroids[:, 5]
    out = np.empty((tesseroids.shape[0], 2))
    for k, node in enumerate((-GLQ_NODE, GLQ_NODE)):
        radius_p = 0.5 * (top - bottom) * node + 0.5 * (top + bottom)
        out[:, k] = [density(r) for r in radius_p]
    return out


#: Chebyshev coefficients per tesseroid of the density's radial profile
#: (hmx_tesseroid_gravity evaluates them at the radial nodes of the parts
#: that the radial discretisation creates).
CHEB_N = 32


def _chebyshev_densities(tesseroids, density):
    """(m, CHEB_N) Chebyshev coefficients c_k of each tesseroid's density
    over its radial extent: the interpolant sum_k c_k T_k(x) through the
    density at the CHEB_N Chebyshev points x_j = cos(pi (j + 1/2) / CHEB_N)
    of [bottom, top] (c_k = (2 / N) sum_j f_j T_k(x_j), c_0 halved).  For
    the density profiles of the reference's tests (polynomials, exponentials)
    it matches the function to rounding; the tests state its error."""
    j = np.arange(CHEB_N)
    x = np.cos(np.pi * (j + 0.5) / CHEB_N)
    bottom, top = tesseroids[:, 4], tesseroids[:, 5]
    radii = 0.5 * (top + bottom)[:, None] + 0.5 * (top - bottom)[:, None] * x[None, :]
    f = np.array([[density(r) for r in row] for row in radii], dtype=np.float64)
    T = np.cos(np.outer(np.arange(CHEB_N), np.pi * (j + 0.5) / CHEB_N))  # T_k(x_j)
    c = (2.0 / CHEB_N) * f @ T.T
    c[:, 0] *= 0.5
    return c


def tesseroid_gravity(
    coordinates,
    tesseroids,
    density,
    field,
    parallel=True,
    radial_adaptive_discretization=False,
    dtype=np.float64,
    progressbar=False,
    disable_checks=False,
):
    """
    Gravitational field of tesseroids on computation points.

    Same signature, units and conventions as the reference
    (src/harmonica/_forward/tesseroid_gravity.py:35-213): ``coordinates``
    are (longitude, latitude, radius) in degrees and meters, ``tesseroids``
    rows (west, east, south, north, bottom, top), ``field`` "potential"
    (J/kg) or "g_z" (mGal, downward).  ``density`` is an array (kg/m^3) or a
    function of the radius; a function runs the reference's density-based
    discretisation on the host, which evaluates it at the radial quadrature
    nodes of every piece and, for ``radial_adaptive_discretization``,
    tabulates it per piece as 32 Chebyshev coefficients that the GPU
    evaluates at the nodes of the radial parts (to within 1e-13 of the
    function for smooth profiles).  The adaptive discretisation and the
    quadrature run on the GPU in FP64 (``parallel`` is accepted).
    """
    if field not in ("potential", "g_z"):
        msg = f"Gravitational field {field} not recognized"
        raise ValueError(msg)
    cast, coords = _prepare_coordinates(coordinates)
    tesseroids = np.atleast_2d(np.asarray(tesseroids, dtype=np.float64))
    if not disable_checks:
        # the checks move west > east tesseroids to [-180, 180); without
        # them such tesseroids are evaluated as given, like the reference
        tesseroids = _check_tesseroids(tesseroids)
        check_points_outside_tesseroids(coords, tesseroids)
    cheb = None
    if callable(density):
        tesseroids = density_based_discretization(tesseroids, density)
        rho = _node_densities(tesseroids, density)
        if radial_adaptive_discretization:
            cheb = _chebyshev_densities(tesseroids, density)
    else:
        density = np.atleast_1d(np.asarray(density, dtype=np.float64)).ravel()
        if not disable_checks and density.size != tesseroids.shape[0]:
            raise ValueError(
                f"Number of elements in density ({density.size}) "
                + f"mismatch the number of tesseroids ({tesseroids.shape[0]})"
            )
        tesseroids, density = _discard_null_tesseroids(tesseroids, density)
        rho = np.repeat(density[:, None], 2, axis=1)
    with progressbar_context(coords[0].size, progressbar):
        if cast.size == 0:
            return np.zeros(cast.shape, dtype=dtype)
        # Points in Morton order of (longitude, latitude): a wavefront then
        # holds a compact patch of points, so the lanes that must divide a
        # nearby tesseroid tend to do so together rather than one at a
        # time while the rest of the wave waits.  Each point's sum does not
        # depend on the order of the points.
        order = spatial_order(coords[0].ravel(), coords[1].ravel())
        lon, lat, rad = (to_device(np.ascontiguousarray(c.ravel()[order])) for c in coords)
        n, m = lon.numel(), tesseroids.shape[0]
        T = to_device(np.ascontiguousarray(tesseroids.reshape(-1, 6)))
        D = to_device(np.ascontiguousarray(rho.reshape(-1, 2)))
        C = None if cheb is None else to_device(np.ascontiguousarray(cheb))
        nbytes = ctypes.c_size_t(0)
        check(lib().hmx_tesseroid_workspace_bytes(n, m, ctypes.byref(nbytes)))
        ws = torch.empty(max(nbytes.value, 8), dtype=torch.uint8, device=lon.device)
        out = torch.empty(n, dtype=torch.float64, device=lon.device)
        ovf = torch.zeros(1, dtype=torch.int32, device=lon.device)
        bit = _lib.FIELD_BIT[field]
        check(lib().hmx_tesseroid_gravity(ptr(lon), ptr(lat), ptr(rad), n, ptr(T), ptr(D), ptr(C),
                                          m, bit,
                                          int(bool(radial_adaptive_discretization)), ptr(out),
                                          ptr(ovf), ptr(ws), nbytes.value, stream_handle(None)))
        flag = int(ovf.item())
        if flag == 1:
            raise OverflowError("Stack Overflow. Try to increase the stack size.")
        if flag == 2:
            raise OverflowError("Exceeded maximum discretizations. Please increase the "
                                "MAX_DISCRETIZATIONS.")
        result = np.empty(n)
        result[order] = out.cpu().numpy()
    return result.astype(dtype, copy=False).reshape(cast.shape)
