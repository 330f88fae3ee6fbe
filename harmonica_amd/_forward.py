"""
Drop-in forward models: ``prism_gravity`` and ``point_gravity`` with the
reference's signatures, validation, messages, warnings and unit/sign
conventions, computed by the HIP kernels of libhmx.

Reference: src/harmonica/_forward/prisms/gravity.py:51-236 (prism_gravity),
src/harmonica/_forward/prisms/utils.py:12-43 (check_prisms),
src/harmonica/_forward/point.py:30-316 (point_gravity),
src/harmonica/_forward/utils.py:71-87,333-392 (coordinate-system check,
progress bar).
"""

import contextlib
import warnings

import numpy as np
import torch

from . import _lib
from ._lib import FIELDS, check, field_mask, lib, ptr, stream_handle

TENSOR_FIELDS = ("g_ee", "g_nn", "g_zz", "g_en", "g_ez", "g_nz")


# ----------------------------------------------------------------- helpers
def to_device(a, dtype=torch.float64):
    """Copy a host array to a contiguous device tensor."""
    dev = _lib.require_device()
    a = np.ascontiguousarray(a)
    if not a.flags.writeable:
        a = a.copy()
    t = torch.as_tensor(a)
    return t.to(device=dev, dtype=dtype).contiguous()


def _workspace(n_obs, n_src, mask):
    lib()
    nbytes = ctypes_size(n_obs, n_src, mask)
    dev = _lib.require_device()
    return torch.empty(max(nbytes, 8), dtype=torch.uint8, device=dev), nbytes


def ctypes_size(n_obs, n_src, mask):
    import ctypes

    out = ctypes.c_size_t(0)
    check(lib().hmx_forward_workspace_bytes(int(n_obs), int(n_src), int(mask), ctypes.byref(out)))
    return out.value


def _prepare_coordinates(coordinates):
    """Broadcast and ravel the first three coordinates (float64)."""
    cast = np.broadcast(*coordinates[:3])
    arrays = tuple(
        np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64), cast.shape).ravel())
        for c in coordinates[:3]
    )
    return cast, arrays


def progressbar_context(total, use_progressbar):
    """Same contract as initialize_progressbar (_forward/utils.py:333-392):
    a null context when not requested, ImportError when numba_progress is
    missing.  The GPU kernels run as one launch, so no per-point progress is
    reported."""
    if not use_progressbar:
        return contextlib.nullcontext()
    try:
        import numba_progress  # noqa: F401, PLC0415
    except ModuleNotFoundError as original:
        error = ImportError(
            "Cannot import the optional dependency 'numba_progress'. "
            "It must be installed to be able to show a progressbar."
        )
        raise error from original
    return contextlib.nullcontext()


# -------------------------------------------------------- device-level API
def on_stream(stream):
    """Context in which the output and workspace of a device-level call are
    allocated: the caller's stream when one is given (so torch's caching
    allocator orders their reuse after the kernel on that stream), else the
    current stream."""
    return torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()


def prism_fields_device(easting, northing, upward, prisms, density, fields, stream=None,
                        flags=None):
    """Fields of prisms on observation points, all tensors already in HBM.

    easting/northing/upward: (n,) float64; prisms: (m, 6) float64 row-major
    [W, E, S, N, B, T]; density: (m,) float64.  Returns a (len(fields), n)
    float64 tensor in harmonica's units.  ``flags`` (int32 device tensor of
    one element) receives OR-ed singular-point bits per field.
    """
    mask = field_mask(fields)
    n_obs, n_prisms = easting.numel(), prisms.shape[0]
    with on_stream(stream):
        out = torch.empty((bin(mask).count("1"), n_obs), dtype=torch.float64,
                          device=easting.device)
        ws, nbytes = _workspace(n_obs, n_prisms, mask)
        check(lib().hmx_prism_gravity(ptr(easting), ptr(northing), ptr(upward), n_obs,
                                      ptr(prisms), ptr(density), n_prisms, mask, ptr(out),
                                      ptr(flags), ptr(ws), nbytes, stream_handle(stream)))
    return out


def layer_fields_device(easting, northing, upward, centre_easting, centre_northing, bottom, top,
                        density, fields, thickness_threshold=0.0, stream=None, flags=None):
    """Fields of a prism layer (centres + (n_north, n_east) grids in HBM)."""
    mask = field_mask(fields)
    n_obs = easting.numel()
    n_e, n_n = centre_easting.numel(), centre_northing.numel()
    with on_stream(stream):
        out = torch.empty((bin(mask).count("1"), n_obs), dtype=torch.float64,
                          device=easting.device)
        ws, nbytes = _workspace(n_obs, n_e * n_n, mask)
        check(lib().hmx_prism_layer_gravity(ptr(easting), ptr(northing), ptr(upward), n_obs,
                                            ptr(centre_easting), n_e, ptr(centre_northing), n_n,
                                            ptr(bottom), ptr(top), ptr(density),
                                            float(thickness_threshold), mask, ptr(out),
                                            ptr(flags), ptr(ws), nbytes, stream_handle(stream)))
    return out


def point_fields_device(c0, c1, c2, p0, p1, p2, masses, fields, coordinate_system=0, stream=None):
    """Fields of point masses (all tensors in HBM)."""
    mask = field_mask(fields)
    n_obs, n_pts = c0.numel(), p0.numel()
    with on_stream(stream):
        out = torch.empty((bin(mask).count("1"), n_obs), dtype=torch.float64, device=c0.device)
        ws, nbytes = _workspace(n_obs, n_pts, mask)
        check(lib().hmx_point_gravity(ptr(c0), ptr(c1), ptr(c2), n_obs, ptr(p0), ptr(p1),
                                      ptr(p2), ptr(masses), n_pts, int(coordinate_system), mask,
                                      ptr(out), ptr(ws), nbytes, stream_handle(stream)))
    return out


# ---------------------------------------------------------- prism_gravity
def check_prisms(prisms):
    """Check that prism boundaries are well defined (prisms/utils.py:12-43)."""
    west, east, south, north, bottom, top = tuple(prisms[:, i] for i in range(6))
    err_msg = "Invalid prism or prisms. "
    bad_we = west > east
    bad_sn = south > north
    bad_bt = bottom > top
    if bad_we.any():
        err_msg += "The west boundary can't be greater than the east one.\n"
        for prism in prisms[bad_we]:
            err_msg += f"\tInvalid prism: {prism}\n"
        raise ValueError(err_msg)
    if bad_sn.any():
        err_msg += "The south boundary can't be greater than the north one.\n"
        for prism in prisms[bad_sn]:
            err_msg += f"\tInvalid prism: {prism}\n"
        raise ValueError(err_msg)
    if bad_bt.any():
        err_msg += "The bottom boundary can't be greater than the top one.\n"
        for prism in prisms[bad_bt]:
            # The reference's message says "tesseroid" here (utils.py:42).
            err_msg += f"\tInvalid tesseroid: {prism}\n"
        raise ValueError(err_msg)


def discard_null_prisms(prisms, density):
    """Drop zero-volume and zero-density prisms (gravity.py:452-486)."""
    west, east, south, north, bottom, top = tuple(prisms[:, i] for i in range(6))
    null_prisms = (west == east) | (south == north) | (bottom == top)
    null_prisms[density == 0] = True
    if not null_prisms.any():  # nothing to drop: no copy of the arrays
        return prisms, density
    keep = np.logical_not(null_prisms)
    return prisms[keep, :], density[keep]


def _warn_singular():
    warnings.warn(
        "Found observation point on singular point of a prism.",
        UserWarning,
        stacklevel=3,
    )


def prism_gravity(
    coordinates,
    prisms,
    density,
    field,
    parallel=True,
    dtype="float64",
    progressbar=False,
    disable_checks=False,
):
    """
    Gravitational fields of right-rectangular prisms in Cartesian coordinates.

    Same signature, units and conventions as the reference
    (src/harmonica/_forward/prisms/gravity.py:51-236): potential in J/kg,
    accelerations in mGal, tensor components in Eotvos; ``g_z``, ``g_ez`` and
    ``g_nz`` point downward; tensor components are NaN on singular points
    (prism vertices and the edges listed in the reference docstring) and a
    UserWarning is issued when ``disable_checks`` is False.  The sum over
    prisms runs on the GPU in FP64 (``parallel`` is accepted for
    compatibility; the GPU path is always parallel).  With a ``dtype`` other
    than float64 the FP64 result is cast at the end.

    Examples
    --------
    >>> prism = [-34, 5, -18, 14, -345, -146]
    >>> coordinates = ([-40, 0, 40], [0, 0, 0], [30, 30, 30])
    >>> gz = prism_gravity(coordinates, prism, 2670, field="g_z")  # doctest: +SKIP
    >>> print("({:.5f}, {:.5f}, {:.5f})".format(*gz))  # doctest: +SKIP
    (0.06552, 0.06629, 0.06174)
    """
    if field not in FIELDS:
        msg = f"Gravitational field {field} not recognized"
        raise ValueError(msg)
    cast, coords = _prepare_coordinates(coordinates)
    prisms = np.atleast_2d(np.asarray(prisms, dtype=np.float64))
    density = np.atleast_1d(np.asarray(density, dtype=np.float64)).ravel()
    if prisms.ndim != 2 or prisms.shape[1] != 6:
        raise ValueError(f"Invalid prisms array with shape {prisms.shape}: expected (n, 6).")
    singular_check = False
    if not disable_checks:
        if density.size != prisms.shape[0]:
            raise ValueError(
                f"Number of elements in density ({density.size}) "
                + f"mismatch the number of prisms ({prisms.shape[0]})"
            )
        check_prisms(prisms)
        singular_check = field in TENSOR_FIELDS
    all_prisms = prisms
    prisms, density = discard_null_prisms(prisms, density)
    with progressbar_context(coords[0].size, progressbar):
        if cast.size == 0:
            return np.zeros(cast.shape, dtype=dtype)
        e, n, u = (to_device(c) for c in coords)
        out = prism_fields_device(e, n, u, to_device(prisms.reshape(-1, 6)), to_device(density),
                                  (field,))
        if singular_check:
            # The reference scans every prism, null ones included, before
            # discarding (gravity.py:207-215).
            if prism_any_singular_device(e, n, u, to_device(all_prisms), field):
                _warn_singular()
        result = out[0].cpu().numpy()
    return result.astype(dtype, copy=False).reshape(cast.shape)


def prism_any_singular_device(easting, northing, upward, prisms, field):
    """True if any point is a singular point of ``field`` for any prism
    (replaces the serial scans of gravity.py:272-449)."""
    flags = torch.zeros(1, dtype=torch.int32, device=easting.device)
    bit = _lib.FIELD_BIT[field]
    check(lib().hmx_prism_singular_scan(ptr(easting), ptr(northing), ptr(upward), easting.numel(),
                                        ptr(prisms), prisms.shape[0], bit, ptr(flags),
                                        stream_handle(None)))
    return bool(int(flags.item()) & bit)


# ---------------------------------------------------------- point_gravity
_POINT_FIELDS = {
    "cartesian": ("potential", "g_e", "g_n", "g_z", "g_ee", "g_nn", "g_zz", "g_en", "g_ez",
                  "g_nz", "g_ne", "g_ze", "g_zn"),
    "spherical": ("potential", "g_z", "g_n", "g_e"),
}
_ALIASES = {"g_ne": "g_en", "g_ze": "g_ez", "g_zn": "g_nz"}


def check_coordinate_system(coordinate_system,
                            valid_coord_systems=("cartesian", "spherical", "geodetic")):
    """_forward/utils.py:71-87."""
    if coordinate_system not in valid_coord_systems:
        msg = f"Coordinate system {coordinate_system} not recognized."
        raise ValueError(msg)


def point_gravity(
    coordinates,
    points,
    masses,
    field,
    coordinate_system="cartesian",
    parallel=True,
    dtype="float64",
):
    """
    Gravitational fields of point masses (reference: point.py:30-261).

    Cartesian coordinates support all ten fields plus the aliases ``g_ne``,
    ``g_ze`` and ``g_zn``; spherical coordinates (longitude and latitude in
    degrees, radius in meters) support ``potential`` and ``g_z`` and raise
    NotImplementedError for ``g_e`` and ``g_n``.  Units and signs as in
    prism_gravity.
    """
    check_coordinate_system(coordinate_system, valid_coord_systems=("cartesian", "spherical"))
    cast, coords = _prepare_coordinates(coordinates)
    points = tuple(np.atleast_1d(np.asarray(p, dtype=np.float64)).ravel() for p in points[:3])
    masses = np.atleast_1d(np.asarray(masses, dtype=np.float64)).ravel()
    if masses.size != points[0].size:
        raise ValueError(
            f"Number of elements in masses ({masses.size}) "
            + f"mismatch the number of points ({points[0].size})"
        )
    if field not in _POINT_FIELDS[coordinate_system]:
        msg = f"Gravitational field '{field}' not recognized"
        raise ValueError(msg)
    if coordinate_system == "spherical" and field in ("g_e", "g_n"):
        raise NotImplementedError
    if cast.size == 0:
        return np.zeros(cast.shape, dtype=dtype)
    system = _lib.HMX_CARTESIAN if coordinate_system == "cartesian" else _lib.HMX_SPHERICAL
    dev_coords = [to_device(c) for c in coords]
    dev_points = [to_device(p) for p in points]
    out = point_fields_device(*dev_coords, *dev_points, to_device(masses),
                              (_ALIASES.get(field, field),), coordinate_system=system)
    return out[0].cpu().numpy().astype(dtype, copy=False).reshape(cast.shape)
