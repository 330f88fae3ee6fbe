"""
Equivalent sources in spherical coordinates
(reference: src/harmonica/_equivalent_sources/spherical.py:28-424).

The Green's function is 1 / l with l the distance between two points given
as (longitude, latitude, radius) (_forward/utils.py:163-201).  On the GPU
the points are first turned into geocentric Cartesian coordinates
(``hmx_geocentric``), where l = |X - X'|, and the Cartesian jacobian,
predict and fit kernels run on them unchanged.  The reference evaluates
l^2 = (r - r')^2 + 2 r r' (1 - cos psi); for nearby points 1 - cos psi
loses digits there (about 1e-8 relative in l for sources 500 m below the
data), which the geocentric difference avoids.
"""

import ctypes
import warnings

import numpy as np
import torch
from sklearn.base import BaseEstimator
from sklearn.utils.validation import check_is_fitted

from . import _lib
from ._eqs import (
    GridDataset,
    _dev,
    check_fit_input,
    fit_device,
    get_region,
    n_1d_arrays,
)
from ._lib import check, lib, ptr, stream_handle


def geocentric_device(longitude, latitude, radius):
    """Device tensors (x, y, z) of the points (hmx_geocentric)."""
    lon, lat, r = (_dev(np.asarray(c, dtype=np.float64).ravel(), torch.float64)
                   for c in (longitude, latitude, radius))
    out = [torch.empty_like(lon) for _ in range(3)]
    check(lib().hmx_geocentric(ptr(lon), ptr(lat), ptr(r), lon.numel(), *(ptr(o) for o in out),
                               stream_handle(None)))
    return tuple(out)


def _jacobian_geocentric(xyz, pxyz):
    n, m = xyz[0].numel(), pxyz[0].numel()
    jac = torch.empty((n, m), dtype=torch.float64, device=xyz[0].device)
    check(lib().hmx_eqs_jacobian(*(ptr(x) for x in xyz), n, *(ptr(p) for p in pxyz), m,
                                 _lib.HMX_F64, ptr(jac), m, stream_handle(None)))
    return jac


def _predict_geocentric(xyz, pxyz, coefs):
    n, m = xyz[0].numel(), pxyz[0].numel()
    cf = _dev(np.asarray(coefs, dtype=np.float64), torch.float64)
    out = torch.empty(n, dtype=torch.float64, device=xyz[0].device)
    nbytes = ctypes.c_size_t(0)
    check(lib().hmx_eqs_predict_workspace_bytes(n, m, ctypes.byref(nbytes)))
    ws = torch.empty(max(nbytes.value, 8), dtype=torch.uint8, device=xyz[0].device)
    check(lib().hmx_eqs_predict(*(ptr(x) for x in xyz), n, *(ptr(p) for p in pxyz), ptr(cf), m,
                                _lib.HMX_F64, ptr(out), ptr(ws), nbytes.value,
                                stream_handle(None)))
    return out


class EquivalentSourcesSph(BaseEstimator):
    r"""
    Equivalent sources for generic harmonic functions in spherical
    coordinates (reference: spherical.py:28-124).

    Parameters are those of the reference: ``damping`` (None for no
    regularisation), ``points`` (custom (longitude, latitude, radius)
    sources), ``relative_depth`` (sources this far below each datum,
    default 500 m), ``parallel`` (accepted; the GPU is always parallel),
    plus ``fit_precision`` as in :class:`EquivalentSources`.  Attributes
    ``points_``, ``coefs_``, ``region_``; ``fit_info_`` reports the solver
    statistics.
    """

    dims = ("spherical_latitude", "longitude")
    extra_coords_name = "radius"

    def __init__(self, damping=None, points=None, relative_depth=500, parallel=True,
                 fit_precision="float64"):
        if fit_precision not in ("float64", "mixed", "mixed-bf16", "mixed-fp32"):
            raise ValueError(f"Invalid fit_precision '{fit_precision}': 'float64', 'mixed', "
                             "'mixed-bf16' or 'mixed-fp32'.")
        self.damping = damping
        self.points = points
        self.relative_depth = relative_depth
        self.parallel = parallel
        self.fit_precision = fit_precision

    def estimate_required_memory(self, coordinates):
        """Bytes of the (n_data, n_points) jacobian (spherical.py:126-174)."""
        coordinates = n_1d_arrays(coordinates, 3)
        if self.points is None:
            points = (coordinates[0], coordinates[1], coordinates[2] - self.relative_depth)
        else:
            points = n_1d_arrays(self.points, 3)
        return coordinates[0].size * points[0].size * np.dtype(coordinates[0].dtype).itemsize

    def fit(self, coordinates, data, weights=None):
        """Fit the source coefficients (spherical.py:176-217)."""
        coordinates, data, weights = check_fit_input(coordinates, data, weights)
        self.region_ = get_region(coordinates[:2])
        coordinates = n_1d_arrays(coordinates, 3)
        if self.points is None:
            self.points_ = (coordinates[0], coordinates[1], coordinates[2] - self.relative_depth)
        else:
            self.points_ = n_1d_arrays(self.points, 3)
        xyz = geocentric_device(*coordinates)
        pxyz = geocentric_device(*self.points_)
        self.coefs_, self.fit_info_ = fit_device(
            xyz, np.asarray(data, dtype=np.float64).ravel(),
            None if weights is None else np.asarray(weights, dtype=np.float64).ravel(),
            pxyz, self.damping, precision=self.fit_precision, order_xy=coordinates[:2])
        return self

    def predict(self, coordinates):
        """Evaluate the sources on the given points (spherical.py:219-248)."""
        check_is_fitted(self, ["coefs_"])
        shape = np.broadcast(*coordinates[:3]).shape
        dtype = np.asarray(coordinates[0]).dtype
        coords = tuple(np.atleast_1d(np.broadcast_to(np.asarray(c), shape)).ravel()
                       for c in coordinates[:3])
        out = _predict_geocentric(geocentric_device(*coords), geocentric_device(*self.points_),
                                  self.coefs_)
        return out.cpu().numpy().astype(dtype, copy=False).reshape(shape)

    def jacobian(self, coordinates, points, dtype="float64"):
        """The (n_data, n_points) jacobian 1 / l (spherical.py:250-282)."""
        coords = tuple(np.asarray(c).ravel() for c in coordinates[:3])
        pts = tuple(np.asarray(p).ravel() for p in points[:3])
        jac = _jacobian_geocentric(geocentric_device(*coords), geocentric_device(*pts))
        return jac.cpu().numpy().astype(dtype, copy=False)

    def grid(self, coordinates, dims=None, data_names=None, **kwargs):
        """Interpolate onto a regular grid of (longitude, latitude, radius)
        (spherical.py:284-371)."""
        deprecated_args = ("upward" in kwargs, "shape" in kwargs, "region" in kwargs,
                           "spacing" in kwargs)
        if any(deprecated_args):
            raise ValueError(
                "The 'upward', 'region', 'shape' and 'spacing' arguments have been "
                + "deprecated. "
                + "Please pass the coordinates of the target grid through the "
                + "'coordinates' argument."
            )
        if kwargs:
            args = "'" + "', '".join(list(kwargs.keys())) + "'"
            warnings.warn(
                f"The {args} arguments are being ignored. The 'grid' method "
                + "will not take any keyword arguments in the next Harmonica release",
                FutureWarning,
                stacklevel=1,
            )
        dims = self.dims if dims is None else tuple(dims)
        data_names = ["scalars"] if data_names is None else list(data_names)
        longitude, latitude, radius = (np.asarray(c, dtype=float) for c in coordinates[:3])
        if longitude.ndim == 1 and latitude.ndim == 1:
            longitude, latitude = np.meshgrid(longitude, latitude)
        radius = np.broadcast_to(radius, longitude.shape)
        values = self.predict((longitude, latitude, radius))
        grid = GridDataset(longitude[0, :], latitude[:, 0], radius, {data_names[0]: values},
                           dims, {"metadata": f"Generated by {self!r}"})
        grid.coords[self.extra_coords_name] = grid.coords.pop("upward")
        return grid

    def scatter(self, region=None, size=300, random_state=0, dims=None, data_names=None,
                projection=None, **kwargs):
        """Not implemented (spherical.py:373-390)."""
        raise NotImplementedError

    def profile(self, point1, point2, size, dims=None, data_names=None, projection=None,
                **kwargs):
        """Not implemented (spherical.py:392-410)."""
        raise NotImplementedError
