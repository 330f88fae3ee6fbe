"""
Prism layers: ``prism_layer`` and the ``.prism_layer`` accessor
(reference: src/harmonica/_forward/prisms/layer.py:23-519).

xarray is not part of this stack, so ``prism_layer`` returns a
``PrismLayerDataset``: a small dataset object with the attributes the
reference's Dataset exposes to its users and tests (``easting``,
``northing``, ``top``, ``bottom`` coordinates with ``.values``, ``coords``,
``data_vars``, item and attribute access to properties, ``attrs``) and the
``prism_layer`` accessor (``DatasetAccessorPrismLayer``) with the same
properties and methods, whose ``gravity`` runs the HIP layer kernel.
"""

import warnings

import numpy as np

from . import _lib
from ._forward import FIELDS, layer_fields_device, progressbar_context, to_device
from ._grid_cells import all_cells, one_cell, top_and_bottom


class DataArrayLite(np.lib.mixins.NDArrayOperatorsMixin):
    """A named numpy array with ``.values`` and ``.dims`` (enough of
    xarray.DataArray for the prism-layer API)."""

    def __init__(self, values, dims):
        self.values = np.asarray(values)
        self.dims = tuple(dims)

    def __array__(self, dtype=None, copy=None):
        return self.values if dtype is None else self.values.astype(dtype)

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        args = [i.values if isinstance(i, DataArrayLite) else i for i in inputs]
        return getattr(ufunc, method)(*args, **kwargs)

    def __getitem__(self, key):
        return self.values[key]

    def __len__(self):
        return len(self.values)

    @property
    def shape(self):
        return self.values.shape

    @property
    def size(self):
        return self.values.size

    @property
    def ndim(self):
        return self.values.ndim

    def ravel(self):
        return self.values.ravel()

    def __repr__(self):
        return f"DataArrayLite(dims={self.dims}, values={self.values!r})"


class PrismLayerDataset:
    """Dataset-like container of a prism layer: coordinates ``easting``
    (n_east,), ``northing`` (n_north,), ``top`` and ``bottom`` (n_north,
    n_east), and physical properties as data variables."""

    def __init__(self, easting, northing, data=None, attrs=None):
        self.coords = {
            "easting": DataArrayLite(easting, ("easting",)),
            "northing": DataArrayLite(northing, ("northing",)),
        }
        self.data_vars = dict(data or {})
        self.attrs = dict(attrs or {})

    def _set_coord(self, name, dims, values):
        self.coords[name] = DataArrayLite(values, dims)

    def __getitem__(self, name):
        if name in self.data_vars:
            return self.data_vars[name]
        if name in self.coords:
            return self.coords[name]
        raise KeyError(name)

    def __getattr__(self, name):
        if name.startswith("_") or name in ("coords", "data_vars", "attrs"):
            raise AttributeError(name)
        try:
            return self[name]
        except KeyError:
            raise AttributeError(name) from None

    def __contains__(self, name):
        return name in self.coords or name in self.data_vars

    @property
    def prism_layer(self):
        return DatasetAccessorPrismLayer(self)

    def __repr__(self):
        dims = f"(northing: {self.northing.size}, easting: {self.easting.size})"
        return (f"<PrismLayerDataset {dims}> coords={list(self.coords)} "
                f"data_vars={list(self.data_vars)} attrs={self.attrs}")


def _centres(coordinates):
    """Centre vectors from 1d arrays or meshgrid 2d arrays (what
    verde.make_xarray_grid accepts)."""
    easting, northing = (np.asarray(c, dtype=float) for c in coordinates[:2])
    if easting.ndim == 2:
        easting = easting[0, :]
    if northing.ndim == 2:
        northing = northing[:, 0]
    return easting, northing


def prism_layer(coordinates, surface, reference, properties=None):
    """
    Create a layer of prisms of equal size (reference: layer.py:23-138).

    Parameters and semantics as in the reference: ``coordinates`` are the
    prism centres (1d arrays or a meshgrid), ``surface`` and ``reference``
    set top and bottom (swapped where the surface lies below the reference),
    ``properties`` maps names to (n_north, n_east) arrays.

    >>> import numpy as np
    >>> layer = prism_layer((np.linspace(0, 10, 5), np.linspace(2, 8, 4)),
    ...                     np.arange(20.0).reshape(4, 5), reference=0,
    ...                     properties={"density": 2670.0 * np.ones((4, 5))})
    >>> [float(b) for b in layer.prism_layer.boundaries]
    [-1.25, 11.25, 1.0, 9.0]
    >>> [float(b) for b in layer.prism_layer.get_prism((0, 2))]
    [3.75, 6.25, 1.0, 3.0, 0.0, 2.0]
    """
    easting, northing = _centres(coordinates)
    data = None
    if properties:
        data = {}
        for name, values in properties.items():
            values = np.asarray(values)
            if values.shape != (northing.size, easting.size):
                raise ValueError(
                    f"Invalid property '{name}' with shape {values.shape}: expected "
                    f"{(northing.size, easting.size)}."
                )
            data[name] = DataArrayLite(values, ("northing", "easting"))
    layer = PrismLayerDataset(easting, northing, data=data)
    _check_regular_grid(layer.easting.values, layer.northing.values)
    layer.attrs = {"coords_units": "meters", "properties_units": "SI"}
    layer.prism_layer.update_top_bottom(surface, reference)
    return layer


def _check_regular_grid(easting, northing):
    """layer.py:141-154."""
    if not np.allclose(easting[1] - easting[0], easting[1:] - easting[:-1]):
        msg = "Passed easting coordinates are not evenly spaced."
        raise ValueError(msg)
    if not np.allclose(northing[1] - northing[0], northing[1:] - northing[:-1]):
        msg = "Passed northing coordinates are not evenly spaced."
        raise ValueError(msg)


class DatasetAccessorPrismLayer:
    """Accessor of a prism layer (reference: layer.py:157-519)."""

    def __init__(self, obj):
        self._obj = obj

    @property
    def dims(self):
        return ("northing", "easting")

    @property
    def spacing(self):
        easting, northing = self._obj.easting.values, self._obj.northing.values
        _check_regular_grid(easting, northing)
        s_north, s_east = northing[1] - northing[0], easting[1] - easting[0]
        return s_north, s_east

    @property
    def boundaries(self):
        s_north, s_east = self.spacing
        west = self._obj.easting.values.min() - s_east / 2
        east = self._obj.easting.values.max() + s_east / 2
        south = self._obj.northing.values.min() - s_north / 2
        north = self._obj.northing.values.max() + s_north / 2
        return west, east, south, north

    @property
    def size(self):
        return self._obj.northing.size * self._obj.easting.size

    @property
    def shape(self):
        return (self._obj.northing.size, self._obj.easting.size)

    def update_top_bottom(self, surface, reference):
        """layer.py:264-312."""
        top, bottom = top_and_bottom(surface, reference, self.shape, "prisms")
        self._obj._set_coord("top", self.dims, top)
        self._obj._set_coord("bottom", self.dims, bottom)

    def gravity(
        self,
        coordinates,
        field,
        *,
        density_name="density",
        thickness_threshold=None,
        parallel=True,
        progressbar=False,
    ):
        """
        Gravity field of the layer on observation points (layer.py:314-433),
        computed by the HIP layer kernel in FP64.  Prisms with NaN top or
        bottom, zero or NaN density or thickness below ``thickness_threshold``
        are ignored; NaN densities raise a warning.  Units and signs as in
        prism_gravity; the output is always float64.
        """
        if field not in FIELDS:
            msg = f"Gravitational field '{field}' not recognized."
            raise ValueError(msg)
        _check_regular_grid(self._obj.easting.values, self._obj.northing.values)
        cast = np.broadcast(*coordinates[:3])
        coords = tuple(
            np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64), cast.shape).ravel())
            for c in coordinates[:3]
        )
        thickness_threshold = 0.0 if thickness_threshold is None else thickness_threshold
        density = np.asarray(self._obj[density_name].values, dtype=np.float64)
        if np.isnan(density).any():
            msg = (
                "Found NaN values in 'density' property of the prisms layer. "
                "Their respective prisms will be ignored."
            )
            warnings.warn(msg, stacklevel=2)
        with progressbar_context(coords[0].size, progressbar):
            if cast.size == 0:
                return np.zeros(cast.shape, dtype=np.float64)
            e, n, u = (to_device(c) for c in coords)
            out = layer_fields_device(
                e, n, u,
                to_device(self._obj.easting.values), to_device(self._obj.northing.values),
                to_device(self._obj.bottom.values), to_device(self._obj.top.values),
                to_device(density), (field,), thickness_threshold=thickness_threshold,
            )
            result = out[0].cpu().numpy()
        return result.reshape(cast.shape)

    def _to_prisms(self):
        s_north, s_east = self.spacing
        return all_cells(self._obj.easting.values, self._obj.northing.values, s_east, s_north,
                         self._obj.bottom.values, self._obj.top.values)

    def get_prism(self, indices):
        s_north, s_east = self.spacing
        return one_cell(self._obj.easting.values, self._obj.northing.values, s_east, s_north,
                        self._obj.bottom.values, self._obj.top.values, tuple(indices))

    def to_pyvista(self, drop_null_prisms=True):
        """Export to pyvista (layer.py:485-519); needs pyvista."""
        try:
            import pyvista  # noqa: F401, PLC0415
        except ModuleNotFoundError as err:
            raise ImportError("to_pyvista requires pyvista, which is not installed") from err
        raise NotImplementedError("pyvista export is outside the GPU forward-modelling path")


def layer_gravity_fields(layer, coordinates, fields, density_name="density",
                         thickness_threshold=None):
    """Several fields of a layer in one fused launch (device arrays out):
    returns a (len(fields), n) float64 torch tensor on the GPU."""
    _lib.require_device()
    cast, coords = np.broadcast(*coordinates[:3]), coordinates
    e, n, u = (to_device(np.broadcast_to(np.asarray(c, float), cast.shape).ravel()) for c in coords[:3])
    return layer_fields_device(
        e, n, u, to_device(layer.easting.values), to_device(layer.northing.values),
        to_device(layer.bottom.values), to_device(layer.top.values),
        to_device(np.asarray(layer[density_name].values, float)), tuple(fields),
        thickness_threshold=0.0 if thickness_threshold is None else thickness_threshold,
    )
