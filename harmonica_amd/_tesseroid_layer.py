"""
Layer of tesseroids of equal angular size (reference:
src/harmonica/_forward/tesseroid_layer.py:20-560): a dataset-like object
with ``longitude`` / ``latitude`` centre coordinates, ``top`` / ``bottom``
radii and physical properties, and the ``tesseroid_layer`` accessor whose
``gravity`` calls :func:`harmonica_amd.tesseroid_gravity` (the GPU path).
"""

import numpy as np

from ._grid_cells import all_cells, one_cell, present_cells, top_and_bottom
from ._layer import DataArrayLite, _centres
from ._tesseroid import tesseroid_gravity


class TesseroidLayerDataset:
    """Dataset-like container of a tesseroid layer: coordinates
    ``longitude`` (n_lon,), ``latitude`` (n_lat,), ``top`` and ``bottom``
    (n_lat, n_lon), and physical properties as data variables."""

    def __init__(self, longitude, latitude, data=None, attrs=None):
        self.coords = {
            "longitude": DataArrayLite(longitude, ("longitude",)),
            "latitude": DataArrayLite(latitude, ("latitude",)),
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
    def tesseroid_layer(self):
        return DatasetAccessorTesseroidLayer(self)

    def __repr__(self):
        dims = f"(latitude: {self.latitude.size}, longitude: {self.longitude.size})"
        return (f"<TesseroidLayerDataset {dims}> coords={list(self.coords)} "
                f"data_vars={list(self.data_vars)} attrs={self.attrs}")


def _check_regular_grid(longitude, latitude):
    """tesseroid_layer.py:153-167."""
    if not np.allclose(longitude[1] - longitude[0], longitude[1:] - longitude[:-1]):
        raise ValueError("Passed longitude coordinates are not evenly spaced.")
    if not np.allclose(latitude[1] - latitude[0], latitude[1:] - latitude[:-1]):
        raise ValueError("Passed latitude coordinates are not evenly spaced.")


def _check_overlap(longitude):
    """tesseroid_layer.py:170-180."""
    spacing = longitude[1] - longitude[0]
    if longitude.max() - longitude.min() >= 360 - spacing:
        raise ValueError(
            "Found invalid longitude coordinates that would create overlapping "
            "tesseroids around the globe."
        )


def tesseroid_layer(coordinates, surface, reference, properties=None):
    """
    Create a layer of tesseroids of equal size (tesseroid_layer.py:20-150).

    ``coordinates`` are the centres (longitude, latitude in degrees; 1d
    arrays or a meshgrid), ``surface`` and ``reference`` the radii of the
    top and bottom (swapped where the surface lies below the reference),
    ``properties`` maps names to (n_lat, n_lon) arrays.

    >>> import numpy as np
    >>> layer = tesseroid_layer((np.arange(0.5, 4), np.arange(-1.5, 2)),
    ...                         6.4e6 + np.zeros((4, 4)), 6.3e6)
    >>> [float(b) for b in layer.tesseroid_layer.boundaries]
    [0.0, 4.0, -2.0, 2.0]
    """
    longitude, latitude = _centres(coordinates)
    data = None
    if properties:
        data = {}
        for name, values in properties.items():
            values = np.asarray(values)
            if values.shape != (latitude.size, longitude.size):
                raise ValueError(
                    f"Invalid property '{name}' with shape {values.shape}: expected "
                    f"{(latitude.size, longitude.size)}."
                )
            data[name] = DataArrayLite(values, ("latitude", "longitude"))
    layer = TesseroidLayerDataset(longitude, latitude, data=data)
    _check_regular_grid(layer.longitude.values, layer.latitude.values)
    _check_overlap(layer.longitude.values)
    layer.attrs = {"longitude_units": "degrees", "latitude_units": "degrees",
                   "radius_units": "meters", "properties_units": "SI"}
    layer.tesseroid_layer.update_top_bottom(surface, reference)
    return layer


def _discard_thin_tesseroids(tesseroids, density, thickness_threshold):
    """tesseroid_layer.py:523-560."""
    thickness = tesseroids[:, -1] - tesseroids[:, -2]
    keep = np.logical_not(thickness < thickness_threshold)
    return tesseroids[keep, :], density[keep]


class DatasetAccessorTesseroidLayer:
    """The ``tesseroid_layer`` accessor (tesseroid_layer.py:183-520)."""

    def __init__(self, obj):
        self._obj = obj

    @property
    def dims(self):
        return ("latitude", "longitude")

    @property
    def spacing(self):
        latitude, longitude = self._obj.latitude.values, self._obj.longitude.values
        _check_regular_grid(longitude, latitude)
        return latitude[1] - latitude[0], longitude[1] - longitude[0]

    @property
    def size(self):
        return self._obj.latitude.size * self._obj.longitude.size

    @property
    def shape(self):
        return (self._obj.latitude.size, self._obj.longitude.size)

    @property
    def boundaries(self):
        s_lat, s_lon = self.spacing
        return (self._obj.longitude.values.min() - s_lon / 2,
                self._obj.longitude.values.max() + s_lon / 2,
                self._obj.latitude.values.min() - s_lat / 2,
                self._obj.latitude.values.max() + s_lat / 2)

    def update_top_bottom(self, surface, reference):
        """tesseroid_layer.py:272-320."""
        top, bottom = top_and_bottom(surface, reference, self.shape, "tesseroids")
        self._obj._set_coord("top", self.dims, top)
        self._obj._set_coord("bottom", self.dims, bottom)

    def gravity(self, coordinates, field, progressbar=False, density_name="density",
                thickness_threshold=None, **kwargs):
        """Gravitational field of the layer (tesseroid_layer.py:322-404):
        tesseroids with NaN top, bottom or density ignored (warning for the
        density), thinner than ``thickness_threshold`` discarded, then
        :func:`harmonica_amd.tesseroid_gravity` with ``**kwargs``."""
        boundaries = self._to_tesseroids()
        density = np.asarray(self._obj[density_name].values)
        mask = self._get_nonans_mask(property_name=density_name)
        boundaries = boundaries[mask.ravel()]
        density = density[mask]
        if thickness_threshold is not None:
            boundaries, density = _discard_thin_tesseroids(boundaries, density,
                                                           thickness_threshold)
        return tesseroid_gravity(coordinates, tesseroids=boundaries, density=density,
                                 field=field, progressbar=progressbar, **kwargs)

    def _get_nonans_mask(self, property_name=None):
        """tesseroid_layer.py:406-441."""
        values = None if property_name is None else self._obj[property_name].values
        return present_cells(self._obj.top.values, self._obj.bottom.values, values,
                             property_name)

    def _to_tesseroids(self):
        s_lat, s_lon = self.spacing
        return all_cells(self._obj.longitude.values, self._obj.latitude.values, s_lon, s_lat,
                         self._obj.bottom.values, self._obj.top.values)

    def get_tesseroid(self, indices):
        s_lat, s_lon = self.spacing
        return one_cell(self._obj.longitude.values, self._obj.latitude.values, s_lon, s_lat,
                        self._obj.bottom.values, self._obj.top.values, tuple(indices))
