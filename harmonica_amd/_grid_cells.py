"""
Cells of a regular layer (prism or tesseroid layers): shared helpers of
the layer accessors.

The accessors' public behaviour follows the reference
(src/harmonica/_forward/prisms/layer.py:264-312, 434-483 and
src/harmonica/_forward/tesseroids/tesseroid_layer.py:406-483): a cell's
horizontal bounds are its centre -+ half the first spacing, cells are
enumerated northing-major (meshgrid ravel), the top is the larger of the
surface and the reference (the surface where it is not below), and cells
with NaN top or bottom (or property) are the missing ones.
"""

import warnings

import numpy as np


def cell_bounds(cx, cy, spacing_x, spacing_y):
    """(west, east, south, north) of cells centred on (cx, cy)."""
    hx, hy = spacing_x / 2, spacing_y / 2
    return cx - hx, cx + hx, cy - hy, cy + hy


def all_cells(x, y, spacing_x, spacing_y, bottom, top):
    """(n_y * n_x, 6) [west, east, south, north, bottom, top] of every cell,
    northing-major."""
    cells = np.empty((y.size * x.size, 6))
    bounds = cell_bounds(np.tile(x, y.size), np.repeat(y, x.size), spacing_x, spacing_y)
    for k, column in enumerate(bounds):
        cells[:, k] = column
    cells[:, 4] = np.ravel(bottom)
    cells[:, 5] = np.ravel(top)
    return cells


def one_cell(x, y, spacing_x, spacing_y, bottom, top, indices):
    """(west, east, south, north, bottom, top) of the cell (i_y, i_x)."""
    i_y, i_x = indices
    return (*cell_bounds(x[i_x], y[i_y], spacing_x, spacing_y), bottom[indices], top[indices])


def top_and_bottom(surface, reference, shape, body):
    """Top and bottom of a layer between ``surface`` and ``reference``
    (scalar or array of the layer's shape), with the reference's messages
    (``body``: "prisms" or "tesseroids")."""
    surface, reference = np.asarray(surface), np.asarray(reference)
    for name, values in (("surface", surface), ("reference", reference)):
        if (name == "surface" or values.ndim != 0) and values.shape != shape:
            raise ValueError(
                f"Invalid {name} array with shape '{values.shape}'. "
                + "Its shape should be compatible with the coordinates "
                + f"of the layer of {body}."
            )
    reference = np.broadcast_to(reference, shape)
    below = surface < reference  # False where either is NaN
    return np.where(below, reference, surface), np.where(below, surface, reference)


def present_cells(top, bottom, values=None, name=None, body="tesseroid"):
    """Mask of the cells with finite-or-infinite (not NaN) top and bottom,
    and, with ``values``, property; warns when a present cell misses the
    property."""
    present = ~(np.isnan(top) | np.isnan(bottom))
    if values is None:
        return present
    missing = np.isnan(values)
    if (missing & present).any():
        warnings.warn(
            f"Found missing values in '{name}' property "
            + f"of the {body} layer. The {body}s with nan as "
            + f"'{name}' will be ignored.",
            stacklevel=2,
        )
    return present & ~missing
