"""
Drop-in ``prism_magnetic``: the magnetic field of uniformly magnetised
right-rectangular prisms, with the reference's signature, checks, messages
and units, computed by the HIP kernel of libhmx (``hmx_prism_magnetic``).

Reference: src/harmonica/_forward/prisms/magnetic.py:20-466.
"""

import ctypes

import numpy as np
import torch

from ._forward import (_prepare_coordinates, check_prisms, on_stream, progressbar_context,
                       to_device)
from ._lib import check, lib, ptr, stream_handle

VALID_FIELDS = ("b", "b_e", "b_n", "b_u")
#: Component bits of hmx_prism_magnetic (HMX_B_E, HMX_B_N, HMX_B_U).
COMPONENT_BIT = {"b_e": 1, "b_n": 2, "b_u": 4}


def magnetic_fields_device(easting, northing, upward, prisms, mag_e, mag_n, mag_u, components,
                           stream=None):
    """(len(components), n_obs) device tensor of the listed components (nT)
    for device inputs (float64, contiguous)."""
    mask = 0
    for c in components:
        mask |= COMPONENT_BIT[c]
    n_obs, n_prisms = easting.numel(), prisms.shape[0]
    nbytes = ctypes.c_size_t(0)
    check(lib().hmx_prism_magnetic_workspace_bytes(n_obs, n_prisms, mask, ctypes.byref(nbytes)))
    nc = bin(mask).count("1")
    with on_stream(stream):
        ws = torch.empty(max(nbytes.value, 8), dtype=torch.uint8, device=easting.device)
        out = torch.empty((nc, n_obs), dtype=torch.float64, device=easting.device)
        check(lib().hmx_prism_magnetic(ptr(easting), ptr(northing), ptr(upward), n_obs,
                                       ptr(prisms), ptr(mag_e), ptr(mag_n), ptr(mag_u), n_prisms,
                                       mask, ptr(out), ptr(ws), nbytes.value,
                                       stream_handle(stream)))
    # Rows come in bit order (b_e, b_n, b_u); reorder to the request.
    order = sorted(components, key=lambda c: COMPONENT_BIT[c])
    if list(order) == list(components):
        return out
    return out[[order.index(c) for c in components]]


def _run_sanity_checks(prisms, magnetization):
    """magnetic.py:443-457."""
    if (size := len(magnetization)) != 3:
        raise ValueError(
            f"Invalid magnetization vectors with '{size}' elements. "
            + "Magnetization vectors should have only 3 elements."
        )
    if magnetization[0].size != prisms.shape[0]:
        raise ValueError(
            f"Number of magnetization vectors ({magnetization[0].size}) "
            + f"mismatch the number of prisms ({prisms.shape[0]})"
        )
    check_prisms(prisms)


def _discard_null_prisms(prisms, magnetization):
    """Drop zero-volume prisms and prisms whose three magnetisation
    components are all zero (magnetic.py:403-440)."""
    west, east, south, north, bottom, top = tuple(prisms[:, i] for i in range(6))
    null_prisms = (west == east) | (south == north) | (bottom == top)
    mag_e, mag_n, mag_u = magnetization
    null_mag = (mag_e == 0) & (mag_n == 0) & (mag_u == 0)
    null_prisms[null_mag] = True
    if not null_prisms.any():  # nothing to drop: no copy of the arrays
        return prisms, tuple(magnetization)
    keep = np.logical_not(null_prisms)
    return prisms[keep, :], tuple(m[keep] for m in magnetization)


def prism_magnetic(
    coordinates,
    prisms,
    magnetization,
    field,
    parallel=True,
    dtype=np.float64,
    progressbar=False,
    disable_checks=False,
):
    """
    Magnetic field of right-rectangular prisms in Cartesian coordinates.

    Same signature, units and conventions as the reference
    (src/harmonica/_forward/prisms/magnetic.py:28-238): ``magnetization`` is
    a tuple of the east, north and up components of each prism's
    magnetisation in A/m; the result is in nT, one array for ``b_e``,
    ``b_n`` or ``b_u`` and the tuple ``(b_e, b_n, b_u)`` for ``b``.  Values
    on an edge or vertex of a prism are NaN.  The sum over prisms runs on
    the GPU in FP64 (``parallel`` is accepted for compatibility); with a
    ``dtype`` other than float64 the FP64 result is cast at the end.
    """
    if field not in VALID_FIELDS:
        msg = (
            f"Invalid field '{field}'. Please choose one of '{','.join(VALID_FIELDS)}'."
        )
        raise ValueError(msg)
    cast, coords = _prepare_coordinates(coordinates)
    prisms = np.atleast_2d(np.asarray(prisms, dtype=np.float64))
    magnetization = tuple(np.atleast_1d(np.asarray(m, dtype=np.float64)).ravel()
                          for m in magnetization)
    if not disable_checks:
        _run_sanity_checks(prisms, magnetization)
    if prisms.ndim != 2 or prisms.shape[1] != 6:
        raise ValueError(f"Invalid prisms array with shape {prisms.shape}: expected (n, 6).")
    prisms, magnetization = _discard_null_prisms(prisms, magnetization)
    components = ("b_e", "b_n", "b_u") if field == "b" else (field,)
    with progressbar_context(coords[0].size, progressbar):
        if cast.size == 0:
            result = np.zeros((len(components), 0))
        else:
            e, n, u = (to_device(c) for c in coords)
            me, mn, mu = (to_device(m) for m in magnetization)
            out = magnetic_fields_device(e, n, u, to_device(prisms.reshape(-1, 6)), me, mn, mu,
                                         components)
            result = out.cpu().numpy()
    arrays = tuple(r.astype(dtype, copy=False).reshape(cast.shape) for r in result)
    return arrays if field == "b" else arrays[0]

