"""
Equivalent sources in Cartesian coordinates
(reference: src/harmonica/_equivalent_sources/cartesian.py:33-644 and
src/harmonica/_equivalent_sources/utils.py:16-103).

The jacobian, predict and fit run on the GPU (hmx_eqs_jacobian,
hmx_eqs_predict, hmx_eqs_fit).  The reference builds on verde's
BaseGridder and least_squares and on bordado's helpers, none of which is in
this stack; the host-side pieces the path needs are restated here:

* fit-input checks, region and n_1d_arrays (verde.base / bordado);
* the default source depth, 4.5 x the mean nearest-neighbour distance
  (bordado.neighbor_distance_statistics with k=1, cartesian.py:314-319);
* block-averaged source locations: median of the points in each block of
  a pixel-registered grid of blocks (verde.BlockReduce(np.median),
  cartesian.py:326-351);
* grid() / profile() outputs: a dataset-like grid with a ``scalars``
  variable and a pandas DataFrame.
"""

import warnings

import numpy as np
import torch
from sklearn.base import BaseEstimator
from sklearn.utils.validation import check_is_fitted

from . import _lib
from ._forward import to_device
from ._layer import DataArrayLite
from ._lib import check, lib, ptr, stream_handle


# ----------------------------------------------------------- host helpers
def n_1d_arrays(arrays, n):
    """First n arrays, raveled to 1d (verde.base.n_1d_arrays)."""
    return tuple(np.atleast_1d(np.asarray(a)).ravel() for a in arrays[:n])


def get_region(coordinates):
    """(west, east, south, north) of the points (bordado.get_region)."""
    easting, northing = (np.asarray(c) for c in coordinates[:2])
    return (float(np.min(easting)), float(np.max(easting)), float(np.min(northing)),
            float(np.max(northing)))


def check_fit_input(coordinates, data, weights):
    """Shape checks of the fit inputs (verde.base.check_fit_input with a
    single data array)."""
    coordinates = tuple(np.asarray(c) for c in coordinates)
    data = np.asarray(data)
    shapes = [c.shape for c in coordinates]
    if any(s != shapes[0] for s in shapes):
        raise ValueError(f"Coordinate arrays must have the same shape. Coordinate shapes: {shapes}")
    if data.shape != shapes[0]:
        raise ValueError(
            "Data arrays must have the same shape as coordinate arrays. "
            f"Given data with shapes '{[data.shape]}' and coordinates with shapes '{shapes}'."
        )
    if weights is not None:
        weights = np.asarray(weights)
        if weights.shape != data.shape:
            raise ValueError(
                "Weights must have the same shape as the data array. "
                f"Given weights with shape '{weights.shape}' and data with shape '{data.shape}'."
            )
    return coordinates, data, weights


def cast_fit_input(coordinates, data, weights, dtype):
    """_equivalent_sources/utils.py:16-39."""
    coordinates = tuple(c.astype(dtype) for c in coordinates)
    data = data.astype(dtype)
    if weights is not None:
        weights = weights.astype(dtype)
    return coordinates, data, weights


def pop_extra_coords(kwargs):
    """_equivalent_sources/utils.py:42-51."""
    if "extra_coords" in kwargs:
        warnings.warn(
            "EQL gridder will ignore extra_coords: {}.".format(kwargs["extra_coords"]),
            stacklevel=2,
        )
        kwargs.pop("extra_coords")


def nearest_neighbour_distances(coordinates):
    """Distance from every point to its nearest other point (2d)."""
    from scipy.spatial import cKDTree

    pts = np.column_stack([np.ravel(coordinates[0]), np.ravel(coordinates[1])])
    dist, _ = cKDTree(pts).query(pts, k=2)
    return dist[:, 1]


def _block_axis(values, lo, hi, spacing):
    """Index of the nearest block centre along one axis: blocks of a
    pixel-registered grid spanning [lo, hi] with the spacing adjusted to fit
    (n = round((hi - lo) / spacing) blocks); ties go to the lower block."""
    n_blocks = max(int(round((hi - lo) / spacing)), 1)
    width = (hi - lo) / n_blocks
    centres = lo + width * (np.arange(n_blocks) + 0.5)
    dist = np.abs(values[:, None] - centres[None, :])
    return np.argmin(dist, axis=1), n_blocks


def block_median(coordinates, block_size):
    """Median of every coordinate over the points of each block, blocks in
    northing-major order, empty blocks dropped (verde.BlockReduce with
    np.median and drop_coords=False)."""
    coordinates = n_1d_arrays(coordinates, len(coordinates))
    if np.ndim(block_size) == 0:
        s_north = s_east = float(block_size)
    else:
        s_north, s_east = (float(v) for v in block_size)
    w, e, s, n = get_region(coordinates)
    ie, ne = _block_axis(coordinates[0], w, e, s_east)
    inn, _ = _block_axis(coordinates[1], s, n, s_north)
    labels = inn * ne + ie
    order = np.unique(labels)
    return tuple(np.array([np.median(c[labels == lab]) for lab in order]) for c in coordinates)


class GridDataset:
    """Dataset-like result of EquivalentSources.grid: coordinates
    ``easting`` and ``northing`` (1d), ``upward`` (2d) and one data variable
    per name (default ``scalars``), each with ``.values``."""

    def __init__(self, easting, northing, upward, data, dims, attrs):
        self.coords = {
            dims[1]: DataArrayLite(easting, (dims[1],)),
            dims[0]: DataArrayLite(northing, (dims[0],)),
            "upward": DataArrayLite(upward, dims),
        }
        self.data_vars = {k: DataArrayLite(v, dims) for k, v in data.items()}
        self.dims = tuple(dims)
        self.attrs = dict(attrs)

    def __getitem__(self, name):
        if name in self.data_vars:
            return self.data_vars[name]
        return self.coords[name]

    def __getattr__(self, name):
        if name.startswith("_") or name in ("coords", "data_vars", "attrs", "dims"):
            raise AttributeError(name)
        try:
            return self[name]
        except KeyError:
            raise AttributeError(name) from None


# ----------------------------------------------------------- device calls
def _dev(a, dtype):
    """Contiguous device tensor of the given dtype (host arrays are copied)."""
    if isinstance(a, torch.Tensor):
        return a.to(device=_lib.require_device(), dtype=dtype).contiguous()
    return to_device(np.asarray(a), dtype=dtype)


def jacobian_device(coords, points, dtype=torch.float64):
    """(n_data, n_src) jacobian on the GPU (tensor)."""
    c = [_dev(x, dtype) for x in coords]
    p = [_dev(x, dtype) for x in points]
    n, m = c[0].numel(), p[0].numel()
    jac = torch.empty((n, m), dtype=dtype, device=c[0].device)
    code = _lib.HMX_F64 if dtype == torch.float64 else _lib.HMX_F32
    check(lib().hmx_eqs_jacobian(*(ptr(x) for x in c), n, *(ptr(x) for x in p), m, code, ptr(jac), m,
                                 stream_handle(None)))
    return jac


def predict_device(coords, points, coefs, dtype=torch.float64):
    """sum_j c_j / |x_i - p_j| on the GPU (tensor of the coordinates' size)."""
    import ctypes

    c = [_dev(x, dtype) for x in coords]
    p = [_dev(x, dtype) for x in points]
    cf = _dev(coefs, dtype)
    n, m = c[0].numel(), p[0].numel()
    out = torch.empty(n, dtype=dtype, device=c[0].device)
    nbytes = ctypes.c_size_t(0)
    check(lib().hmx_eqs_predict_workspace_bytes(n, m, ctypes.byref(nbytes)))
    ws = torch.empty(max(nbytes.value, 8), dtype=torch.uint8, device=c[0].device)
    code = _lib.HMX_F64 if dtype == torch.float64 else _lib.HMX_F32
    check(lib().hmx_eqs_predict(*(ptr(x) for x in c), n, *(ptr(x) for x in p), ptr(cf), m, code,
                                ptr(out), ptr(ws), nbytes.value, stream_handle(None)))
    return out


def spatial_order(easting, northing, bits=16):
    """Permutation that sorts points along a Morton (Z-order) curve of their
    horizontal coordinates, so that runs of consecutive points are compact
    regions."""
    def cells(v):
        v = np.asarray(v, dtype=np.float64)
        span = v.max() - v.min()
        q = (v - v.min()) / (span if span > 0 else 1.0) * (2**bits - 1)
        return q.astype(np.uint64)

    x, y = cells(easting), cells(northing)
    code = np.zeros_like(x)
    for b in range(bits):
        code |= ((x >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
        code |= ((y >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b + 1)
    return np.argsort(code, kind="stable")


def strip_order(easting, northing, ratio=0.25, run=64):
    """Permutation that walks the points in boustrophedon strips along
    easting: strips of northing height h, easting ascending in even strips
    and descending in odd ones.  h = sqrt(ratio * run * area / n) over the
    points' bounding box, so that ``run`` consecutive points cover about
    h x (h / ratio): ``ratio`` < 1 makes the runs narrower in northing than
    in easting."""
    e = np.asarray(easting, dtype=np.float64).ravel()
    n = np.asarray(northing, dtype=np.float64).ravel()
    if e.size < 2:
        return np.arange(e.size)
    span_e, span_n = np.ptp(e), np.ptp(n)
    if span_n == 0.0:
        return np.argsort(e, kind="stable")
    if span_e == 0.0:
        return np.argsort(n, kind="stable")
    h = np.sqrt(ratio * run * span_e * span_n / e.size)
    strip = np.floor((n - n.min()) / h).astype(np.int64)
    return np.lexsort((np.where(strip % 2 == 0, e, -e), strip))


#: Accuracy bar of the mixed-precision fits' predictions (relative to their
#: largest value, against the FP64 fit) and the conjugate-gradient stop rule
#: derived from it: the relative update of the data-space prediction in the
#: last iteration estimates the error left (measured on the 60,000-datum
#: fit: the error on the 512 x 512 grid is about 1.5 times the last
#: update), so stopping at a tenth of the bar keeps a margin of ~6.
MIXED_PREDICTION_BAR = 1e-6
MIXED_RTOL = MIXED_PREDICTION_BAR / 10


def fit_device(coords, data, weights, points, damping, precision="float64", max_iter=None,
               rtol=None, order_xy=None):
    """Coefficients of the scaled, damped, weighted least-squares fit on the
    GPU (host arrays or device tensors, computed in float64).  Returns
    (coefs numpy array, info dict).  ``order_xy``: host horizontal
    coordinates of the data for the spatial order of the split mode (default
    the first two coordinates).

    Every mode refines its solution by conjugate gradients on the FP64
    normal equations, preconditioned by its Cholesky factor: the FP64 mode
    for a few iterations (at cond ~ 1e11 the direct solve alone is off by
    ~1e-6 in the predictions), the mixed modes until the relative update of
    the predictions falls below rtol."""
    import ctypes

    modes = {"float64": _lib.HMX_FIT_FP64, "mixed": _lib.HMX_FIT_INT8_SPLIT,
             "mixed-bf16": _lib.HMX_FIT_BF16_SPLIT, "mixed-fp32": _lib.HMX_FIT_FP32_MFMA}
    mode = modes[precision]
    split = mode in (_lib.HMX_FIT_BF16_SPLIT, _lib.HMX_FIT_INT8_SPLIT)
    if max_iter is None:
        max_iter = 10 if mode == _lib.HMX_FIT_FP64 else 300
    if rtol is None:
        rtol = 1e-12 if mode == _lib.HMX_FIT_FP64 else MIXED_RTOL
    if split:
        # The normal equations sum over the data, so their order is free.  In
        # spatial order each group of 1024 data the split normal matrix
        # scales separately is a compact region, where the columns of far
        # sources are uniformly small: they keep more significant bits, and
        # the matrix preconditions better.
        xy = coords[:2] if order_xy is None else order_xy
        order = spatial_order(*(np.asarray(v.cpu() if isinstance(v, torch.Tensor) else v)
                                for v in xy[:2]))
    c = [_dev(x, torch.float64) for x in coords]
    p = [_dev(x, torch.float64) for x in points]
    d = _dev(data, torch.float64)
    w = None if weights is None else _dev(weights, torch.float64)
    if split:
        idx = torch.as_tensor(order).to(c[0].device)
        c = [x[idx].contiguous() for x in c]
        d = d[idx].contiguous()
        w = None if w is None else w[idx].contiguous()
    n, m = c[0].numel(), p[0].numel()
    nbytes = ctypes.c_size_t(0)
    if damping is None:  # LinearRegression's minimum-norm least squares
        mode |= _lib.HMX_FIT_NO_DAMPING
    check(lib().hmx_eqs_fit_workspace_bytes(n, m, mode, ctypes.byref(nbytes)))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=c[0].device)
    coefs = torch.empty(m, dtype=torch.float64, device=c[0].device)
    info = (ctypes.c_double * len(FIT_INFO_KEYS))()
    check(lib().hmx_eqs_fit(*(ptr(x) for x in c), n, ptr(d), ptr(w), *(ptr(x) for x in p), m,
                            float(0.0 if damping is None else damping), mode, int(max_iter),
                            float(rtol), ptr(coefs), info, ptr(ws), nbytes.value,
                            stream_handle(None)))
    return coefs.cpu().numpy(), dict(zip(FIT_INFO_KEYS, list(info)))


#: Keys of the solver statistics of hmx_eqs_fit (``fit_info_``): conjugate
#: gradient iterations, their last relative update, the diagonal shift and
#: factorisation retries (fits larger than the SVD route), the largest
#: diagonal entry of the normal matrix (when a shift was needed, else 0), whether the SVD route solved the fit
#: (1.0), its Jacobi sweeps and kept singular values, and the seconds spent
#: forming the normal matrix, factorising it, and solving and refining.
FIT_INFO_KEYS = ("iterations", "last_relative_update", "shift", "retries", "max_diagonal",
                 "svd", "svd_sweeps", "svd_rank", "normal_matrix_s", "factorisation_s",
                 "solve_s")


# ----------------------------------------------------------- the estimator
class EquivalentSources(BaseEstimator):
    r"""
    Equivalent sources for generic harmonic functions (gravity, magnetics),
    with the Green's function 1 / |x - x'| (reference: cartesian.py:33-191).

    Parameters are those of the reference: ``damping`` (None for no
    regularisation), ``points`` (custom source coordinates), ``depth``
    (``"default"`` = 4.5 x mean nearest-neighbour distance, or a non-zero
    number), ``block_size`` (block-averaged sources), ``parallel`` (accepted;
    the GPU is always parallel), ``dtype`` (of the jacobian and predictions).
    ``fit_precision`` selects the GPU least-squares mode: ``"float64"``
    (default; FP64 normal matrix and Cholesky, like the reference's LAPACK
    solve), ``"mixed"`` (normal matrix from exact int8 MFMA products of
    8-bit slices of the scaled jacobian, about 2^-31 relative),
    ``"mixed-bf16"`` (the same slices on bf16 MFMA, at half the rate) or
    ``"mixed-fp32"`` (FP32-MFMA normal matrix); in the mixed modes the
    FP64 factor of the low-precision matrix preconditions conjugate
    gradients on the FP64 normal equations.

    Attributes ``points_``, ``coefs_``, ``depth_``, ``region_`` as in the
    reference; ``fit_info_`` reports the solver statistics.
    """

    dims = ("northing", "easting")
    extra_coords_name = "upward"

    def __init__(
        self,
        damping=None,
        points=None,
        depth="default",
        block_size=None,
        parallel=True,
        dtype="float64",
        fit_precision="float64",
    ):
        if isinstance(depth, str) and depth != "default":
            msg = (
                f"Found invalid 'depth' value equal to '{depth}'. "
                "It should be 'default' or a numeric value."
            )
            raise ValueError(msg)
        if depth == 0:
            msg = "Depth value cannot be zero. It should be a non-zero numeric value."
            raise ValueError(msg)
        if fit_precision not in ("float64", "mixed", "mixed-bf16", "mixed-fp32"):
            raise ValueError(f"Invalid fit_precision '{fit_precision}': 'float64', 'mixed', "
                             "'mixed-bf16' or 'mixed-fp32'.")
        self.damping = damping
        self.points = points
        self.depth = depth
        self.block_size = block_size
        self.parallel = parallel
        self.dtype = dtype
        self.fit_precision = fit_precision

    def estimate_required_memory(self, coordinates):
        """Bytes of the (n_data, n_points) jacobian (cartesian.py:193-234)."""
        coordinates = n_1d_arrays(coordinates, 3)
        points = self._build_points(coordinates)
        return coordinates[0].size * points[0].size * np.dtype(self.dtype).itemsize

    def fit(self, coordinates, data, weights=None):
        """Fit the source coefficients (cartesian.py:236-281)."""
        coordinates, data, weights = check_fit_input(coordinates, data, weights)
        coordinates, data, weights = cast_fit_input(coordinates, data, weights, self.dtype)
        self.region_ = get_region(coordinates[:2])
        coordinates = n_1d_arrays(coordinates, 3)
        if self.points is None:
            self.points_ = tuple(p.astype(self.dtype) for p in self._build_points(coordinates))
        else:
            self.depth_ = None
            self.points_ = tuple(p.astype(self.dtype) for p in n_1d_arrays(self.points, 3))
        _lib.require_device()
        self.coefs_, self.fit_info_ = fit_device(
            tuple(c.astype(np.float64) for c in coordinates), data.ravel().astype(np.float64),
            None if weights is None else weights.ravel().astype(np.float64),
            tuple(p.astype(np.float64) for p in self.points_), self.damping,
            precision=self.fit_precision,
        )
        return self

    def _build_points(self, coordinates):
        """Sources beneath the (block-averaged) points (cartesian.py:283-324)."""
        if self.block_size is not None:
            coordinates = self._block_average_coordinates(coordinates)
        if self.depth == "default":
            self.depth_ = 4.5 * np.mean(nearest_neighbour_distances(coordinates[:2]))
        else:
            self.depth_ = self.depth
        return (coordinates[0], coordinates[1], coordinates[2] - self.depth_)

    def _block_average_coordinates(self, coordinates):
        """Median location per block (cartesian.py:326-351)."""
        return block_median(coordinates, self.block_size)

    def predict(self, coordinates):
        """Evaluate the sources on the given points (cartesian.py:353-383)."""
        check_is_fitted(self, ["coefs_"])
        shape = np.broadcast(*coordinates[:3]).shape
        coords = tuple(
            np.atleast_1d(np.broadcast_to(np.asarray(c).astype(self.dtype), shape)).ravel()
            for c in coordinates[:3]
        )
        tdtype = torch.float64 if np.dtype(self.dtype) == np.float64 else torch.float32
        if np.dtype(self.dtype) not in (np.float64, np.float32):
            raise ValueError(f"Unsupported dtype {self.dtype}")
        out = predict_device(coords, self.points_, np.asarray(self.coefs_), dtype=tdtype)
        return out.cpu().numpy().astype(self.dtype, copy=False).reshape(shape)

    def jacobian(self, coordinates, points):
        """The (n_data, n_points) jacobian 1 / |x_i - p_j| (cartesian.py:385-415)."""
        tdtype = torch.float64 if np.dtype(self.dtype) == np.float64 else torch.float32
        coords = tuple(np.asarray(c).astype(self.dtype).ravel() for c in coordinates[:3])
        pts = tuple(np.asarray(p).astype(self.dtype).ravel() for p in points[:3])
        return jacobian_device(coords, pts, dtype=tdtype).cpu().numpy()

    def grid(self, coordinates, dims=None, data_names=None, projection=None, **kwargs):
        """Interpolate onto a regular grid (cartesian.py:417-515)."""
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
        easting, northing, upward = (np.asarray(c, dtype=float) for c in coordinates[:3])
        if easting.ndim == 1 and northing.ndim == 1:
            easting, northing = np.meshgrid(easting, northing)
        upward = np.broadcast_to(upward, easting.shape)
        pe, pn = (easting, northing) if projection is None else projection(easting, northing)
        values = self.predict((pe, pn, upward))
        return GridDataset(easting[0, :], northing[:, 0], upward, {data_names[0]: values}, dims,
                           {"metadata": f"Generated by {self!r}"})

    def scatter(self, region=None, size=300, random_state=0, dims=None, data_names=None,
                projection=None, **kwargs):
        """Not implemented (as in the reference, cartesian.py:517-534)."""
        raise NotImplementedError

    def profile(self, point1, point2, upward, size, dims=None, data_names=None, projection=None,
                **kwargs):
        """Interpolate along a straight profile (cartesian.py:536-631); returns
        a pandas DataFrame with northing, easting, upward, distance and the
        data column(s)."""
        import pandas as pd

        pop_extra_coords(kwargs)
        dims = self.dims if dims is None else tuple(dims)
        data_names = ["scalars"] if data_names is None else list(data_names)
        p1, p2 = tuple(point1), tuple(point2)
        if projection is not None:
            p1 = tuple(projection(*p1))
            p2 = tuple(projection(*p2))
        easting = np.linspace(p1[0], p2[0], size)
        northing = np.linspace(p1[1], p2[1], size)
        distances = np.hypot(easting - easting[0], northing - northing[0])
        up = np.full(size, float(upward))
        values = self.predict((easting, northing, up))
        if projection is not None:
            easting, northing = projection(easting, northing, inverse=True)
        return pd.DataFrame({dims[0]: northing, dims[1]: easting, "upward": up,
                             "distance": distances, data_names[0]: values})
