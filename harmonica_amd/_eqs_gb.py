"""
Gradient-boosted equivalent sources in Cartesian coordinates
(reference: src/harmonica/_equivalent_sources/gradient_boosted.py:24-397).

The window loop (fit each window's sources to the current residue of its
data, subtract their prediction from every datum, accumulate the
coefficients) runs on the GPU in one call, ``hmx_eqs_gb_fit``.  The window
geometry is host bookkeeping, restated here from the reference's helpers:

* ``rolling_window``: bordado.rolling_window as the reference calls it
  (gradient_boosted.py:343-354): window centres on a grid spanning the
  region shrunk by half a window, spaced ``window_size * (1 - overlap)``
  (rounded to a whole number of steps, the spacing then adjusted so the
  centres span the region exactly), and the points within half a window of
  each centre in the infinity norm, found with a KD-tree (scipy's
  ``cKDTree.query_ball_point``, sorted indices); centres in northing-major
  order;
* the window order shuffled by ``sklearn.utils.shuffle`` with the
  estimator's ``random_state`` (gradient_boosted.py:358-362), and empty
  windows dropped (:364-369).
"""

import ctypes
import warnings

import numpy as np
import torch

from . import _lib
from ._eqs import (
    FIT_INFO_KEYS,
    MIXED_RTOL,
    EquivalentSources,
    cast_fit_input,
    check_fit_input,
    get_region,
    n_1d_arrays,
)
from ._forward import to_device
from ._lib import check, lib, ptr, stream_handle

_MODES = {"float64": _lib.HMX_FIT_FP64, "mixed": _lib.HMX_FIT_INT8_SPLIT,
          "mixed-bf16": _lib.HMX_FIT_BF16_SPLIT, "mixed-fp32": _lib.HMX_FIT_FP32_MFMA}


def rolling_window(coordinates, window_size, overlap, region):
    """(centres, indices): window centres (two 2d arrays, northing-major) and,
    per window in raveled order, the sorted indices of the points within
    ``window_size / 2`` of its centre in both directions."""
    from scipy.spatial import cKDTree

    w, e, s, n = region
    half = window_size / 2
    spacing = (1 - overlap) * window_size
    axes = []
    for lo, hi in ((w + half, e - half), (s + half, n - half)):
        length = hi - lo
        if length < 0:
            raise ValueError(
                f"Window size '{window_size}' is larger than the region '{region}'."
            )
        count = int(round(length / spacing)) + 1
        axes.append(np.linspace(lo, hi, count))
    centre_e, centre_n = np.meshgrid(axes[0], axes[1])
    points = np.column_stack([np.asarray(c, dtype=np.float64).ravel() for c in coordinates[:2]])
    tree = cKDTree(points)
    found = tree.query_ball_point(np.column_stack([centre_e.ravel(), centre_n.ravel()]), r=half,
                                  p=np.inf, return_sorted=True)
    return (centre_e, centre_n), [np.asarray(i, dtype=np.int64) for i in found]


def _get_region_data_sources(coordinates, points):
    """Region holding every datum and every source (gradient_boosted.py:372-397)."""
    data_region = get_region(coordinates)
    sources_region = get_region(points)
    return (
        min(data_region[0], sources_region[0]),
        max(data_region[1], sources_region[1]),
        min(data_region[2], sources_region[2]),
        max(data_region[3], sources_region[3]),
    )


#: Windows factorised per batch in the FP64 mode with damping: None picks
#: the largest batch whose buffers fit half of the free device memory (at
#: most all windows); 1 fits window by window.
FACTOR_BATCH = None


def _factor_batch(n, max_wd, max_ws, mode, damping, n_windows):
    """Windows per batch of the batched factorisation (1: none)."""
    if FACTOR_BATCH is not None:
        return max(1, int(FACTOR_BATCH))
    if mode != _lib.HMX_FIT_FP64 or not damping > 0 or n_windows < 2:
        return 1
    sizes = []
    for batch in (1, 2):
        nbytes = ctypes.c_size_t(0)
        check(lib().hmx_eqs_gb_fit_workspace_bytes(n, max_wd, max_ws, mode, damping, batch,
                                                   ctypes.byref(nbytes)))
        sizes.append(nbytes.value)
    per_window = max(sizes[1] - sizes[0], 1) / 2
    free, _ = torch.cuda.mem_get_info()
    fits = int((0.5 * free - sizes[0]) // per_window)
    return max(1, min(n_windows, fits))


class EquivalentSourcesGB(EquivalentSources):
    r"""
    Gradient-boosted equivalent sources (reference:
    gradient_boosted.py:24-138): the coefficients are fitted window by window
    over overlapping windows of equal size (50% overlap), each window fitting
    the residue left by the previous ones.

    Parameters are those of the reference (``damping``, ``points``,
    ``depth``, ``block_size``, ``window_size`` (``"default"``: about 5000
    data per window), ``parallel`` (accepted), ``random_state`` (window
    order), ``dtype``) plus ``fit_precision`` as in
    :class:`EquivalentSources` (the least-squares mode of every window).
    Attributes ``points_``, ``coefs_``, ``region_``, ``depth_``,
    ``window_size_`` and ``rmse_per_iteration_`` as in the reference;
    ``fit_info_`` reports the solver statistics summed over the windows.
    """

    # Overlap between adjacent windows (gradient_boosted.py:112-113).
    overlapping = 0.5

    def __init__(
        self,
        damping=None,
        points=None,
        depth="default",
        block_size=None,
        window_size="default",
        parallel=True,
        random_state=None,
        dtype="float64",
        fit_precision="float64",
    ):
        if isinstance(window_size, str) and window_size != "default":
            msg = (
                f"Found invalid 'window_size' value equal to '{window_size}'."
                "It should be 'default' or a numeric value."
            )
            raise ValueError(msg)
        super().__init__(damping=damping, points=points, depth=depth, block_size=block_size,
                         parallel=parallel, dtype=dtype, fit_precision=fit_precision)
        self.random_state = random_state
        self.window_size = window_size

    def estimate_required_memory(self, coordinates):
        """Bytes of the largest window jacobian (gradient_boosted.py:140-183)."""
        coordinates = n_1d_arrays(coordinates, 3)
        self.points_ = self._build_points(coordinates)
        source_windows, data_windows = self._create_windows(coordinates)
        sizes = np.array([s.size * d.size for s, d in zip(source_windows, data_windows)])
        return sizes.max() * np.dtype(self.dtype).itemsize

    def fit(self, coordinates, data, weights=None):
        """Fit the coefficients by gradient boosting (gradient_boosted.py:185-229)."""
        coordinates, data, weights = check_fit_input(coordinates, data, weights)
        coordinates, data, weights = cast_fit_input(coordinates, data, weights, self.dtype)
        self.region_ = get_region(coordinates[:2])
        coordinates = n_1d_arrays(coordinates, 3)
        data = np.ravel(data)
        if weights is not None:
            weights = np.ravel(weights)
        if self.points is None:
            self.points_ = tuple(p.astype(self.dtype) for p in self._build_points(coordinates))
        else:
            self.depth_ = None
            self.points_ = tuple(p.astype(self.dtype) for p in n_1d_arrays(self.points, 3))
        self.coefs_ = np.zeros_like(self.points_[0])
        self._gradient_boosting(coordinates, data, weights)
        return self

    def _gradient_boosting(self, coordinates, data, weights):
        """The window loop (gradient_boosted.py:231-283), on the GPU."""
        source_windows, data_windows = self._create_windows(coordinates)
        n_windows = len(source_windows)
        dev = _lib.require_device()
        c = [to_device(np.asarray(x, dtype=np.float64)) for x in coordinates[:3]]
        p = [to_device(np.asarray(x, dtype=np.float64)) for x in self.points_[:3]]
        d = to_device(np.asarray(data, dtype=np.float64))
        w = None if weights is None else to_device(np.asarray(weights, dtype=np.float64))
        n, m = c[0].numel(), p[0].numel()
        data_offset = np.zeros(n_windows + 1, dtype=np.int64)
        source_offset = np.zeros(n_windows + 1, dtype=np.int64)
        data_offset[1:] = np.cumsum([x.size for x in data_windows])
        source_offset[1:] = np.cumsum([x.size for x in source_windows])
        cat = lambda parts: (np.concatenate(parts).astype(np.int64) if parts
                             else np.zeros(1, dtype=np.int64))
        didx = torch.as_tensor(cat(data_windows)).to(dev)
        sidx = torch.as_tensor(cat(source_windows)).to(dev)
        max_wd = int(max((x.size for x in data_windows), default=1))
        max_ws = int(max((x.size for x in source_windows), default=1))
        mode = _MODES[self.fit_precision]
        max_iter = 10 if mode == _lib.HMX_FIT_FP64 else 300
        rtol = 1e-12 if mode == _lib.HMX_FIT_FP64 else MIXED_RTOL
        if self.damping is None:  # LinearRegression's minimum-norm least squares
            mode |= _lib.HMX_FIT_NO_DAMPING
        damping = float(0.0 if self.damping is None else self.damping)
        batch = _factor_batch(n, max_wd, max_ws, mode, damping, n_windows)
        nbytes = ctypes.c_size_t(0)
        check(lib().hmx_eqs_gb_fit_workspace_bytes(n, max_wd, max_ws, mode, damping, batch,
                                                   ctypes.byref(nbytes)))
        ws = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
        coefs = torch.empty(m, dtype=torch.float64, device=dev)
        residue = torch.empty(n, dtype=torch.float64, device=dev)
        rmse = torch.empty(n_windows + 1, dtype=torch.float64, device=dev)
        info = (ctypes.c_double * len(FIT_INFO_KEYS))()
        i64p = ctypes.POINTER(ctypes.c_int64)
        check(lib().hmx_eqs_gb_fit(
            *(ptr(x) for x in c), n, ptr(d), ptr(w), *(ptr(x) for x in p), m, ptr(didx),
            data_offset.ctypes.data_as(i64p), ptr(sidx), source_offset.ctypes.data_as(i64p),
            n_windows, damping, mode, batch, max_iter, rtol, ptr(coefs), ptr(residue), ptr(rmse),
            info, ptr(ws), nbytes.value, stream_handle(None)))
        self.coefs_ = coefs.cpu().numpy().astype(self.points_[0].dtype, copy=False)
        self.rmse_per_iteration_ = rmse.cpu().numpy()
        # summed over the windows, except the last update and the largest
        # shift and diagonal; "svd": windows solved by the SVD route
        self.fit_info_ = dict(zip(FIT_INFO_KEYS, list(info)))
        self.fit_info_["windows"] = n_windows
        self.fit_info_["factor_batch"] = batch

    def _create_windows(self, coordinates, shuffle=True):
        """Source and data indices of each window (gradient_boosted.py:285-369)."""
        from sklearn import utils

        region = _get_region_data_sources(coordinates, self.points_)
        if self.window_size == "default":
            area = (region[1] - region[0]) * (region[3] - region[2])
            ndata = coordinates[0].size
            if ndata <= 5e3:
                warnings.warn(
                    f"Found {ndata} number of coordinates (<= 5e3). "
                    "Only one window will be used.",
                    stacklevel=1,
                )
                self.window_size_ = None
                return [np.arange(self.points_[0].size)], [np.arange(ndata)]
            points_per_m2 = ndata / area
            window_area = 5e3 / points_per_m2
            self.window_size_ = np.sqrt(window_area)
        else:
            self.window_size_ = self.window_size
        _, source_windows = rolling_window(self.points_[:2], self.window_size_,
                                           self.overlapping, region)
        _, data_windows = rolling_window(coordinates[:2], self.window_size_, self.overlapping,
                                         region)
        if shuffle:
            source_windows, data_windows = utils.shuffle(source_windows, data_windows,
                                                         random_state=self.random_state)
        source_nonempty, data_nonempty = [], []
        for src, dat in zip(source_windows, data_windows):
            if src.size > 0 and dat.size > 0:
                source_nonempty.append(src)
                data_nonempty.append(dat)
        return source_nonempty, data_nonempty
