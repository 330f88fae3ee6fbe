"""
Host-side logic of EquivalentSources that needs no GPU
(test_eq_sources_cartesian.py:201-252,407-498): source placement, block
averaging against the reference's known arrays, default depth, memory
estimate, parameter validation.
"""

import numpy as np
import numpy.testing as npt
import pytest

from conftest import grid_coordinates
from harmonica_amd import EquivalentSources
from harmonica_amd._eqs import block_median, check_fit_input, get_region

REGION = (-3e3, -1e3, 5e3, 7e3)


@pytest.fixture
def coordinates():
    return grid_coordinates(REGION, shape=(40, 40), extra=0)


@pytest.fixture
def coordinates_9x9():
    e, n = grid_coordinates(REGION, shape=(9, 9))
    return e, n, np.arange(81, dtype=float).reshape(9, 9)


def test_build_points(coordinates):
    eqs = EquivalentSources(depth=1.5e3)
    points = eqs._build_points(coordinates)
    npt.assert_allclose(points, (*coordinates[:2], coordinates[-1] - 1.5e3))


@pytest.mark.parametrize("block_size", [750, (750, 1e3)], ids=["float", "tuple"])
def test_block_averaging_coordinates(coordinates_9x9, block_size):
    """Known arrays of test_eq_sources_cartesian.py:215-236."""
    eqs = EquivalentSources(depth=1.5e3, block_size=block_size)
    if np.ndim(block_size):
        expected = ([-2500.0, -1375.0, -2500.0, -1375.0, -2500.0, -1375.0],
                    [5250.0, 5250.0, 6000.0, 6000.0, 6750.0, 6750.0],
                    [11.0, 15.5, 38.0, 42.5, 65.0, 69.5])
    else:
        expected = ([-2750, -2000, -1250, -2750, -2000, -1250, -2750, -2000, -1250],
                    [5250, 5250, 5250, 6000, 6000, 6000, 6750, 6750, 6750],
                    [10.0, 13.0, 16.0, 37.0, 40.0, 43.0, 64.0, 67.0, 70.0])
    npt.assert_allclose(expected, eqs._block_average_coordinates(coordinates_9x9))


def test_build_points_block_average(coordinates_9x9):
    eqs = EquivalentSources(depth=1.5e3, block_size=750)
    expected = [np.array([-2750, -2000, -1250] * 3), np.repeat([5250, 6000, 6750], 3),
                np.array([10.0, 13.0, 16.0, 37.0, 40.0, 43.0, 64.0, 67.0, 70.0]) - 1.5e3]
    npt.assert_allclose(expected, eqs._build_points(coordinates_9x9))


def test_default_depth(coordinates):
    """4.5 x the first-neighbour distance on a regular grid (:448-459)."""
    eqs = EquivalentSources()
    eqs._build_points(tuple(c.ravel() for c in coordinates))
    spacing = min(coordinates[0][1, 1] - coordinates[0][0, 0], coordinates[1][1, 1] - coordinates[1][0, 0])
    npt.assert_allclose(eqs.depth_, 4.5 * spacing)


def test_invalid_and_zero_depth():
    with pytest.raises(ValueError, match="Found invalid 'depth' value equal to 'this is not'"):
        EquivalentSources(depth="this is not")
    with pytest.raises(ValueError, match="Depth value cannot be zero"):
        EquivalentSources(depth=0)


def test_scatter_not_implemented():
    with pytest.raises(NotImplementedError):
        EquivalentSources().scatter()


@pytest.mark.parametrize("spacing", [100, 500, 1e3])
@pytest.mark.parametrize(("dtype", "itemsize"), [("float32", 4), ("float64", 8)])
def test_memory_estimation(spacing, dtype, itemsize):
    coords = grid_coordinates((-1e3, 5e3, 2e3, 8e3), spacing=spacing, extra=0)
    n = coords[0].size
    assert EquivalentSources(dtype=dtype).estimate_required_memory(coords) == n * n * itemsize


def test_memory_estimation_doctest():
    """cartesian.py:211-223: 100 points -> 80000 bytes."""
    rng = np.random.default_rng(42)
    coords = (rng.uniform(-1e3, 3e3, 100), rng.uniform(2e3, 5e3, 100), np.full(100, 100.0))
    assert int(EquivalentSources().estimate_required_memory(coords)) == 80000


def test_fit_input_checks():
    c = (np.zeros(5), np.zeros(5), np.zeros(5))
    with pytest.raises(ValueError, match="Data arrays must have the same shape"):
        check_fit_input(c, np.zeros(4), None)
    with pytest.raises(ValueError, match="Weights must have the same shape"):
        check_fit_input(c, np.zeros(5), np.zeros(3))
    assert get_region((np.array([1.0, -2.0]), np.array([5.0, 3.0]))) == (-2.0, 1.0, 3.0, 5.0)


def test_block_median_drops_empty_blocks():
    e = np.array([0.0, 0.1, 9.9, 10.0])
    n = np.array([0.0, 0.0, 10.0, 10.0])
    out = block_median((e, n, np.arange(4.0)), 5.0)
    npt.assert_allclose(out[0], [0.05, 9.95])
    npt.assert_allclose(out[2], [0.5, 2.5])


def test_get_params_roundtrip():
    eqs = EquivalentSources(damping=0.1, depth=500, fit_precision="mixed")
    params = eqs.get_params()
    assert params["damping"] == 0.1 and params["fit_precision"] == "mixed"
    with pytest.raises(ValueError, match="fit_precision"):
        EquivalentSources(fit_precision="half")


def test_fit_workspace_bytes_per_mode():
    """hmx_eqs_fit_workspace_bytes (host arithmetic only): the FP64 mode
    holds the FP64 scaled jacobian, the FP32 mode an FP32 one plus a copy of
    the normal matrix, the split modes four slices per entry (2 bytes each
    for bf16, 1 for int8, rows padded to 32 or 64) and one scale per column
    and group of 1024 rows.  Fits small enough for the SVD route (at most
    8192 sources and 2^27 jacobian entries) get at least the FP64 jacobian
    behind the normal matrix's copy; the damping=None flag does not change
    the size; unknown modes are rejected."""
    import ctypes

    from harmonica_amd import _lib

    L = _lib.lib()

    def nbytes(n, m, mode):
        b = ctypes.c_size_t(0)
        st = L.hmx_eqs_fit_workspace_bytes(n, m, mode, ctypes.byref(b))
        return st, b.value

    def region(n, m, x_bytes, svd=True):
        return max(x_bytes, n * m * 8 - m * m * 8) if svd else x_bytes

    n, m = 5000, 700
    st, b64 = nbytes(n, m, _lib.HMX_FIT_FP64)
    assert st == _lib.HMX_OK
    common = b64 - n * m * 8
    st, b32 = nbytes(n, m, _lib.HMX_FIT_FP32_MFMA)
    assert st == 0 and b32 == common + m * m * 8 + region(n, m, n * m * 4)
    n_fold = 5
    st, bb = nbytes(n, m, _lib.HMX_FIT_BF16_SPLIT)
    assert st == 0 and bb == common + m * m * 8 + region(n, m, 4 * 5056 * m * 2 + n_fold * m * 8)
    st, bi = nbytes(n, m, _lib.HMX_FIT_INT8_SPLIT)
    assert st == 0 and bi == common + m * m * 8 + region(n, m, 4 * 5120 * m + n_fold * m * 8)
    for mode, b in ((_lib.HMX_FIT_FP64, b64), (_lib.HMX_FIT_INT8_SPLIT, bi)):
        assert nbytes(n, m, mode | _lib.HMX_FIT_NO_DAMPING) == (0, b)
    # beyond the SVD route's sizes the mode's own jacobian only
    n, m = 20000, 7000
    st, b64 = nbytes(n, m, _lib.HMX_FIT_FP64)
    common = b64 - n * m * 8
    st, bi = nbytes(n, m, _lib.HMX_FIT_INT8_SPLIT)
    assert st == 0 and bi == common + m * m * 8 + 4 * (20096 // 16) * 16 * m + 20 * m * 8
    assert nbytes(n, m, 4)[0] == _lib.HMX_ERR_ARG
    assert nbytes(0, m, _lib.HMX_FIT_FP64)[0] == _lib.HMX_ERR_ARG
