"""
Host-side behaviour of EquivalentSourcesGB that needs no GPU: the window
geometry, the default window size, the memory estimate and the messages
(test/test_gradient_boosted_eqs.py:94-156, 238-301, 404-442), the KD-tree
window membership against a direct comparison (the oracle's restatement),
and argument checks of hmx_eqs_gb_fit that return before any device call.
"""

import ctypes

import numpy as np
import numpy.testing as npt
import pytest

import oracle
from conftest import grid_coordinates
from harmonica_amd import EquivalentSourcesGB, _lib
from harmonica_amd._eqs_gb import _get_region_data_sources, rolling_window


def grid3(region, spacing=None, shape=None, upward=0.0):
    return grid_coordinates(region, shape=shape, spacing=spacing, extra=upward)


@pytest.mark.parametrize(
    ("data_region", "sources_region", "expected_region"),
    [
        ((-1, 1, -1, 1), (-2, 2, -2, 2), (-2, 2, -2, 2)),
        ((-3, 3, -3, 3), (-3, 3, -3, 3), (-3, 3, -3, 3)),
        ((-2, 2, -2, 2), (-2, 1, -2, 1), (-2, 2, -2, 2)),
        ((-2, 2, -2, 2), (-2, 1, -1, 2), (-2, 2, -2, 2)),
        ((-2, 2, -2, 2), (1, 2, 1, 2), (-2, 2, -2, 2)),
        ((-2, 2, -2, 2), (1, 2, -2, 1), (-2, 2, -2, 2)),
        ((-2, 2, -2, 2), (-3, 1, -2, 1), (-3, 2, -2, 2)),
        ((-2, 2, -2, 2), (-3, 3, -3, 3), (-3, 3, -3, 3)),
    ],
)
def test_get_region_data_sources(data_region, sources_region, expected_region):
    """test_gradient_boosted_eqs.py:94-111."""
    coordinates = grid3(data_region, shape=(10, 10))
    points = grid3(sources_region, shape=(10, 10), upward=-10.0)
    npt.assert_allclose(_get_region_data_sources(coordinates, points), expected_region)


def test_same_number_of_windows_data_and_sources():
    """test_gradient_boosted_eqs.py:238-260."""
    coordinates = grid3((1, 3, 1, 3), spacing=1)
    points = grid3((1.5, 2.5, 1.5, 2.5), spacing=1, upward=-10.0)
    eqs = EquivalentSourcesGB(window_size=1)
    eqs.points_ = points
    source_windows, data_windows = eqs._create_windows(coordinates)
    assert len(source_windows) == len(data_windows)


def test_same_windows_data_and_sources():
    """test_gradient_boosted_eqs.py:263-295: 9 windows of the expected
    sizes, northing-major."""
    coordinates = grid3((1, 3, 1, 3), spacing=1)
    points = grid3((1, 2, 1, 3), spacing=1, upward=-10.0)
    eqs = EquivalentSourcesGB(window_size=1)
    eqs.points_ = points
    source_windows, data_windows = eqs._create_windows(coordinates, shuffle=False)
    assert len(source_windows) == 9 and len(data_windows) == 9
    expected_data = np.array([[4, 2, 4], [2, 1, 2], [4, 2, 4]]).ravel()
    expected_sources = np.array([[4, 2, 2], [2, 1, 1], [4, 2, 2]]).ravel()
    assert [w.size for w in data_windows] == list(expected_data)
    assert [w.size for w in source_windows] == list(expected_sources)


@pytest.mark.parametrize("spacing", [100, 500, 1e3])
@pytest.mark.parametrize("window_size", [1e3, 2e3, 4e3])
@pytest.mark.parametrize(("dtype", "itemsize"), [("float32", 4), ("float64", 8)])
def test_memory_estimation(spacing, window_size, dtype, itemsize):
    """test_gradient_boosted_eqs.py:131-149."""
    coordinates = grid3((-1e3, 5e3, 2e3, 8e3), spacing=spacing)
    per_window = (int(window_size / spacing) + 1) ** 2
    eqs = EquivalentSourcesGB(window_size=window_size, dtype=dtype)
    assert eqs.estimate_required_memory(coordinates) == per_window * per_window * itemsize


def test_window_size_less_than_5000():
    """test_gradient_boosted_eqs.py:404-423."""
    coordinates = grid3((0, 10e3, -7e3, 3e3), shape=(20, 20))
    eqs = EquivalentSourcesGB()
    eqs.points_ = eqs._build_points(coordinates)
    with pytest.warns(UserWarning, match="Only one window will be used"):
        source_windows, data_windows = eqs._create_windows(coordinates)
    assert eqs.window_size_ is None
    assert len(source_windows) == 1 and len(data_windows) == 1
    npt.assert_array_equal(source_windows[0], np.arange(400))
    npt.assert_array_equal(data_windows[0], np.arange(400))


def test_window_size():
    """test_gradient_boosted_eqs.py:426-437."""
    coordinates = grid3((0, 10e3, -7e3, 3e3), shape=(100, 100))
    eqs = EquivalentSourcesGB()
    eqs.points_ = eqs._build_points(coordinates)
    eqs._create_windows(coordinates)
    npt.assert_allclose(eqs.window_size_, np.sqrt(5e3 / (100**2 / 10e3**2)))


def test_invalid_window_size():
    with pytest.raises(ValueError, match="Found invalid 'window_size' value equal to"):
        EquivalentSourcesGB(window_size="Chuckie took my soul!")


def test_invalid_and_zero_depth():
    """test_gradient_boosted_eqs.py:459-478."""
    with pytest.raises(ValueError, match="Found invalid 'depth' value equal to"):
        EquivalentSourcesGB(depth="This is not a valid depth")
    with pytest.raises(ValueError, match="Depth value cannot be zero"):
        EquivalentSourcesGB(depth=0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_kdtree_windows_match_direct_comparison(seed):
    """The KD-tree membership against the oracle's direct comparison, on
    random points and on a grid whose points fall on window edges."""
    rng = np.random.default_rng(seed)
    pts = (rng.uniform(-3e3, 2e3, 3000), rng.uniform(1e3, 4.5e3, 3000))
    grid = grid_coordinates((-3e3, 2e3, 1e3, 4.5e3), spacing=125.0)
    for coords in (pts, grid):
        for size in (400.0, 750.0, 1234.5):
            region = _get_region_data_sources(coords, coords)
            _, got = rolling_window(coords, size, 0.5, region)
            expected = oracle.rolling_window_windows(coords, size, 0.5, region)
            assert len(got) == len(expected)
            for g, x in zip(got, expected):
                npt.assert_array_equal(g, x)


def test_windows_cover_every_point_and_random_state():
    rng = np.random.default_rng(5)
    coords = (rng.uniform(0, 1e4, 20000), rng.uniform(0, 1e4, 20000), np.zeros(20000))
    eqs = EquivalentSourcesGB(random_state=3)
    eqs.points_ = eqs._build_points(coords)
    src_a, dat_a = eqs._create_windows(coords)
    other = EquivalentSourcesGB(random_state=3)
    other.points_ = eqs.points_
    src_b, dat_b = other._create_windows(coords)
    assert len(src_a) == len(src_b)
    for a, b in zip(dat_a, dat_b):
        npt.assert_array_equal(a, b)
    covered = np.zeros(20000, dtype=bool)
    for d in dat_a:
        covered[d] = True
    assert covered.all()
    # about 5000 data per window by default
    assert 2500 < np.median([d.size for d in dat_a]) < 7500


def test_gb_fit_argument_checks_before_device_use():
    L = _lib.lib()
    info = (ctypes.c_double * 11)()
    off_bad = np.array([1, 3], dtype=np.int64)
    off_empty = np.array([0, 0], dtype=np.int64)
    i64p = ctypes.POINTER(ctypes.c_int64)
    dummy = ctypes.c_void_p(8)
    for offsets, msg in ((off_bad, b"start at 0"), (off_empty, b"is empty")):
        st = L.hmx_eqs_gb_fit(dummy, dummy, dummy, 10, dummy, None, dummy, dummy, dummy, 10,
                              dummy, offsets.ctypes.data_as(i64p), dummy,
                              offsets.ctypes.data_as(i64p), 1, 0.0, 0, 1, 10, 1e-12, dummy,
                              dummy, dummy, info, dummy, 0, None)
        assert st == -1
        assert msg in L.hmx_last_error()
