"""
Host-side behaviour of the drop-in API that needs no GPU: validation and
messages raised before any kernel call (test_prism.py:42-91,
test_point_gravity.py:41-113, test_prism_layer.py:79-279), prism-layer
construction and accessor attributes, and the C-ABI library exporting
every symbol of include/hmx.h.
"""

import ctypes
import os
import re

import numpy as np
import numpy.testing as npt
import pytest

import harmonica_amd as hm
from harmonica_amd import _forward, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    header = open(os.path.join(ROOT, "include", "hmx.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(hmx_\w+)\s*\(", header, re.M))
    assert len(declared) >= 8
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert L.hmx_version() >= 10000
    count = ctypes.c_int(-1)
    assert L.hmx_device_count(ctypes.byref(count)) == 0
    assert count.value >= 0


def test_invalid_field():
    msg = re.escape("Gravitational field this-field-does-not-exist not recognized")
    with pytest.raises(ValueError, match=msg):
        hm.prism_gravity([0, 0, 0], [-100, 100, -100, 100, -200, -100], 1000,
                         field="this-field-does-not-exist")


def test_invalid_density_array():
    prisms = [[-100, 0, -100, 0, -200, -100], [-100, 0, 0, 100, -200, -100],
              [0, 100, -100, 0, -200, -100], [0, 100, 0, 100, -200, -100]]
    msg = r"Number of elements in density \([0-9]+\) mismatch the number of prisms \([0-9]+\)"
    with pytest.raises(ValueError, match=msg):
        hm.prism_gravity([0, 0, 0], prisms, [1000, 2000], field="potential")


def test_invalid_prisms():
    w, e, s, n, b, t = -100, 100, -100, 100, -200, -100
    _forward.check_prisms(np.atleast_2d([w, e, s, n, b, t]))
    _forward.check_prisms(np.atleast_2d([w, w, s, n, b, t]))
    with pytest.raises(ValueError, match="The west boundary can't be greater than the east one"):
        _forward.check_prisms(np.atleast_2d([e, w, s, n, b, t]))
    with pytest.raises(ValueError, match="The south boundary can't be greater than the north one"):
        _forward.check_prisms(np.atleast_2d([w, e, n, s, b, t]))
    with pytest.raises(ValueError, match="The bottom boundary can't be greater than the top one"):
        _forward.check_prisms(np.atleast_2d([w, e, s, n, t, b]))
    with pytest.raises(ValueError, match="The bottom boundary can't be greater than the top one"):
        hm.prism_gravity([0, 0, 0], [0, 10, 0, 10, 0, -10], 100, field="potential")


def test_discard_null_prisms():
    """test_prism.py:94-112."""
    prisms = np.array([
        [0, 10, -50, 33, -2e3, 150], [-10, 0, 33, 66, -1e3, -100], [7, 7, -50, 50, -3e3, -1e3],
        [-50, 50, 7, 7, -3e3, -1e3], [-50, 50, -50, 50, -3e3, -3e3], [7, 7, 7, 7, -200, -200],
    ], dtype=float)
    density = np.array([2670, 0, 3300, 3000, 2800, 2700], dtype=float)
    p, d = _forward.discard_null_prisms(prisms, density)
    npt.assert_allclose(p, [[0.0, 10.0, -50.0, 33.0, -2e3, 150.0]])
    npt.assert_allclose(d, [2670])


def test_progressbar_requires_numba_progress():
    try:
        import numba_progress  # noqa: F401
        pytest.skip("numba_progress is installed")
    except ImportError:
        pass
    with pytest.raises(ImportError, match="numba_progress"):
        hm.prism_gravity([0, 0, 100], [-1, 1, -1, 1, -2, -1], 1.0, field="g_z", progressbar=True)


def test_point_gravity_errors():
    with pytest.raises(ValueError, match=re.escape("Coordinate system bla not recognized")):
        hm.point_gravity([0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 1.0, "potential", "bla")
    for field in ("g_n", "g_e"):
        with pytest.raises(NotImplementedError):
            hm.point_gravity([0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 1.0, field, "spherical")
    for system in ("spherical", "cartesian"):
        with pytest.raises(ValueError, match=re.escape("Gravitational field 'bla' not recognized")):
            hm.point_gravity([0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 1.0, "bla", system)
    msg = r"Number of elements in masses \([0-9]+\) mismatch the number of points \([0-9]+\)"
    with pytest.raises(ValueError, match=msg):
        hm.point_gravity([0, 0, 250], [[-10, 0, 10], [-10, 0, 10], [-100, 0, 100]], [1000, 2000],
                         field="potential")


# ----------------------------------------------------------- prism layers
@pytest.fixture
def dummy_layer():
    easting = np.linspace(-1, 3, 5)
    northing = np.linspace(7, 13, 4)
    surface = np.arange(20, dtype=float).reshape(4, 5)
    return (easting, northing), surface, 0, 2670 * np.ones((4, 5))


def test_prism_layer_construction(dummy_layer):
    """test_prism_layer.py:79-106."""
    (easting, northing), surface, reference, _ = dummy_layer
    layer = hm.prism_layer((easting, northing), surface, reference)
    for name in ("easting", "northing", "top", "bottom"):
        assert name in layer.coords
    npt.assert_allclose(layer.easting, easting)
    npt.assert_allclose(layer.northing, northing)
    npt.assert_allclose(layer.top, surface)
    npt.assert_allclose(layer.bottom, reference)
    surface[1, 1] = -1
    layer = hm.prism_layer((easting, northing), surface, reference)
    assert layer.top.values[1, 1] == 0 and layer.bottom.values[1, 1] == -1
    # meshgrid centres are accepted too
    E, N = np.meshgrid(easting, northing)
    layer2 = hm.prism_layer((E, N), surface, reference)
    npt.assert_allclose(layer2.easting, easting)


def test_prism_layer_invalid_surface_reference(dummy_layer):
    coordinates, surface, reference, _ = dummy_layer
    msg = r"Invalid surface array with shape '.+'\. Its shape should be compatible"
    with pytest.raises(ValueError, match=msg):
        hm.prism_layer(coordinates, np.arange(20, dtype=float), reference)
    msg = r"Invalid reference array with shape '.+'\. Its shape should be compatible"
    with pytest.raises(ValueError, match=msg):
        hm.prism_layer(coordinates, surface, np.zeros(20))


def test_prism_layer_properties_and_irregular(dummy_layer):
    coordinates, surface, reference, density = dummy_layer
    layer = hm.prism_layer(coordinates, surface, reference,
                           properties={"density": density, "suceptibility": 0 * density + 1e-3})
    npt.assert_allclose(layer.density, density)
    npt.assert_allclose(layer["suceptibility"], 1e-3)
    easting, northing = coordinates
    bad = easting.copy()
    bad[3] = -22
    with pytest.raises(ValueError, match="Passed easting coordinates are not evenly spaced"):
        hm.prism_layer((bad, northing), surface, reference)
    bad = northing.copy()
    bad[3] = -22
    with pytest.raises(ValueError, match="Passed northing coordinates are not evenly spaced"):
        hm.prism_layer((easting, bad), surface, reference)


def test_prism_layer_attributes():
    """test_prism_layer.py:184-201."""
    easting = np.linspace(1, 3, 5)
    northing = np.linspace(7, 10, 4)
    layer = hm.prism_layer((easting, northing), np.arange(20, dtype=float).reshape(4, 5), 0)
    assert layer.prism_layer.dims == ("northing", "easting")
    assert layer.prism_layer.spacing == (1, 0.5)
    assert layer.prism_layer.boundaries == (easting[0] - 0.25, easting[-1] + 0.25,
                                            northing[0] - 0.5, northing[-1] + 0.5)
    assert layer.prism_layer.size == 20
    assert layer.prism_layer.shape == (4, 5)


def test_prism_layer_to_prisms_and_get_prism():
    """test_prism_layer.py:204-237 (known arrays) and the layer.py doctest."""
    layer = hm.prism_layer((np.array([0, 1]), np.array([0, 1])), (np.arange(4) + 10).reshape(2, 2),
                           np.arange(4).reshape(2, 2))
    expected = [[-0.5, 0.5, -0.5, 0.5, 0, 10], [0.5, 1.5, -0.5, 0.5, 1, 11],
                [-0.5, 0.5, 0.5, 1.5, 2, 12], [0.5, 1.5, 0.5, 1.5, 3, 13]]
    npt.assert_allclose(layer.prism_layer._to_prisms(), expected)
    for i in range(2):
        for j in range(2):
            npt.assert_allclose(layer.prism_layer.get_prism((i, j)), expected[2 * i + j])
    doc = hm.prism_layer((np.linspace(0, 10, 5), np.linspace(2, 8, 4)),
                         np.arange(20, dtype=float).reshape((4, 5)), reference=0,
                         properties={"density": 2670.0 * np.ones((4, 5))})
    assert [float(b) for b in doc.prism_layer.boundaries] == [-1.25, 11.25, 1.0, 9.0]
    assert [float(b) for b in doc.prism_layer.get_prism((0, 2))] == [3.75, 6.25, 1.0, 3.0, 0.0, 2.0]


def test_prism_layer_invalid_field(dummy_layer):
    coordinates, surface, reference, density = dummy_layer
    layer = hm.prism_layer(coordinates, surface, reference, properties={"density": density})
    msg = re.escape("Gravitational field 'invalid field' not recognized.")
    with pytest.raises(ValueError, match=msg):
        layer.prism_layer.gravity((0, 0, 30), field="invalid field")


def test_workloads_are_seeded():
    from harmonica_amd import workloads

    a = workloads.terrain_layer_small()
    b = workloads.terrain_layer_small()
    npt.assert_array_equal(a["surface"], b["surface"])
    assert a["surface"].shape == (100, 100)
    assert a["coordinates"][0].shape == (64, 64)
    big = workloads.terrain_layer_large(n_prisms_side=100, n_obs_side=8)
    assert big["surface"].min() == 0.0 and big["surface"].max() == 2000.0


def test_batched_plan_on_host():
    """The work-list planner runs on the host: every (tile, chunk) slot once,
    longest items first, per-model slot bases and chunk counts."""
    obs_off = np.array([0, 300, 300, 1000], dtype=np.int64)
    pr_off = np.array([0, 5000, 6000, 6100], dtype=np.int64)
    L = _lib.lib()
    n = ctypes.c_int64(0)
    assert L.hmx_batched_plan_size(obs_off.ctypes.data_as(ctypes.c_void_p),
                                   pr_off.ctypes.data_as(ctypes.c_void_p), 3, 2048,
                                   ctypes.byref(n)) == 0
    # model 0: 2 tiles x 3 chunks, model 1: no points, model 2: 3 tiles x 1 chunk
    assert n.value == 2 * 3 + 0 + 3
    items = np.zeros(5 * n.value, dtype=np.int64)
    base, chunks = np.zeros(3, dtype=np.int64), np.zeros(3, dtype=np.int64)
    got = ctypes.c_int64(0)
    assert L.hmx_batched_plan(obs_off.ctypes.data_as(ctypes.c_void_p),
                              pr_off.ctypes.data_as(ctypes.c_void_p), 3, 2048,
                              items.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(got),
                              base.ctypes.data_as(ctypes.c_void_p),
                              chunks.ctypes.data_as(ctypes.c_void_p)) == 0
    it = items.reshape(-1, 5)
    assert sorted(it[:, 4]) == list(range(n.value))
    cost = it[:, 3] - it[:, 2]
    assert np.all(np.diff(cost) <= 0)
    assert list(base) == [0, 6, 6] and list(chunks) == [3, 1, 1]


# ------------------------------------------------------------ prism_magnetic
MAG_PRISMS = [[-100, 0, -100, 0, -200, -100], [-100, 0, 0, 100, -200, -100],
              [0, 100, -100, 0, -200, -100], [0, 100, 0, 100, -200, -100]]


def test_magnetic_invalid_field():
    """test_prism_magnetic.py:29-36."""
    with pytest.raises(ValueError, match="Invalid field"):
        hm.prism_magnetic([0, 0, 0], [-100, 100, -100, 100, -200, -100], [1000, 1, 2],
                          field="Not a valid field")
    msg = re.escape("Invalid field 'g_z'. Please choose one of 'b,b_e,b_n,b_u'.")
    with pytest.raises(ValueError, match=msg):
        hm.prism_magnetic([0, 0, 0], MAG_PRISMS[0], [1.0, 1.0, 1.0], field="g_z")


@pytest.mark.parametrize("field", ["b", "b_e", "b_n", "b_u"])
def test_magnetic_invalid_magnetization(field):
    """test_prism_magnetic.py:169-225: two vectors for four prisms, five
    elements per component, four components."""
    with pytest.raises(ValueError, match="Number of magnetization vectors"):
        hm.prism_magnetic([0, 0, 0], MAG_PRISMS, ([1.0, -2.0], [1.0, 3.0], [1.0, -5.0]), field)
    five = ([1.0, 1.0, 1.0, 3.4, 1.3], [-2.0, 3.0, -5.0, 3.4, 1.3], [-2.0, 3.0, -5.0, 3.4, 1.3])
    with pytest.raises(ValueError, match=re.escape("Number of magnetization vectors (5) "
                                                   "mismatch the number of prisms (4)")):
        hm.prism_magnetic([0, 0, 0], MAG_PRISMS, five, field)
    four = tuple([1.0, 1.0, 1.0, 3.4] for _ in range(4))
    with pytest.raises(ValueError, match=re.escape("Invalid magnetization vectors with '4' "
                                                   "elements. Magnetization vectors should have "
                                                   "only 3 elements.")):
        hm.prism_magnetic([0, 0, 0], MAG_PRISMS, four, field)


@pytest.mark.parametrize("field", ["b", "b_u"])
@pytest.mark.parametrize("bad", [[100, 0, -100, 0, -200, -100], [-100, 0, 200, 0, -200, -100],
                                 [-100, 0, -100, 0, -100, -200]])
def test_magnetic_invalid_prisms(field, bad):
    """test_prism_magnetic.py:120-158."""
    with pytest.raises(ValueError, match="boundary can't be greater than the"):
        hm.prism_magnetic([0, 0, 0], bad, [1.0, 1.0, 1.0], field)


def test_magnetic_discard_null_prisms():
    from harmonica_amd import _magnetic

    prisms = np.array([[0, 10, 0, 10, -10, 0], [5, 5, 0, 10, -10, 0], [0, 10, 0, 10, -10, 0],
                       [0, 10, 3, 3, -10, 0]], dtype=float)
    mag = (np.array([1.0, 1.0, 0.0, 2.0]), np.array([0.0, 1.0, 0.0, 2.0]),
           np.array([0.0, 1.0, 0.0, 2.0]))
    p, m = _magnetic._discard_null_prisms(prisms, mag)
    npt.assert_array_equal(p, prisms[:1])
    npt.assert_array_equal(np.array(m), [[1.0], [0.0], [0.0]])


def test_magnetic_empty_coordinates():
    """No observation points: empty results of the coordinates' shape,
    without a kernel call."""
    b = hm.prism_magnetic((np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3))),
                          MAG_PRISMS[0], [1.0, 1.0, 1.0], "b")
    assert len(b) == 3 and all(x.shape == (0, 3) for x in b)
    assert hm.prism_magnetic(([], [], []), MAG_PRISMS[0], [1.0, 1.0, 1.0], "b_n").shape == (0,)


def test_strip_order_walks_grid_rows_boustrophedon():
    """Strips one grid row high: even rows by ascending easting, odd rows
    descending; the result is a permutation."""
    from harmonica_amd._eqs import strip_order

    e, n = np.meshgrid(np.arange(8.0), np.arange(4.0))
    e, n = e.ravel(), n.ravel()
    rng = np.random.default_rng(0)
    shuffle = rng.permutation(e.size)
    e, n = e[shuffle], n[shuffle]
    ratio = 0.9**2 * e.size / (64 * 7 * 3)  # strip height 0.9
    order = strip_order(e, n, ratio)
    npt.assert_array_equal(np.sort(order), np.arange(e.size))
    walked = np.column_stack([e[order], n[order]])
    expected = [(x, y) for y in range(4) for x in (range(8) if y % 2 == 0 else range(7, -1, -1))]
    npt.assert_array_equal(walked, np.array(expected, dtype=float))


def test_strip_order_runs_are_compact():
    """Runs of 64 consecutive random points stay within about h x h / ratio
    (northing spread below two strip heights), unlike the input order."""
    from harmonica_amd._eqs import strip_order

    rng = np.random.default_rng(1)
    e, n = rng.uniform(0, 2e4, 20000), rng.uniform(0, 2e4, 20000)
    ratio = 0.25
    h = np.sqrt(ratio * 64 * 2e4 * 2e4 / e.size)
    order = strip_order(e, n, ratio)
    npt.assert_array_equal(np.sort(order), np.arange(e.size))
    runs = n[order][: e.size // 64 * 64].reshape(-1, 64)
    spread = runs.max(axis=1) - runs.min(axis=1)
    assert np.all(spread < 2 * h)
    assert np.median(np.ptp(n[: e.size // 64 * 64].reshape(-1, 64), axis=1)) > 10 * h


def test_strip_order_degenerate_inputs():
    from harmonica_amd._eqs import strip_order

    assert strip_order([], []).size == 0
    npt.assert_array_equal(strip_order([5.0], [1.0]), [0])
    npt.assert_array_equal(strip_order([3.0, 1.0, 2.0], [7.0, 7.0, 7.0]), [1, 2, 0])
    npt.assert_array_equal(strip_order([7.0, 7.0, 7.0], [3.0, 1.0, 2.0]), [1, 2, 0])


def test_forward_workspace_queries_reject_sizes_past_int32():
    """The workspace queries refuse what the entry points refuse (2^31 or
    more points or sources) before touching a device, so no size product
    can overflow."""
    L = _lib.lib()
    nbytes = ctypes.c_size_t(0)
    for n_obs, n_src in ((1 << 31, 10), (10, 1 << 31), (1 << 40, 1 << 40)):
        st = L.hmx_forward_workspace_bytes(n_obs, n_src, 0x8, ctypes.byref(nbytes))
        assert st == _lib.HMX_ERR_ARG
        assert b"2^31-1" in L.hmx_last_error()
        st = L.hmx_prism_magnetic_workspace_bytes(n_obs, n_src, 0x7, ctypes.byref(nbytes))
        assert st == _lib.HMX_ERR_ARG
        assert b"2^31-1" in L.hmx_last_error()
