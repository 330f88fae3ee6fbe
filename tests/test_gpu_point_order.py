"""
Point order of the prism entry points (point_order.hip): calls with at
least 2048 points and 2^30 pairs sort the points into wavefront-coherent
strips on the device and scatter the results back.  Checked here at 4096
points x 262,144 prisms (past the threshold) against the same entry points
called on 1024-point pieces (below it, so unsorted), for scattered points,
a raster grid, a line, one repeated point and a cluster with outliers:
every field, NaN at the same places, within 1e-13 x max|field| (the sorted
waves may take other closed forms, whose roundings differ); and against the
CPU oracle at sampled points (1e-10).  Also the layer accessor (with and
without NaN holes), prism_magnetic, the smallest sorted size, a ragged
point count and points with NaN coordinates.
"""

import numpy as np
import numpy.testing as npt
import pytest

import oracle

pytestmark = pytest.mark.gpu

N_OBS, N_PRISMS, PIECE = 4096, 262144, 1024
TEN = ("potential", "g_e", "g_n", "g_z", "g_ee", "g_nn", "g_zz", "g_en", "g_ez", "g_nz")


@pytest.fixture(scope="module")
def hm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import harmonica_amd

    return harmonica_amd


def assert_close(got, ref, tol, err_msg=""):
    got, ref = np.asarray(got), np.asarray(ref)
    npt.assert_array_equal(np.isnan(got), np.isnan(ref), err_msg=err_msg)
    finite = np.isfinite(ref)
    scale = np.abs(ref[finite]).max() if finite.any() else 1.0
    npt.assert_allclose(got[finite], ref[finite], rtol=0, atol=tol * scale, err_msg=err_msg)


@pytest.fixture(scope="module")
def model():
    rng = np.random.default_rng(21)
    w = rng.uniform(0, 2e4, N_PRISMS)
    s = rng.uniform(0, 2e4, N_PRISMS)
    t = -rng.uniform(0, 1500, N_PRISMS)
    b = t - rng.uniform(50, 2000, N_PRISMS)
    prisms = np.column_stack([w, w + rng.uniform(50, 2000, N_PRISMS), s,
                              s + rng.uniform(50, 2000, N_PRISMS), b, t])
    density = rng.normal(0, 300, N_PRISMS)
    mag = tuple(rng.normal(0, 2.0, N_PRISMS) for _ in range(3))
    return prisms, density, mag


def _points(kind):
    rng = np.random.default_rng(22)
    up = rng.uniform(-800, 400, N_OBS)  # some points inside prisms
    if kind == "scattered":
        return rng.uniform(0, 2e4, N_OBS), rng.uniform(0, 2e4, N_OBS), up
    if kind == "grid":
        e, n = np.meshgrid(np.linspace(0, 2e4, 64), np.linspace(0, 2e4, 64))
        return e.ravel(), n.ravel(), np.full(N_OBS, 300.0)
    if kind == "line":
        return np.linspace(0, 2e4, N_OBS), np.full(N_OBS, 9876.5), up
    if kind == "one_point":
        return np.full(N_OBS, 1.2e4), np.full(N_OBS, 7.3e3), np.full(N_OBS, 100.0)
    if kind == "cluster":
        e, n = rng.uniform(5e3, 5.1e3, N_OBS), rng.uniform(8e3, 8.1e3, N_OBS)
        e[::700], n[::700] = rng.uniform(-5e4, 5e4, e[::700].size), rng.uniform(-5e4, 5e4, n[::700].size)
        return e, n, up
    raise ValueError(kind)


def _fields(hm, coords, prisms, density, fields):
    from harmonica_amd._forward import prism_fields_device, to_device

    e, n, u = (to_device(np.ascontiguousarray(c)) for c in coords)
    return prism_fields_device(e, n, u, to_device(prisms), to_device(density), fields).cpu().numpy()


def _pieces(fn, coords):
    parts = [fn(tuple(c[k:k + PIECE] for c in coords)) for k in range(0, N_OBS, PIECE)]
    return np.concatenate(parts, axis=-1)


@pytest.mark.parametrize("kind", ["scattered", "grid", "line", "one_point", "cluster"])
def test_sorted_matches_unsorted_pieces(hm, model, kind):
    prisms, density, _ = model
    coords = _points(kind)
    for fields in (("g_z",), TEN):
        got = _fields(hm, coords, prisms, density, fields)
        ref = _pieces(lambda c: _fields(hm, c, prisms, density, fields), coords)
        for k, f in enumerate(fields):
            assert_close(got[k], ref[k], 1e-13, err_msg=f"{kind} {f} ({len(fields)} fields)")


def test_sorted_matches_oracle(hm, model):
    prisms, density, _ = model
    coords = _points("scattered")
    got = _fields(hm, coords, prisms, density, TEN)
    idx = np.arange(0, N_OBS, N_OBS // 32)
    sample = tuple(c[idx] for c in coords)
    for k, f in enumerate(TEN):
        ref = oracle.prism_gravity(sample, prisms, density, f)
        assert_close(got[k][idx], ref, 1e-10, err_msg=f)


def test_sorted_layer_accessor(hm):
    rng = np.random.default_rng(23)
    easting, northing = np.linspace(0, 2e4, 512), np.linspace(0, 2e4, 512)
    surface = rng.uniform(-500, 800, (512, 512))
    density = rng.uniform(2000, 3000, (512, 512))
    layer = hm.prism_layer((easting, northing), surface, -1000.0, properties={"density": density})
    coords = _points("scattered")
    for field in ("g_z", "g_zz"):
        got = layer.prism_layer.gravity(coords, field=field)
        ref = _pieces(lambda c: layer.prism_layer.gravity(c, field=field), coords)
        assert_close(got, ref, 1e-13, err_msg=field)


def test_sorted_layer_with_holes(hm):
    """A layer whose density has NaN holes (cells the forward model skips,
    as the reference's layer accessor does) and a top surface with NaN
    holes, through the sorted path: the same as 1024-point pieces."""
    rng = np.random.default_rng(24)
    easting, northing = np.linspace(0, 2e4, 512), np.linspace(0, 2e4, 512)
    surface = rng.uniform(-500, 800, (512, 512))
    density = rng.uniform(2000, 3000, (512, 512))
    density[rng.random((512, 512)) < 0.1] = np.nan
    surface[rng.random((512, 512)) < 0.05] = np.nan
    layer = hm.prism_layer((easting, northing), surface, -1000.0, properties={"density": density})
    coords = _points("grid")
    for field in ("g_z", "g_ez"):
        with pytest.warns(UserWarning, match="NaN values in 'density'"):
            got = layer.prism_layer.gravity(coords, field=field, density_name="density")
            ref = _pieces(
                lambda c: layer.prism_layer.gravity(c, field=field, density_name="density"), coords)
        assert np.isfinite(got).all()
        assert_close(got, ref, 1e-13, err_msg=field)


@pytest.mark.parametrize("kind", ["scattered", "grid"])
def test_sorted_magnetic(hm, model, kind):
    prisms, _, mag = model
    coords = _points(kind)
    got = np.asarray(hm.prism_magnetic(coords, prisms, mag, "b"))
    ref = _pieces(lambda c: np.asarray(hm.prism_magnetic(c, prisms, mag, "b")), coords)
    for k in range(3):
        assert_close(got[k], ref[k], 1e-13, err_msg=f"{kind} component {k}")
    if kind == "scattered":
        idx = np.arange(0, N_OBS, N_OBS // 32)
        sample = tuple(c[idx] for c in coords)
        expected = np.asarray(oracle.prism_magnetic(sample, prisms, mag, "b"))
        for k in range(3):
            assert_close(got[k][idx], expected[k], 1e-10, err_msg=f"oracle component {k}")


@pytest.mark.parametrize("n_obs", [2048, 5003])
def test_sorted_ragged_point_counts(hm, model, n_obs):
    """The smallest sorted call (2048 points) and a point count that leaves
    the last radix tile and the last workgroup partial."""
    prisms, density, _ = model
    rng = np.random.default_rng(n_obs)
    coords = (rng.uniform(0, 2e4, n_obs), rng.uniform(0, 2e4, n_obs), rng.uniform(-800, 400, n_obs))
    m = -(-(1 << 30) // n_obs)  # enough prisms for 2^30 pairs (repeated if needed)
    p, d = np.resize(prisms, (m, 6)), np.resize(density, m)
    got = _fields(hm, coords, p, d, TEN)
    parts = [_fields(hm, tuple(c[k:k + PIECE] for c in coords), p, d, TEN)
             for k in range(0, n_obs, PIECE)]
    ref = np.concatenate(parts, axis=-1)
    for k, f in enumerate(TEN):
        assert_close(got[k], ref[k], 1e-13, err_msg=f"{n_obs} points, {f}")


def test_sorted_with_nan_coordinates(hm, model):
    """Points with a NaN coordinate go to the first cell of the sort and give
    NaN at their own index; the other points are unaffected."""
    prisms, density, _ = model
    e, n, u = (np.array(c) for c in _points("scattered"))
    e[[5, 900]] = np.nan
    n[[17, 2048]] = np.nan
    u[[33, 4095]] = np.nan
    coords = (e, n, u)
    got = _fields(hm, coords, prisms, density, ("g_z", "g_zz"))
    ref = _pieces(lambda c: _fields(hm, c, prisms, density, ("g_z", "g_zz")), coords)
    bad = np.zeros(N_OBS, dtype=bool)
    bad[[5, 900, 17, 2048, 33, 4095]] = True
    for k in range(2):
        assert np.isnan(got[k][bad]).all()
        assert np.isfinite(got[k][~bad]).all()
        assert_close(got[k], ref[k], 1e-13)
