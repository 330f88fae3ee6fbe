"""
GPU tests of EquivalentSourcesSph: the geocentric jacobian and predictions
against 40-digit mpmath and the oracle, the fit against the oracle's
least squares on its own jacobian, and the reference's checks
(test/test_eq_sources_spherical.py:73-196, 264-300).

Tolerances: the oracle restates the reference's distance,
l^2 = dr^2 + 2 r r' (1 - cos psi), whose rounding reaches ~2e-8 of l for
sources 500 m below nearby data; the GPU evaluates l as a geocentric
difference (exact to ~1e-12 there).  Entries are compared to mpmath at
1e-11 and to the oracle at its own precision.
"""

import mpmath
import numpy as np
import numpy.testing as npt
import pytest

import oracle
from conftest import grid_coordinates

pytestmark = pytest.mark.gpu

REGION = (-70, -60, -40, -30)
RADIUS = 6400e3
mpmath.mp.dps = 40


@pytest.fixture(scope="module")
def hm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import harmonica_amd

    return harmonica_amd


def checkerboard(amplitude, region, coordinates):
    """verde.synthetic.CheckerBoard with its default wavelengths."""
    w_e = (region[1] - region[0]) / 2
    w_n = (region[3] - region[2]) / 2
    return (amplitude * np.sin(2 * np.pi / w_e * coordinates[0])
            * np.cos(2 * np.pi / w_n * coordinates[1]))


def synthetic(shape, depth=500e3, points_shape=(6, 6)):
    points = grid_coordinates(REGION, shape=points_shape, extra=RADIUS - depth)
    masses = checkerboard(1e13, REGION, points)
    coordinates = grid_coordinates(REGION, shape=shape, extra=RADIUS)
    data = oracle.point_gravity(coordinates, points, masses, "g_z",
                                coordinate_system="spherical")
    return points, masses, coordinates, data


def mp_inv_distance(p, q):
    def xyz(v):
        lon, lat, r = (mpmath.mpf(float(t)) for t in v)
        lon, lat = lon * mpmath.pi / 180, lat * mpmath.pi / 180
        return (r * mpmath.cos(lat) * mpmath.cos(lon), r * mpmath.cos(lat) * mpmath.sin(lon),
                r * mpmath.sin(lat))
    a, b = xyz(p), xyz(q)
    return 1 / mpmath.sqrt(sum((x - y) ** 2 for x, y in zip(a, b)))


def test_jacobian_against_mpmath_and_oracle(hm):
    rng = np.random.default_rng(8)
    n, m = 60, 40
    coords = (rng.uniform(-180, 180, n), rng.uniform(-89, 89, n), 6.4e6 + rng.uniform(0, 1e4, n))
    # half the sources far, half 500 m - 5 km below data points
    far = (rng.uniform(-180, 180, m // 2), rng.uniform(-89, 89, m // 2),
           6.4e6 - rng.uniform(1e3, 5e5, m // 2))
    k = rng.integers(0, n, m // 2)
    near = (coords[0][k] + rng.uniform(-0.02, 0.02, m // 2),
            coords[1][k] + rng.uniform(-0.02, 0.02, m // 2),
            coords[2][k] - rng.uniform(500, 5e3, m // 2))
    pts = tuple(np.concatenate([a, b]) for a, b in zip(far, near))
    jac = hm.EquivalentSourcesSph().jacobian(coords, pts)
    exact = np.array([[float(mp_inv_distance((coords[0][i], coords[1][i], coords[2][i]),
                                             (pts[0][j], pts[1][j], pts[2][j])))
                       for j in range(m)] for i in range(n)])
    npt.assert_allclose(jac, exact, rtol=1e-11)
    ref = oracle.eqs_sph_jacobian(coords, pts)
    npt.assert_allclose(jac[:, : m // 2], ref[:, : m // 2], rtol=1e-13)
    npt.assert_allclose(jac, ref, rtol=5e-8)


def test_predict_and_fit_against_oracle(hm):
    """Moderately damped fit of 400 data: the GPU coefficients against the
    oracle's least squares on the oracle's jacobian, predictions against
    the oracle's loop."""
    points, masses, coordinates, data = synthetic((20, 20), depth=50e3)
    eqs = hm.EquivalentSourcesSph(relative_depth=30e3, damping=1e-3).fit(coordinates, data)
    jac = oracle.eqs_sph_jacobian(coordinates, eqs.points_)
    coefs = oracle.eqs_fit_coefs(jac, data.ravel(), None, 1e-3)
    grid = grid_coordinates(REGION, shape=(30, 30), extra=RADIUS + 2e3)
    pred = eqs.predict(grid)
    expected = oracle.eqs_sph_predict(grid, eqs.points_, coefs)
    npt.assert_allclose(pred, expected, rtol=0, atol=1e-7 * np.abs(expected).max())
    npt.assert_allclose(eqs.predict(coordinates), data, rtol=0, atol=1e-2 * np.abs(data).max())


def test_equivalent_sources_small_data_spherical(hm):
    """test_eq_sources_spherical.py:116-161."""
    points, masses, coordinates, data = synthetic((8, 8))
    eqs = hm.EquivalentSourcesSph(relative_depth=500e3, damping=1e-12)
    eqs.fit(coordinates, data)
    npt.assert_allclose(data, eqs.predict(coordinates), rtol=1e-5)
    tmp = [i.ravel() for i in coordinates]
    npt.assert_allclose(tmp[:2], eqs.points_[:2], rtol=1e-5)
    npt.assert_allclose(tmp[2] - 500e3, eqs.points_[2], rtol=1e-5)
    grid_coords = grid_coordinates(REGION, shape=(8, 8), extra=RADIUS + 2e3)
    true = oracle.point_gravity(grid_coords, points, masses, "g_z", coordinate_system="spherical")
    npt.assert_allclose(true, eqs.predict(grid_coords), rtol=0.05)
    grid = eqs.grid(grid_coords)
    npt.assert_allclose(true, grid.scalars, rtol=0.05)
    assert "radius" in grid.coords and grid.dims == ("spherical_latitude", "longitude")


def test_equivalent_sources_spherical(hm):
    """test_eq_sources_spherical.py:72-113 at the reference's tolerances:
    1600 data, sources 500 km deep, damping 1e-18.  The normal matrix is
    numerically singular, so the FP64 Cholesky factorisation fails and the
    fit takes scikit-learn's SVD route, as the reference's Ridge does."""
    points, masses, coordinates, data = synthetic((40, 40))
    eqs = hm.EquivalentSourcesSph(relative_depth=500e3, damping=1e-18)
    eqs.fit(coordinates, data)
    assert eqs.fit_info_["svd"] == 1.0 and eqs.fit_info_["svd_sweeps"] > 0
    npt.assert_allclose(data, eqs.predict(coordinates), rtol=1.3e-5, atol=1e-11)
    grid_coords = grid_coordinates(REGION, shape=(60, 60), extra=RADIUS)
    true = oracle.point_gravity(grid_coords, points, masses, "g_z", coordinate_system="spherical")
    npt.assert_allclose(true, eqs.predict(grid_coords), rtol=1e-3, atol=1e-11)
    grid = eqs.grid(grid_coords)
    npt.assert_allclose(true, grid.scalars, rtol=1e-3, atol=1e-11)


def test_custom_points_spherical(hm):
    """test_eq_sources_spherical.py:164-196."""
    points, masses, coordinates, data = synthetic((5, 5))
    points_custom = tuple(i.ravel() for i in grid_coordinates(REGION, shape=(3, 3),
                                                              extra=RADIUS - 500e3))
    eqs = hm.EquivalentSourcesSph(points=points_custom)
    eqs.fit(coordinates, data)
    npt.assert_allclose(points_custom, eqs.points_, rtol=1e-5)


@pytest.mark.parametrize(
    "deprecated_args",
    [
        {"upward": 5e3, "spacing": 1},
        {"upward": 5e3, "shape": (6, 6)},
        {"upward": 5e3, "spacing": 1, "region": (-75, -55, -40, -30)},
        {"upward": 5e3, "shape": (6, 6), "region": (-75, -55, -40, -30)},
    ],
)
def test_error_deprecated_args(hm, deprecated_args):
    """test_eq_sources_spherical.py:264-286."""
    small = grid_coordinates(REGION, shape=(5, 5), extra=0.0)
    small = (small[0], small[1], RADIUS + np.arange(25, dtype=float).reshape(5, 5))
    points = grid_coordinates(REGION, shape=(6, 6), extra=RADIUS - 1e3)
    data = oracle.point_gravity(small, points, checkerboard(1e13, REGION, points), "g_z",
                                coordinate_system="spherical")
    eqs = hm.EquivalentSourcesSph().fit(small, data)
    grid_coords = grid_coordinates(REGION, shape=(4, 4), extra=2e3)
    msg = "The 'upward', 'region', 'shape' and 'spacing' arguments have been"
    with pytest.raises(ValueError, match=msg):
        eqs.grid(coordinates=grid_coords, **deprecated_args)
    with pytest.warns(FutureWarning, match="The 'bla' arguments are being ignored."):
        eqs.grid(coordinates=grid_coords, bla="bla")
