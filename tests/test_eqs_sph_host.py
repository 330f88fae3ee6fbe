"""
EquivalentSourcesSph without a GPU: the oracle's spherical Green's function
(oracle.c, or_eqs_sph_*) pinned against 40-digit mpmath distances and the
geocentric identity, and the host-side API of the reference's tests
(test/test_eq_sources_spherical.py:199-225, 300-340).
"""

import mpmath
import numpy as np
import numpy.testing as npt
import pytest

import oracle
from conftest import grid_coordinates
from harmonica_amd import EquivalentSourcesSph

mpmath.mp.dps = 40


def mp_distance(p, q):
    """Distance between (lon, lat, r) points (degrees, m) from geocentric
    Cartesian coordinates in 40-digit arithmetic."""
    def xyz(v):
        lon, lat, r = (mpmath.mpf(float(t)) for t in v)
        lon, lat = lon * mpmath.pi / 180, lat * mpmath.pi / 180
        return (r * mpmath.cos(lat) * mpmath.cos(lon), r * mpmath.cos(lat) * mpmath.sin(lon),
                r * mpmath.sin(lat))
    a, b = xyz(p), xyz(q)
    return mpmath.sqrt(sum((x - y) ** 2 for x, y in zip(a, b)))


def test_green_function_against_mpmath():
    """Far pairs to 1e-13; sources 500 m to 5 km below nearby data to the
    precision of the reference's formula, l^2 = dr^2 + 2 r r' (1 - cos psi),
    whose cos psi carries an absolute rounding of ~1e-16 (relative error
    of l up to ~2e-8 for l = 500 m at r = 6.4e6 m)."""
    rng = np.random.default_rng(3)
    far_err = near_err = 0.0
    for _ in range(200):
        p = (rng.uniform(-180, 180), rng.uniform(-89, 89), 6.4e6 + rng.uniform(-1e4, 1e4))
        q = (rng.uniform(-180, 180), rng.uniform(-89, 89), 6.4e6 - rng.uniform(1e3, 5e5))
        got = 1.0 / oracle.eqs_sph_jacobian(tuple([v] for v in p), tuple([v] for v in q))[0, 0]
        far_err = max(far_err, abs(got / float(mp_distance(p, q)) - 1))
        # a source up to 0.05 degrees off, 500 m to 5 km deeper
        q2 = (p[0] + rng.uniform(-0.05, 0.05), p[1] + rng.uniform(-0.05, 0.05),
              p[2] - rng.uniform(500, 5e3))
        got = 1.0 / oracle.eqs_sph_jacobian(tuple([v] for v in p), tuple([v] for v in q2))[0, 0]
        near_err = max(near_err, abs(got / float(mp_distance(p, q2)) - 1))
    assert far_err < 1e-13
    assert near_err < 5e-8


def test_predict_is_point_mass_potential():
    """predict = sum c_j / l_j = potential of point masses c_j / G (the
    oracle's separately pinned spherical point-mass loop)."""
    rng = np.random.default_rng(4)
    coords = (rng.uniform(-70, -60, 50), rng.uniform(-40, -30, 50), np.full(50, 6.4e6))
    pts = (rng.uniform(-70, -60, 30), rng.uniform(-40, -30, 30),
           6.4e6 - rng.uniform(1e4, 5e5, 30))
    coefs = rng.normal(size=30)
    pred = oracle.eqs_sph_predict(coords, pts, coefs)
    pot = oracle.point_gravity(coords, pts, coefs / oracle.GRAVITATIONAL_CONST, "potential",
                               coordinate_system="spherical")
    npt.assert_allclose(pred, pot, rtol=1e-13)
    jac = oracle.eqs_sph_jacobian(coords, pts)
    npt.assert_allclose(jac @ coefs, pred, rtol=1e-12)


@pytest.mark.parametrize("spacing", [100, 500, 1e3])
def test_memory_estimation(spacing):
    """test_eq_sources_spherical.py:306-320."""
    coordinates = grid_coordinates((-20, 20, -20, 20), spacing=spacing, extra=100.0)
    n = coordinates[0].size
    assert EquivalentSourcesSph().estimate_required_memory(coordinates) == n * n * 8


def test_memory_estimation_supplied_points():
    """test_eq_sources_spherical.py:323-340."""
    coordinates = grid_coordinates((-20, 20, -20, 20), shape=(6, 6), extra=100.0)
    points = grid_coordinates((-20, 20, -20, 20), shape=(3, 3), extra=50.0)
    eqs = EquivalentSourcesSph(points=points)
    assert eqs.estimate_required_memory(coordinates) == 9 * 36 * 8


def test_not_implemented_and_no_projection():
    """test_eq_sources_spherical.py:199-225."""
    eqs = EquivalentSourcesSph()
    with pytest.raises(NotImplementedError):
        eqs.scatter()
    with pytest.raises(NotImplementedError):
        eqs.profile(point1=(1, 1), point2=(2, 2), size=3)
    with pytest.raises(TypeError):
        eqs.grid(upward=10, projection=lambda a, b: (a * 2, b * 2))
