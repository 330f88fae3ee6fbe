"""
GPU tests of tesseroid_gravity (hmx_tesseroid_gravity): parity with the
oracle's restatement of the adaptive discretisation and quadrature, the
analytical spherical shells of the reference's tests (constant, linear and
exponential density; test_tesseroid.py:660-790,
test_tesseroid_variable_density.py:355-525), longitude continuity, inverted
tesseroids with checks disabled and the overflow errors.

Tolerance against the oracle: 1e-9 of the largest value.  Both evaluate the
reference's distance l^2 = dr^2 + 2 r r' (1 - cos psi) with different cos
roundings, which for quadrature nodes about a kilometre from a point differ
by ~1e-9 of the term.  Where the distance-size criterion of the
discretisation meets an exact tie (symmetric grids on tesseroid
boundaries), one ulp of cos/acos can add or remove one split, which moves
a result by up to the quadrature error; such configurations are compared at
the reference's accuracy threshold.
"""

import numpy as np
import numpy.testing as npt
import pytest

import oracle
from conftest import grid_coordinates

pytestmark = pytest.mark.gpu

G = oracle.GRAVITATIONAL_CONST
R = oracle.WGS84_MEAN_RADIUS
ACCURACY_THRESHOLD = 1e-3
TOL = 1e-9


@pytest.fixture(scope="module")
def hm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import harmonica_amd

    return harmonica_amd


def assert_parity(got, expected, tol=TOL, err_msg=""):
    got, expected = np.asarray(got), np.asarray(expected)
    npt.assert_allclose(got, expected, rtol=0, atol=tol * np.abs(expected).max(), err_msg=err_msg)


def shell_tesseroids(bottom, top, n_lon=12, n_lat=6, lon0=0.0):
    lon = np.linspace(lon0, lon0 + 360, n_lon + 1)
    lat = np.linspace(-90, 90, n_lat + 1)
    return [[w, e, s, n, bottom, top] for w, e in zip(lon[:-1], lon[1:])
            for s, n in zip(lat[:-1], lat[1:])]


@pytest.mark.parametrize("field", ["potential", "g_z"])
@pytest.mark.parametrize("radial", [False, True])
def test_tesseroids_vs_oracle(hm, field, radial):
    """60 random tesseroids (0.5-8 degrees, 1-50 km thick) and 400 points
    from 100 m to 200 km above the top radius."""
    rng = np.random.default_rng(2 + radial)
    m = 60
    w = rng.uniform(-60, 50, m)
    s = rng.uniform(-60, 50, m)
    top = R - rng.uniform(0, 2e3, m)
    tess = np.column_stack([w, w + rng.uniform(0.5, 8, m), s, s + rng.uniform(0.5, 8, m),
                            top - rng.uniform(1e3, 5e4, m), top])
    dens = rng.uniform(-400, 3300, m)
    n = 400
    coords = (rng.uniform(-70, 70, n), rng.uniform(-70, 70, n), R + rng.uniform(100, 2e5, n))
    got = hm.tesseroid_gravity(coords, tess, dens, field=field,
                               radial_adaptive_discretization=radial)
    expected = oracle.tesseroid_gravity(coords, tess, dens, field, radial=radial)
    assert_parity(got, expected, err_msg=f"{field} radial={radial}")


@pytest.mark.parametrize("field", ["potential", "g_z"])
def test_spherical_shell_two_dim(hm, field):
    """test_tesseroid.py:683-715 (2D discretisation, points on the shell's
    outer radius, thicknesses 10 m to 100 km)."""
    coordinates = grid_coordinates((0, 350, -90, 90), spacing=10, extra=R)
    for thickness in np.logspace(1, 5, 5):
        tess = shell_tesseroids(R - thickness, R)
        got = hm.tesseroid_gravity(coordinates, tess, 1000.0 * np.ones(len(tess)), field=field)
        potential = 4 / 3 * np.pi * G * 1000 * (R**3 - (R - thickness) ** 3) / R
        analytical = potential if field == "potential" else 1e5 * potential / R
        npt.assert_allclose(got, analytical, rtol=ACCURACY_THRESHOLD)
        # These points sit on the shell's outer radius and on meridians and
        # parallels of the tesseroid boundaries, where the distance-size
        # criterion meets exact ties; a tie decided the other way by one ulp
        # of cos/acos splits one piece more or less, which moves the result
        # by up to the quadrature error (~1e-4 here).  So the oracle is
        # compared at the reference's accuracy threshold on this
        # configuration, and at 1e-9 on tie-free models above.
        expected = oracle.tesseroid_gravity(coordinates, tess, 1000.0 * np.ones(len(tess)), field)
        assert_parity(got, expected, tol=ACCURACY_THRESHOLD)


@pytest.mark.parametrize("field", ["potential", "g_z"])
@pytest.mark.parametrize("thickness", [10, 1e3, 1e5])
def test_spherical_shell_three_dim(hm, field, thickness):
    """test_tesseroid.py:718-763 (3D discretisation, a point 1 km above)."""
    radius = R + 1e3
    tess = shell_tesseroids(R - thickness, R, 6, 6)
    got = hm.tesseroid_gravity([0, 0, radius], tess, 1000.0 * np.ones(36), field=field,
                               radial_adaptive_discretization=True)
    potential = 4 / 3 * np.pi * G * 1000 * (R**3 - (R - thickness) ** 3) / radius
    analytical = potential if field == "potential" else 1e5 * potential / radius
    npt.assert_allclose(got, analytical, rtol=ACCURACY_THRESHOLD)


@pytest.mark.parametrize("radial", [False, True])
@pytest.mark.parametrize("field", ["potential", "g_z"])
@pytest.mark.parametrize("thickness", [1e2, 1e3, 1e6])
def test_spherical_shell_linear_density(hm, field, thickness, radial):
    """test_tesseroid_variable_density.py:440-478, and the same with radial
    discretisation for points 1 km above the shell (on its top the radial
    splits never end, in the reference too: OverflowError)."""
    top = 6371e3
    bottom = top - thickness
    tess = shell_tesseroids(bottom, top, 12, 6, lon0=-180.0)
    density_outer, density_inner = 2500.0, 3300.0
    slope = (density_outer - density_inner) / (top - bottom)
    constant_term = density_outer - slope * top
    height = top + 1e3 if radial else top
    coordinates = grid_coordinates((-180, 180, -90, 90), shape=(19, 13), extra=height)
    constant = (np.pi * G * slope * (top**4 - bottom**4)
                + 4 / 3.0 * np.pi * G * constant_term * (top**3 - bottom**3))
    potential = constant / coordinates[2]
    analytical = potential if field == "potential" else 1e5 * potential / coordinates[2]
    got = hm.tesseroid_gravity(coordinates, tess, lambda r: slope * r + constant_term,
                               field=field, radial_adaptive_discretization=radial)
    npt.assert_allclose(got, analytical, rtol=ACCURACY_THRESHOLD)


@pytest.mark.parametrize("radial", [False, True])
@pytest.mark.parametrize("field", ["potential", "g_z"])
@pytest.mark.parametrize("thickness", [1e2, 1e6])
@pytest.mark.parametrize("b_factor", [5, 100])
def test_spherical_shell_exponential_density(hm, field, thickness, b_factor, radial):
    """test_tesseroid_variable_density.py:481-525, and the same with radial
    discretisation for points 1 km above the shell."""
    top = 6371e3
    bottom = top - thickness
    tess = shell_tesseroids(bottom, top, 12, 6, lon0=-180.0)
    density_outer, density_inner = 2670.0, 3300.0
    a_factor = (density_inner - density_outer) / (1 - np.exp(-b_factor))
    constant_term = density_inner - a_factor

    def density(radius):
        return a_factor * np.exp(-b_factor * (radius - bottom) / thickness) + constant_term

    height = top + 1e3 if radial else top
    coordinates = grid_coordinates((-180, 180, -90, 90), shape=(19, 13), extra=height)
    radius = coordinates[2]
    k = b_factor / thickness
    potential = (4 * np.pi * G * a_factor / k**3 / radius
                 * (((bottom * k) ** 2 + 2 * bottom * k + 2)
                    - ((top * k) ** 2 + 2 * top * k + 2) * np.exp(-k * thickness))
                 + 4 / 3 * np.pi * G * constant_term * (top**3 - bottom**3) / radius)
    analytical = potential if field == "potential" else 1e5 * potential / radius
    got = hm.tesseroid_gravity(coordinates, tess, density, field=field,
                               radial_adaptive_discretization=radial)
    npt.assert_allclose(got, analytical, rtol=ACCURACY_THRESHOLD)


@pytest.mark.parametrize("field", ["potential", "g_z"])
def test_variable_density_radial_discretisation_vs_oracle(hm, field):
    """Density functions with radial_adaptive_discretization=True: the GPU
    evaluates the density at the parts' radial nodes from each tesseroid's
    Chebyshev table; the oracle evaluates the function itself there
    (oracle.tesseroid_gravity_variable) on the same pieces of the
    density-based discretisation.  Points close to the tesseroids so that
    they divide radially; a linear profile (interpolated exactly) and an
    exponential one (interpolated to rounding)."""
    tess = np.array([[-2.0, 2.0, -2.0, 2.0, R - 4e4, R - 1e4],
                     [5.0, 8.0, -1.0, 3.0, R - 3e4, R - 2.5e4]])
    rng = np.random.default_rng(17)
    coords = (rng.uniform(-4, 10, 12), rng.uniform(-4, 5, 12), R + rng.uniform(1e2, 2e3, 12))
    for density in (lambda r: 3000.0 - 0.01 * (r - (R - 4e4)),
                    lambda r: 2700.0 + 100.0 * np.exp((r - R) / 2e4)):
        from harmonica_amd import _tesseroid

        pieces = _tesseroid.density_based_discretization(tess, density)
        got = hm.tesseroid_gravity(coords, tess, density, field=field,
                                   radial_adaptive_discretization=True)
        expected = oracle.tesseroid_gravity_variable(coords, pieces, density, field, radial=True)
        assert_parity(got, expected)
        if field == "g_z":  # (for the potential, ratio 1, these points divide no tesseroid radially)
            flat = hm.tesseroid_gravity(coords, tess, density, field=field)
            assert np.abs(flat - got).max() > 1e-7 * np.abs(got).max()


def test_variable_density_vs_oracle(hm):
    """The GPU with the host's density-based pieces and node densities
    against the oracle on the same pieces."""
    from harmonica_amd import _tesseroid

    tess = np.array(shell_tesseroids(R - 3e5, R, 8, 4))

    def density(r):
        return 3300.0 - 0.002 * (r - (R - 3e5)) + 200 * np.sin(r / 5e4)

    coordinates = grid_coordinates((-170, 170, -80, 80), shape=(9, 9), extra=R + 5e3)
    got = hm.tesseroid_gravity(coordinates, tess, density, field="g_z")
    pieces = _tesseroid.density_based_discretization(tess, density)
    rho = _tesseroid._node_densities(pieces, density)
    expected = oracle.tesseroid_gravity(coordinates, pieces, rho, "g_z")
    assert_parity(got, expected)


@pytest.mark.parametrize("field", ["potential", "g_z"])
def test_longitude_continuity_equivalent_tesseroids(hm, field):
    """test_tesseroid.py:445-463."""
    bottom, top = R - 1e4, R
    coordinates = [0, 0, R + 1e3]
    a = hm.tesseroid_gravity(coordinates, [-10, 10, -10, 10, bottom, top], 1e3, field=field)
    b = hm.tesseroid_gravity(coordinates, [350, 10, -10, 10, bottom, top], 1e3, field=field)
    npt.assert_allclose(a, b)


def test_disable_checks(hm):
    """test_tesseroid.py:224-245."""
    valid = [0.0, 10.0, 0.0, 10.0, 10.0, 20.0]
    invalid = [0.0, 10.0, 0.0, 10.0, 20.0, 10.0]
    with pytest.raises(ValueError, match="The bottom radius boundary can't be greater"):
        hm.tesseroid_gravity([0.0, 0.0, 10.0], invalid, 100.0, field="potential")
    inv = hm.tesseroid_gravity([0.0, 0.0, 10.0], invalid, 100.0, field="potential",
                               disable_checks=True)
    val = hm.tesseroid_gravity([0.0, 0.0, 10.0], valid, 100.0, field="potential")
    npt.assert_allclose(inv, -val)


@pytest.mark.parametrize("field", ["potential", "g_z"])
def test_west_greater_than_east_without_checks(hm, field):
    """With disable_checks=True the reference skips _check_tesseroids and
    with it the longitude continuity (tesseroid_gravity.py:171-174): a
    tesseroid whose west boundary exceeds its east one is discretised and
    integrated as given (the longitude steps are negative).  The GPU result
    matches the oracle's loop on the tesseroids as given, and with the
    checks on it matches the oracle after the continuity move."""
    tess = np.array([[350.0, 10.0, -10.0, 10.0, R - 1e4, R],
                     [340.0, -10.0, 20.0, 30.0, R - 2e4, R - 5e3],
                     [-20.0, 15.0, -40.0, -30.0, R - 1e4, R]])
    dens = np.array([1e3, 2e3, -5e2])
    rng = np.random.default_rng(5)
    coords = (rng.uniform(-180, 180, 60), rng.uniform(-60, 60, 60), np.full(60, R + 5e4))
    raw = hm.tesseroid_gravity(coords, tess, dens, field=field, disable_checks=True)
    assert_parity(raw, oracle.tesseroid_gravity(coords, tess, dens, field, continuity=False))
    moved = hm.tesseroid_gravity(coords, tess, dens, field=field)
    assert_parity(moved, oracle.tesseroid_gravity(coords, tess, dens, field))
    assert np.abs(raw - moved).max() > 1e-3 * np.abs(moved).max()


def test_stack_overflow_raises(hm):
    """A point on the centre of a tesseroid's top face with radial
    discretisation: the distance to the centre of the piece below the point
    stays below ten times its size, so the stack of 100 overflows
    (test_tesseroid.py:318-337 through the public function; checks disabled
    since the point is on the tesseroid)."""
    tess = [-10.0, 10.0, -10.0, 10.0, 0.5, 1.0]
    with pytest.raises(OverflowError):
        hm.tesseroid_gravity([0.0, 0.0, 0.75], tess, 1000.0, field="g_z",
                             radial_adaptive_discretization=True, disable_checks=True)


def test_shape_and_dtype(hm):
    coordinates = grid_coordinates((-10, 10, -10, 10), shape=(3, 4), extra=R + 1e4)
    g = hm.tesseroid_gravity(coordinates, [-1, 1, -1, 1, R - 1e4, R], 2670.0, field="g_z",
                             dtype=np.float32)
    assert g.shape == (3, 4) and g.dtype == np.float32


# ----------------------------------------------------------- tesseroid_layer
def _dummy_layer(hm):
    latitude = np.linspace(-10, 10, 6)
    longitude = np.linspace(-10, 10, 5)
    shape = (latitude.size, longitude.size)
    surface = R * np.ones(shape) + 1e3
    density = 2670 * np.ones(shape)
    layer = hm.tesseroid_layer((longitude, latitude), surface, R,
                               properties={"density": density})
    return layer, density


@pytest.mark.parametrize("field", ["potential", "g_z"])
def test_tesseroid_layer_gravity(hm, field):
    """test_tesseroid_layer.py:385-409: the accessor equals
    tesseroid_gravity on the layer's tesseroids."""
    layer, density = _dummy_layer(hm)
    coordinates = grid_coordinates((-12, 12, -12, 12), shape=(5, 5), extra=R + 2e3)
    tess = [layer.tesseroid_layer.get_tesseroid((i, j)) for i in range(6) for j in range(5)]
    expected = hm.tesseroid_gravity(coordinates, tess, density.ravel(), field=field)
    npt.assert_allclose(layer.tesseroid_layer.gravity(coordinates, field=field), expected)
    assert_parity(expected, oracle.tesseroid_gravity(coordinates, tess, density.ravel(), field))


def test_tesseroid_layer_holes_and_threshold(hm):
    """test_tesseroid_layer.py:412-532: NaN tops and densities are skipped
    (with a warning for the density), thin tesseroids discarded."""
    layer, density = _dummy_layer(hm)
    top = layer.top.values.copy()
    top[3, 3] = np.nan
    layer.tesseroid_layer.update_top_bottom(top, R)
    layer.data_vars["density"].values[2, 1] = np.nan
    coordinates = grid_coordinates((-12, 12, -12, 12), shape=(4, 4), extra=R + 2e3)
    with pytest.warns(UserWarning, match="Found missing values in 'density' property"):
        got = layer.tesseroid_layer.gravity(coordinates, field="g_z")
    keep = [(i, j) for i in range(6) for j in range(5) if (i, j) not in ((3, 3), (2, 1))]
    tess = [layer.tesseroid_layer.get_tesseroid(ij) for ij in keep]
    expected = hm.tesseroid_gravity(coordinates, tess, [2670.0] * len(keep), field="g_z")
    npt.assert_allclose(got, expected)
    # thin tesseroids
    top = R + 1e3 * np.ones((6, 5))
    top[0, :] = R + 10.0
    layer.tesseroid_layer.update_top_bottom(top, R)
    layer.data_vars["density"].values[:] = 2670.0
    got = layer.tesseroid_layer.gravity(coordinates, field="g_z", thickness_threshold=100.0)
    tess = [layer.tesseroid_layer.get_tesseroid((i, j)) for i in range(1, 6) for j in range(5)]
    npt.assert_allclose(got, hm.tesseroid_gravity(coordinates, tess, [2670.0] * 25, field="g_z"))
