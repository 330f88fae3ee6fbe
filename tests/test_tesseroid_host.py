"""
Tesseroids without a GPU: the oracle's restatement of the adaptive
discretisation and quadrature (oracle.c, or_tesseroid_*) pinned against the
analytical field of homogeneous spherical shells built of tesseroids (2D
and 3D discretisation, test_tesseroid.py:660-790) and a distant point
mass, the discretisation properties of the reference's tests
(test_tesseroid.py:318-337, 586-657), and the host-side checks and
messages of tesseroid_gravity (test_tesseroid.py:101-316).
"""

import re

import numpy as np
import numpy.testing as npt
import pytest

import oracle
from conftest import grid_coordinates
from harmonica_amd import _tesseroid
from harmonica_amd import tesseroid_gravity

G = oracle.GRAVITATIONAL_CONST
R = oracle.WGS84_MEAN_RADIUS
ACCURACY_THRESHOLD = 1e-3


def shell(top, bottom, density, radius):
    potential = 4 / 3 * np.pi * G * density * (top**3 - bottom**3) / radius
    return {"potential": potential, "g_z": 1e5 * potential / radius}


def shell_tesseroids(bottom, top, n_lon=12, n_lat=6):
    lon = np.linspace(0, 360, n_lon + 1)
    lat = np.linspace(-90, 90, n_lat + 1)
    return [[w, e, s, n, bottom, top] for w, e in zip(lon[:-1], lon[1:])
            for s, n in zip(lat[:-1], lat[1:])]


@pytest.mark.parametrize("field", ["potential", "g_z"])
def test_oracle_spherical_shell_two_dim(field):
    """test_tesseroid.py:683-715."""
    coordinates = grid_coordinates((0, 350, -90, 90), spacing=10, extra=R)
    for thickness in np.logspace(1, 5, 5):
        tess = shell_tesseroids(R - thickness, R)
        got = oracle.tesseroid_gravity(coordinates, tess, 1000.0 * np.ones(len(tess)), field)
        npt.assert_allclose(got, shell(R, R - thickness, 1000, R)[field], rtol=ACCURACY_THRESHOLD)


@pytest.mark.parametrize("field", ["potential", "g_z"])
@pytest.mark.parametrize("thickness", [10, 100, 1e3, 1e4, 1e5])
def test_oracle_spherical_shell_three_dim(field, thickness):
    """test_tesseroid.py:718-763."""
    radius = R + 1e3
    tess = shell_tesseroids(R - thickness, R, 6, 6)
    got = oracle.tesseroid_gravity(([0.0], [0.0], [radius]), tess, 1000.0 * np.ones(36), field,
                                   radial=True)
    npt.assert_allclose(got, shell(R, R - thickness, 1000, radius)[field],
                        rtol=ACCURACY_THRESHOLD)


@pytest.mark.parametrize("field", ["potential", "g_z"])
def test_oracle_far_field_point_mass(field):
    """A 1 x 1 degree, 10 km thick tesseroid seen from 3000 km: its mass
    rho (r_t^3 - r_b^3) / 3 (lon_e - lon_w) (sin lat_n - sin lat_s) at its
    centre of mass, to the size/distance squared."""
    t = [10.0, 11.0, 20.0, 21.0, R - 1e4, R]
    rho = 2670.0
    mass = rho * (t[5]**3 - t[4]**3) / 3 * np.radians(1.0) * (np.sin(np.radians(21))
                                                              - np.sin(np.radians(20)))
    point = (10.5, 48.0, R + 1e3)
    got = oracle.tesseroid_gravity(tuple([v] for v in point), [t], [rho], field)[0]
    pm = oracle.point_gravity(tuple([v] for v in point), ([10.5], [20.5], [R - 5e3]), [mass],
                              field, coordinate_system="spherical")[0]
    npt.assert_allclose(got, pm, rtol=2e-4)


def test_oracle_stack_and_discretisation_overflow():
    """test_tesseroid.py:318-336."""
    tess = [-10.0, 10.0, -10.0, 10.0, 0.5, 1.0]
    with pytest.raises(OverflowError):
        oracle.tesseroid_split_list((0.0, 0.0, 1.0), tess, 10, stack_size=2)
    with pytest.raises(OverflowError):
        oracle.tesseroid_split_list((0.0, 0.0, 1.0), tess, 10, max_out=2)


@pytest.mark.parametrize("radial", [True, False])
def test_oracle_discretisation_monotonic(radial):
    """test_tesseroid.py:586-639: fewer pieces farther away, more pieces for
    larger distance-size ratios."""
    tess = [-10.0, 10.0, -10.0, 10.0, 1.0, 10.0]
    radii = [10.1 if radial else 10.0, 10.5, 12.0, 13.0, 15.0, 20.0, 30.0]
    counts = [len(oracle.tesseroid_split_list((0.0, 0.0, r), tess, 10, radial)) for r in radii]
    assert all(a >= b for a, b in zip(counts, counts[1:]))
    counts = [len(oracle.tesseroid_split_list((0.0, 0.0, 10.2), tess, ratio, radial))
              for ratio in np.linspace(1, 10, 10)]
    assert all(a <= b for a, b in zip(counts, counts[1:]))


def test_oracle_two_dimensional_keeps_radii():
    """test_tesseroid.py:642-657."""
    pieces = oracle.tesseroid_split_list((0.0, 0.0, 10.0), [-10.0, 10.0, -10.0, 10.0, 1.0, 10.0],
                                         10)
    assert len(pieces) > 1
    assert np.all(pieces[:, 4] == 1.0) and np.all(pieces[:, 5] == 10.0)


def test_invalid_field_and_density():
    """test_tesseroid.py:101-131."""
    tess = [-10, 10, -10, 10, R - 1e3, R]
    with pytest.raises(ValueError, match="Gravitational field Not a valid field not recognized"):
        tesseroid_gravity([0, 0, R + 100], tess, 1000, field="Not a valid field")
    msg = r"Number of elements in density \(2\) mismatch the number of tesseroids \(4\)"
    four = [[-10, 0, -10, 0, R - 1e3, R], [-10, 0, 0, 10, R - 1e3, R],
            [0, 10, -10, 0, R - 1e3, R], [0, 10, 0, 10, R - 1e3, R]]
    with pytest.raises(ValueError, match=msg):
        tesseroid_gravity([0, 0, R + 100], four, [1000, 2000], field="potential")


def test_invalid_tesseroid():
    """test_tesseroid.py:155-221."""
    check = _tesseroid._check_tesseroids
    w, e, s, n, bottom, top = -10, 10, -10, 10, 100, 200
    cases = [
        ([w, e, n, s, bottom, top], "The south boundary can't be greater than the north one"),
        ([w, e, s, -100, bottom, top], re.escape("inside the [-90, 90] degrees interval")),
        ([w, e, s, 100, bottom, top], re.escape("inside the [-90, 90] degrees interval")),
        ([w, e, -100, n, bottom, top], re.escape("inside the [-90, 90] degrees interval")),
        ([w, e, s, n, top, bottom], "The bottom radius boundary can't be greater than the top"),
        ([w, e, s, n, bottom, -1], "The bottom and top radii should be positive or zero"),
        ([w, e, s, n, -1, top], "The bottom and top radii should be positive or zero"),
        ([-200, e, s, n, bottom, top], re.escape("inside the [-180, 360] degrees interval")),
        ([400, e, s, n, bottom, top], re.escape("inside the [-180, 360] degrees interval")),
        ([w, -200, s, n, bottom, top], re.escape("inside the [-180, 360] degrees interval")),
        ([w, 400, s, n, bottom, top], re.escape("inside the [-180, 360] degrees interval")),
        ([30, 0, s, n, bottom, top], "The west boundary can't be greater than the east one"),
        ([-60, -70, s, n, bottom, top], "The west boundary can't be greater than the east one"),
        ([300, -150, s, n, bottom, top], "The west boundary can't be greater than the east one"),
        ([350, 340, s, n, bottom, top], "The west boundary can't be greater than the east one"),
        ([-150, 300, s, n, bottom, top], "cannot be greater than one turn around the globe"),
    ]
    for tess, msg in cases:
        with pytest.raises(ValueError, match=msg):
            check(np.atleast_2d(np.array(tess, dtype=float)))
    # valid ones pass (test_tesseroid.py:134-152)
    for tess in ([w, e, s, n, bottom, top], [0, 360, s, n, bottom, top], [-180, 180, s, n, 0, top],
                 [350, 10, s, n, bottom, top]):
        check(np.atleast_2d(np.array(tess, dtype=float)))


def test_longitude_continuity_and_discard():
    """test_tesseroid.py:339-443."""
    cont = _tesseroid._longitude_continuity
    npt.assert_array_equal(cont(np.atleast_2d([-10.0, 10, -10, 10, 1, 2]))[0, :2], [-10, 10])
    npt.assert_array_equal(cont(np.atleast_2d([-70.0, -60, -10, 10, 1, 2]))[0, :2], [-70, -60])
    npt.assert_array_equal(cont(np.atleast_2d([350.0, 10, -10, 10, 1, 2]))[0, :2], [-10, 10])
    top, bottom = R, R - 1e3
    tess = np.array([
        [-10, -5, -10, -5, bottom, top], [-10, -5, -5, 0, bottom, top],
        [-10, -10, 0, 5, bottom, top], [-10, -5, 5, 5, bottom, top],
        [-5, 0, -10, -5, top, top], [-5, 0, -5, 0, bottom, top],
        [-5, -5, 5, 5, top, top], [-5, 0, 5, 10, bottom, top]])
    dens = np.array([2400, 0, 2500, 2600, 2700, 2800, 2900, 3000], dtype=float)
    t, d = _tesseroid._discard_null_tesseroids(tess, dens)
    npt.assert_allclose(t, tess[[0, 5, 7]])
    npt.assert_allclose(d, [2400, 2800, 3000])


def test_points_inside_tesseroids():
    """test_tesseroid.py:248-315."""
    check = _tesseroid.check_points_outside_tesseroids
    tess = np.atleast_2d([-10.0, 10, -10, 10, 100, 200])
    outside = np.array([[0, 0, 250], [20, 0, 150], [0, 20, 150], [0, 0, 200], [0, 0, 100],
                        [-10, 0, 150], [10, 0, 150], [0, -10, 150], [0, 10, 150]], float).T
    check(tuple(outside), tess)
    msg = re.escape("Found computation point(s) inside tesseroid(s)")
    for p, t in (([0, 0, 150], tess[0]), ([360, 0, 150], tess[0]),
                 ([-90, 0, 150], [260, 280, -10, 10, 100, 200])):
        with pytest.raises(ValueError, match=msg):
            check(tuple(np.atleast_2d(np.array(p, float)).T), np.atleast_2d(np.array(t, float)))
    tesss = np.atleast_2d([[-10, 10, -10, 10, 100, 200], [20, 30, 20, 30, 400, 500],
                           [-50, -40, -30, -20, 100, 500]]).astype(float)
    pts = np.array([[0, 0, 150], [80, 82, 4000], [10, 10, 450]], float).T
    with pytest.raises(ValueError) as err:
        check(tuple(pts), tesss)
    assert str(err.value).count(" - Computation point") == 1
    assert "'(0.0, 0.0, 150.0)' inside tesseroid '(-10.0, 10.0, -10.0, 10.0, 100.0, 200.0)'" \
        in str(err.value)


def test_density_based_discretisation():
    """_tesseroid_variable_density.py: a linear density needs no split
    (test_tesseroid_variable_density.py:276-291), a constant one neither,
    a curved one is split into pieces that tile the radius range."""
    t = np.array([[-10.0, 10, -10, 10, 1e3, 1e5]])
    lin = _tesseroid.density_based_discretization(t, lambda r: 2 * r + 10)
    npt.assert_array_equal(lin, t)
    const = _tesseroid.density_based_discretization(t, lambda r: 3000.0)
    npt.assert_array_equal(const, t)
    pieces = _tesseroid.density_based_discretization(t, lambda r: np.exp(-r / 1e4))
    assert len(pieces) >= 2
    edges = np.sort(np.concatenate([pieces[:, 4], pieces[:, 5]]))
    assert edges[0] == 1e3 and edges[-1] == 1e5


# ----------------------------------------------------------- tesseroid_layer
from harmonica_amd import tesseroid_layer  # noqa: E402
from harmonica_amd._tesseroid_layer import _discard_thin_tesseroids  # noqa: E402


def dummy_layer(reference_as_array=False):
    """test_tesseroid_layer.py:30-52."""
    latitude = np.linspace(-10, 10, 6)
    longitude = np.linspace(-10, 10, 5)
    shape = (latitude.size, longitude.size)
    surface = R * np.ones(shape) + 1e3
    reference = R * np.ones(shape) if reference_as_array else R
    density = 2670 * np.ones(shape)
    return (longitude, latitude), surface, reference, density


@pytest.mark.parametrize(("west", "east"), [(0, 480), (0, 360), (-180, 180), (0, 360 - 9)])
def test_tesseroid_overlap_invalid_coords(west, east):
    """test_tesseroid_layer.py:85-111."""
    latitude = np.linspace(-10, 10, 6)
    longitude = np.linspace(west, east, 21)
    shape = (latitude.size, longitude.size)
    with pytest.raises(ValueError, match="overlapping tesseroids around the globe"):
        tesseroid_layer((longitude, latitude), R * np.ones(shape) + 1e3, R * np.ones(shape))


@pytest.mark.parametrize(("west", "east"), [(0, 360 - 18), (-180, 180 - 18)])
def test_tesseroid_overlap_valid_coords(west, east):
    latitude = np.linspace(-10, 10, 6)
    longitude = np.linspace(west, east, 21)
    shape = (latitude.size, longitude.size)
    tesseroid_layer((longitude, latitude), R * np.ones(shape) + 1e3, R * np.ones(shape))


@pytest.mark.parametrize("as_array", [False, True])
def test_tesseroid_layer_construction(as_array):
    """test_tesseroid_layer.py:135-222."""
    (lon, lat), surface, reference, density = dummy_layer(as_array)
    layer = tesseroid_layer((lon, lat), surface, reference, properties={"density": density})
    npt.assert_allclose(layer.longitude, lon)
    npt.assert_allclose(layer.latitude, lat)
    npt.assert_allclose(layer.top, surface)
    npt.assert_allclose(layer.bottom, R)
    npt.assert_allclose(layer.density, density)
    assert layer.tesseroid_layer.dims == ("latitude", "longitude")
    assert layer.tesseroid_layer.shape == (6, 5) and layer.tesseroid_layer.size == 30
    npt.assert_allclose(layer.tesseroid_layer.spacing, (4, 5))
    npt.assert_allclose(layer.tesseroid_layer.boundaries, (-12.5, 12.5, -12, 12))
    assert layer.attrs["radius_units"] == "meters"
    # surface below the reference swaps top and bottom
    swapped = tesseroid_layer((lon, lat), R - 1e3 + 0 * surface, R)
    npt.assert_allclose(swapped.top, R)
    npt.assert_allclose(swapped.bottom, R - 1e3)
    with pytest.raises(ValueError, match="Invalid surface array with shape"):
        tesseroid_layer((lon, lat), surface[1:], reference)
    with pytest.raises(ValueError, match="Invalid reference array with shape"):
        tesseroid_layer((lon, lat), surface, np.ones((2, 2)))
    with pytest.raises(ValueError, match="not evenly spaced"):
        tesseroid_layer((np.array([0.0, 1, 3, 4, 5]), lat), surface, reference)


def test_tesseroid_layer_to_tesseroids_and_get_tesseroid():
    """test_tesseroid_layer.py:248-293."""
    (lon, lat), surface, reference, density = dummy_layer()
    layer = tesseroid_layer((lon, lat), surface, reference)
    t = layer.tesseroid_layer._to_tesseroids()
    assert t.shape == (30, 6)
    for i in range(6):
        for j in range(5):
            npt.assert_allclose(layer.tesseroid_layer.get_tesseroid((i, j)), t[i * 5 + j])
    npt.assert_allclose(layer.tesseroid_layer.get_tesseroid((0, 0)),
                        (-12.5, -7.5, -12, -8, R, R + 1e3))


def test_nonans_mask_and_thin_discard():
    """test_tesseroid_layer.py:296-384, 535-560."""
    (lon, lat), surface, reference, density = dummy_layer()
    surface = surface.copy()
    surface[1, 2] = np.nan
    density = density.copy()
    density[3, 3] = np.nan
    layer = tesseroid_layer((lon, lat), surface, reference, properties={"density": density})
    mask = layer.tesseroid_layer._get_nonans_mask()
    assert not mask[1, 2] and mask.sum() == 29
    with pytest.warns(UserWarning, match="Found missing values in 'density' property"):
        mask = layer.tesseroid_layer._get_nonans_mask(property_name="density")
    assert not mask[3, 3] and mask.sum() == 28
    tess = np.array([[-1, 1, -1, 1, 1.0, 2.0], [-1, 1, -1, 1, 1.0, 1.5],
                     [-1, 1, -1, 1, 1.0, 1.1]])
    t, d = _discard_thin_tesseroids(tess, np.array([1.0, 2.0, 3.0]), 0.5)
    npt.assert_array_equal(t, tess[:2])
    npt.assert_array_equal(d, [1.0, 2.0])


@pytest.mark.parametrize("profile", ["linear", "cubic", "exponential", "steep exponential"])
def test_chebyshev_density_tables(profile):
    """The per-tesseroid Chebyshev tables that carry a density function to
    the GPU for radial discretisation: evaluated at random radii of each
    tesseroid (numpy's chebval on the mapped radius), they match the
    function to within 1e-13 of its largest value (polynomials up to degree
    31 exactly, exponentials on the pieces of the density-based
    discretisation to rounding)."""
    from numpy.polynomial import chebyshev

    from harmonica_amd import _tesseroid

    R = 6371e3
    rng = np.random.default_rng(4)
    bottom = R - rng.uniform(1e2, 1e6, 20)
    top = bottom + rng.uniform(10.0, 3e5, 20)
    tess = np.column_stack([np.zeros(20), np.ones(20), np.zeros(20), np.ones(20), bottom, top])
    density = {
        "linear": lambda r: 3300.0 - 0.003 * (r - R),
        "cubic": lambda r: 2000.0 + 1e-15 * (r - R + 5e5) ** 3,
        "exponential": lambda r: 2670.0 + 630.0 * np.exp(-5 * (r - R + 1e6) / 1e6),
        "steep exponential": lambda r: 2670.0 + 630.0 * np.exp(-100 * (r - R + 1e6) / 1e6),
    }[profile]
    # the pieces of the density-based discretisation, as tesseroid_gravity
    # tabulates them
    tess = _tesseroid.density_based_discretization(tess, density)
    bottom, top = tess[:, 4], tess[:, 5]
    c = _tesseroid._chebyshev_densities(tess, density)
    assert c.shape == (tess.shape[0], _tesseroid.CHEB_N)
    for q in range(tess.shape[0]):
        r = rng.uniform(bottom[q], top[q], 50)
        x = (2 * r - (bottom[q] + top[q])) / (top[q] - bottom[q])
        f = np.array([density(v) for v in r])
        npt.assert_allclose(chebyshev.chebval(x, c[q]), f, rtol=0, atol=1e-13 * np.abs(f).max())


@pytest.mark.parametrize("radial", [False, True])
@pytest.mark.parametrize("field", ["potential", "g_z"])
def test_variable_density_oracle_with_constant_function(field, radial):
    """The Python variable-density oracle with a constant density function
    equals the C oracle's constant-density loop (itself pinned by the
    analytical shells above)."""
    R = oracle.WGS84_MEAN_RADIUS
    tess = np.array([[-2.0, 2.0, -2.0, 2.0, R - 4e4, R - 1e4],
                     [5.0, 8.0, -1.0, 3.0, R - 3e4, R - 2.5e4]])
    rng = np.random.default_rng(3)
    coords = (rng.uniform(-4, 10, 6), rng.uniform(-4, 5, 6), R + rng.uniform(1e2, 2e3, 6))
    got = oracle.tesseroid_gravity_variable(coords, tess, lambda r: 2670.0, field, radial=radial)
    expected = oracle.tesseroid_gravity(coords, tess, np.full(2, 2670.0), field, radial=radial)
    npt.assert_allclose(got, expected, rtol=1e-12, atol=1e-13 * np.abs(expected).max())
