"""
Pin the CPU oracle's point-mass, prism-layer and equivalent-source parts
against the reference's fixtures and against checks that do not restate the
oracle's formulas (finite differences, Laplace, symmetry, spherical closed
cases, scikit-learn's StandardScaler + Ridge for the fit).
"""

import os

import numpy as np
import numpy.testing as npt
import pytest

import oracle
from conftest import grid_coordinates

G = 6.6743e-11
HERE = os.path.dirname(os.path.abspath(__file__))
TENSORS = ("g_ee", "g_nn", "g_zz", "g_en", "g_ez", "g_nz")


# ---------------------------------------------------------------- point masses
def test_point_csv_fixture():
    """test_point_gravity.py:131-154 with the reference's CSV fixture."""
    e, n, u, pot = np.loadtxt(os.path.join(HERE, "golden", "sample_point_gravity.csv"),
                              delimiter=",", unpack=True)
    res = oracle.point_gravity((e, n, u), [0, 0, 0], [5000], "potential")
    npt.assert_allclose(res, pot)


SETS = [((0, -39, -13), (1, -67, -300.7), 250), ((-3, 24, -10), (20, 54, -500.7), 200)]


@pytest.mark.parametrize("coordinates,point,mass", SETS)
def test_point_finite_differences(coordinates, point, mass):
    """test_point_gravity.py:217-296 and 564-650 (same tolerances)."""
    delta = 0.05
    dist = np.sqrt(sum((a - b) ** 2 for a, b in zip(coordinates, point)))
    idx = {"e": 0, "n": 1, "z": 2}

    def shifted(field, axis):
        pair = [[c, c] for c in coordinates]
        pair[axis][0] -= delta
        pair[axis][1] += delta
        return oracle.point_gravity(tuple(np.array(p, dtype=float) for p in pair), point, mass, field)

    for f in ("e", "n", "z"):
        pot = shifted("potential", idx[f])
        fd = (pot[1] - pot[0]) / (2 * delta) * 1e5 * (-1 if f == "z" else 1)
        npt.assert_allclose(oracle.point_gravity(coordinates, point, mass, f"g_{f}"), fd,
                            rtol=1.5 * (delta / dist) ** 2)
    for field in TENSORS:
        di, dj = field[-2], field[-1]
        acc = shifted(f"g_{di}", idx[dj])
        fd = (acc[1] - acc[0]) / (2 * delta) * 1e-5 * 1e9
        if "z" in (di, dj):
            fd *= -1
        npt.assert_allclose(oracle.point_gravity(coordinates, point, mass, field), fd,
                            rtol=6 * (delta / dist) ** 2)


def test_point_laplace_and_symmetry():
    """test_point_gravity.py:354-373 and 157-214."""
    rng = np.random.default_rng(0)
    pts = (rng.uniform(2e3, 10e3, 30), rng.uniform(-3e3, 5e3, 30), np.full(30, -1e3))
    masses = np.arange(30.0)
    coords = grid_coordinates((2e3, 10e3, -3e3, 5e3), spacing=1e3, extra=0)
    trace = sum(oracle.point_gravity(coords, pts, masses, f) for f in ("g_ee", "g_nn", "g_zz"))
    scale = np.abs(oracle.point_gravity(coords, pts, masses, "g_zz")).max()
    npt.assert_allclose(trace, 0, atol=1e-12 * scale)
    for a, b in (("g_en", "g_ne"), ("g_ez", "g_ze"), ("g_nz", "g_zn")):
        npt.assert_allclose(oracle.point_gravity(coords, pts, masses, a),
                            oracle.point_gravity(coords, pts, masses, b))
    p, d = (1.1, 1.2, 1.3), 3.3
    e, n, u = (np.full(6, c) for c in p)
    e[0] += d; e[1] -= d; n[2] += d; n[3] -= d; u[4] += d; u[5] -= d
    pot = oracle.point_gravity((e, n, u), p, [2670], "potential")
    npt.assert_allclose(pot, pot[0])
    gz = oracle.point_gravity((e[4:], n[4:], u[4:]), p, [2670], "g_z")
    npt.assert_allclose(gz[0], -gz[1])
    assert gz[0] > 0  # mass below the point pulls downward


def test_point_spherical_closed_cases():
    """test_point_gravity.py:656-768."""
    mass = 2.0
    lon, lat, rad = np.meshgrid(np.linspace(-180, 180, 37), np.linspace(-90, 90, 19),
                                np.logspace(1, 8, 5))
    npt.assert_allclose(oracle.point_gravity((lon, lat, rad), [0, 0, 0], mass, "potential", "spherical"),
                        G * mass / rad)
    npt.assert_allclose(oracle.point_gravity((lon, lat, rad), [0, 0, 0], mass, "g_z", "spherical"),
                        G * mass / rad**2 * 1e5)
    radius = 3.0
    for lon_p in np.linspace(0, 350, 8):
        lons = np.array([l for l in np.linspace(0, 350, 36) if l != lon_p])
        res = oracle.point_gravity((lons, np.zeros_like(lons), np.full_like(lons, radius)),
                                   [lon_p, 0, radius], mass, "potential", "spherical")
        dist = 2 * radius * np.sin(0.5 * np.radians(np.abs(lons - lon_p)))
        npt.assert_allclose(res, G * mass / dist)
    for h in np.logspace(0, 4, 5):
        res = oracle.point_gravity(([30.0], [-45.0], [radius + h]), [30.0, -45.0, radius], mass,
                                   "g_z", "spherical")
        npt.assert_allclose(res, G * mass / h**2 * 1e5)


# --------------------------------------------------------------- prism layer
def _layer_prisms(easting, northing, bottom, top):
    E, N = np.meshgrid(easting, northing)
    he, hn = (easting[1] - easting[0]) / 2, (northing[1] - northing[0]) / 2
    return np.column_stack([(E - he).ravel(), (E + he).ravel(), (N - hn).ravel(), (N + hn).ravel(),
                            bottom.ravel(), top.ravel()])


def test_layer_matches_prisms_with_holes_and_threshold():
    """test_prism_layer.py:240-334,422-456 semantics on the oracle."""
    easting, northing = np.linspace(-1, 3, 5), np.linspace(7, 13, 4)
    top = np.arange(20, dtype=float).reshape(4, 5)
    bottom = np.zeros((4, 5))
    density = 2670.0 * np.ones((4, 5))
    coords = grid_coordinates((1, 3, 7, 10), spacing=1, extra=30.0)
    prisms = _layer_prisms(easting, northing, bottom, top)
    for field in oracle.FIELDS:
        lay = oracle.prism_layer_gravity(coords, easting, northing, bottom, top, density, field)
        ref = oracle.prism_gravity(coords, prisms, density.ravel(), field)
        npt.assert_allclose(lay, ref, rtol=1e-12, err_msg=field)
    # NaN density and NaN top are skipped
    d2, t2 = density.copy(), top.copy()
    d2[3, 3] = np.nan
    t2[2, 1] = np.nan
    keep = np.ones(20, bool)
    keep[[3 * 5 + 3, 2 * 5 + 1]] = False
    for field in ("potential", "g_z"):
        lay = oracle.prism_layer_gravity(coords, easting, northing, bottom, t2, d2, field)
        ref = oracle.prism_gravity(coords, prisms[keep], density.ravel()[keep], field)
        npt.assert_allclose(lay, ref, rtol=1e-12)
    # thickness threshold
    thick = (top - bottom).ravel() >= 5.0
    lay = oracle.prism_layer_gravity(coords, easting, northing, bottom, top, density, "g_z",
                                     thickness_threshold=5)
    ref = oracle.prism_gravity(coords, prisms[thick], density.ravel()[thick], "g_z")
    npt.assert_allclose(lay, ref, rtol=1e-12)


# ------------------------------------------------------- equivalent sources
def test_eqs_jacobian_symmetry_and_brute_force():
    """test_eq_sources_cartesian.py:293-318 plus a direct 1/r evaluation."""
    e, n = grid_coordinates((-100, 100, -100, 100), shape=(2, 2))
    coords = (e.ravel(), n.ravel(), np.zeros(4))
    pts = (e.ravel(), n.ravel(), np.full(4, 100.0))
    jac = oracle.eqs_jacobian(coords, pts)
    diag = np.diag_indices(4)
    npt.assert_allclose(jac[diag], jac[diag][0])
    anti = (diag[0], diag[1][::-1])
    npt.assert_allclose(jac[anti], jac[anti][0])
    rng = np.random.default_rng(3)
    c = tuple(rng.uniform(-1e3, 1e3, 17) for _ in range(3))
    p = tuple(rng.uniform(-1e3, 1e3, 11) for _ in range(2)) + (rng.uniform(-3e3, -2e3, 11),)
    direct = 1 / np.sqrt(sum((a[:, None] - b[None, :]) ** 2 for a, b in zip(c, p)))
    npt.assert_allclose(oracle.eqs_jacobian(c, p), direct, rtol=1e-15)
    coefs = rng.normal(size=11)
    npt.assert_allclose(oracle.eqs_predict(c, p, coefs), direct @ coefs, rtol=1e-13)


@pytest.mark.parametrize("damping", [None, 1e-3, 0.1])
@pytest.mark.parametrize("weighted", [False, True])
def test_eqs_fit_restatement_matches_sklearn(damping, weighted):
    """verde.least_squares semantics (cartesian.py:280) checked against
    scikit-learn's StandardScaler(with_mean=False) + Ridge/LinearRegression."""
    from sklearn.linear_model import LinearRegression, Ridge
    from sklearn.preprocessing import StandardScaler

    rng = np.random.default_rng(7)
    n_data, n_src = 300, 120
    coords = (rng.uniform(0, 5e3, n_data), rng.uniform(0, 5e3, n_data), rng.uniform(0, 100, n_data))
    pts = (rng.uniform(0, 5e3, n_src), rng.uniform(0, 5e3, n_src), np.full(n_src, -1500.0))
    J = oracle.eqs_jacobian(coords, pts)
    data = J @ rng.normal(size=n_src) + 1e-6 * rng.normal(size=n_data)
    w = rng.uniform(0.5, 2.0, n_data) if weighted else None
    scaler = StandardScaler(copy=True, with_mean=False, with_std=True)
    Js = scaler.fit_transform(J)
    regr = LinearRegression(fit_intercept=False) if damping is None else Ridge(alpha=damping, fit_intercept=False)
    regr.fit(Js, data, sample_weight=w)
    expected = regr.coef_ / scaler.scale_
    got = oracle.eqs_fit_coefs(J, data, w, damping)
    # compare predictions (the coefficient vector itself is ill-conditioned)
    npt.assert_allclose(J @ got, J @ expected, rtol=1e-6, atol=1e-9 * np.abs(J @ expected).max())
    if damping is not None:
        npt.assert_allclose(got, expected, rtol=1e-6, atol=1e-8 * np.abs(expected).max())


def _deep_sources_problem(depth, n_side=20, spacing=51.0, wide=False):
    """Sources `depth` below a 20 x 20 grid of 51 m spacing (one per datum,
    or two per datum when ``wide``); data from two deeper point masses."""
    e = np.arange(n_side) * spacing
    E, N = np.meshgrid(e, e)
    coords = (E.ravel(), N.ravel(), np.zeros(E.size))
    if wide:
        pts = (np.concatenate([coords[0], coords[0] + spacing / 3]),
               np.concatenate([coords[1], coords[1] + spacing / 4]),
               np.concatenate([coords[2] - depth, coords[2] - 1.3 * depth]))
    else:
        pts = (coords[0], coords[1], coords[2] - depth)
    masses = (np.array([300.0, 700.0]), np.array([500.0, 200.0]), np.array([-800.0, -600.0]))
    data = oracle.eqs_jacobian(coords, masses) @ np.array([1e6, -2e6])
    return oracle.eqs_jacobian(coords, pts), data


@pytest.mark.parametrize("depth,damping,route", [
    (200.0, 1e-18, "cholesky"), (500.0, 1e-18, "svd"), (1000.0, 1e-18, "svd"),
    (500.0, None, "lstsq"), (1000.0, None, "lstsq"), (1000.0, 1e-3, "cholesky"),
])
def test_eqs_fit_restatement_singular_matches_sklearn(depth, damping, route):
    """Near-singular scaled jacobians (condition numbers 2e8 to 1e20): the
    restatement takes scikit-learn's route (Cholesky, the SVD fallback of
    Ridge when the Cholesky solve fails, lstsq's cut-off for
    LinearRegression) and matches its coefficients."""
    import warnings

    from sklearn.linear_model import LinearRegression, Ridge
    from sklearn.preprocessing import StandardScaler

    J, data = _deep_sources_problem(depth)
    got, taken = oracle.eqs_fit_coefs_info(J, data, None, damping)
    assert taken == route
    scaler = StandardScaler(copy=True, with_mean=False, with_std=True)
    Js = scaler.fit_transform(J)
    regr = LinearRegression(fit_intercept=False) if damping is None else Ridge(alpha=damping, fit_intercept=False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        regr.fit(Js, data)
    expected = regr.coef_ / scaler.scale_
    npt.assert_allclose(J @ got, J @ expected, rtol=0, atol=1e-10 * np.abs(J @ expected).max())
    npt.assert_allclose(got, expected, rtol=0, atol=1e-8 * np.abs(expected).max())


@pytest.mark.parametrize("damping", [None, 1e-18, 1e-3])
def test_eqs_fit_restatement_wide_matches_sklearn(damping):
    """More sources than data (m > n): Ridge's dual form (and its lstsq
    fallback) and LinearRegression's minimum-norm solution."""
    import warnings

    from sklearn.linear_model import LinearRegression, Ridge
    from sklearn.preprocessing import StandardScaler

    J, data = _deep_sources_problem(400.0, n_side=12, wide=True)
    assert J.shape[1] > J.shape[0]
    got, taken = oracle.eqs_fit_coefs_info(J, data, None, damping)
    assert taken in (("lstsq",) if damping is None else ("dual", "dual-lstsq"))
    scaler = StandardScaler(copy=True, with_mean=False, with_std=True)
    Js = scaler.fit_transform(J)
    regr = LinearRegression(fit_intercept=False) if damping is None else Ridge(alpha=damping, fit_intercept=False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        regr.fit(Js, data)
    expected = regr.coef_ / scaler.scale_
    npt.assert_allclose(J @ got, J @ expected, rtol=0, atol=1e-10 * np.abs(J @ expected).max())
    npt.assert_allclose(got, expected, rtol=0, atol=1e-8 * np.abs(expected).max())
