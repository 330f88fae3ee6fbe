"""
Pin the CPU oracle's prism magnetic field (oracle/oracle.c,
or_prism_magnetic_pair; the reference calls choclo's prism magnetic
functions at src/harmonica/_forward/prisms/magnetic.py:319-331,384-397)
against checks that do not restate its closed form:

* Gauss-Legendre volume quadrature of the dipole field
  (mu_0 / 4 pi) [3 (M . d) d / |d|^5 - M / |d|^3] over the prism;
* the far field of the equivalent dipole (moment M V);
* the demagnetising field at the centre of a cube, -mu_0 M / 3 (the cube's
  demagnetising factors are 1/3 by symmetry);
* curl- and divergence-free outside the prism (finite differences);
* superposition (a prism cut in two sums to the whole), reversal of M,
  and the exchange of the east and north axes;
* the relation B_i = (c_m / (G rho)) sum_j M_j g_ij to the gravity tensor
  of the (separately pinned) gravity oracle;
* NaN on the edges and vertices, finite on faces and edge extensions;
* the reference's discarding of null prisms (magnetic.py:403-440).

Parity unpinned (choclo is not in this image, and the reference's tests
only compare against choclo): the value of mu_0 (4 pi 1e-7 H/m here; a
CODATA measured value would differ by up to 5.5e-10 relative), the NaN set
(every edge and vertex, for all three components) and the interior
convention (the closed form gives mu_0 H, the field of the magnetisation's
surface charges, without mu_0 M).
"""

import numpy as np
import numpy.testing as npt
import pytest

import oracle

MU0 = 4e-7 * np.pi
CM = MU0 / 4 / np.pi
G = 6.6743e-11
PRISM = [-120.0, 80.0, -50.0, 170.0, -300.0, -40.0]
MAGS = [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (2.5, -1.3, 0.7)]
QUAD_POINTS = [
    (0.0, 0.0, 100.0),
    (300.0, -250.0, 60.0),
    (-400.0, 350.0, -170.0),
    (-30.0, 420.0, -500.0),
    (250.0, 60.0, -600.0),
    (-350.0, -200.0, 200.0),
    (40.0, -260.0, -150.0),
]


def _axis(a, b, nsub, nq):
    x, w = np.polynomial.legendre.leggauss(nq)
    edges = np.linspace(a, b, nsub + 1)
    nodes = [0.5 * (hi - lo) * x + 0.5 * (hi + lo) for lo, hi in zip(edges[:-1], edges[1:])]
    weights = [0.5 * (hi - lo) * w for lo, hi in zip(edges[:-1], edges[1:])]
    return np.concatenate(nodes), np.concatenate(weights)


def quadrature_b(point, prism, mag, nsub=6, nq=16):
    """(b_e, b_n, b_u) in nT by quadrature of the dipole field density."""
    ex, ew = _axis(prism[0], prism[1], nsub, nq)
    nx, nw = _axis(prism[2], prism[3], nsub, nq)
    ux, uw = _axis(prism[4], prism[5], nsub, nq)
    E, N, U = np.meshgrid(ex, nx, ux, indexing="ij")
    W = ew[:, None, None] * nw[None, :, None] * uw[None, None, :]
    d = (point[0] - E, point[1] - N, point[2] - U)
    l2 = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
    l3 = l2 * np.sqrt(l2)
    l5 = l3 * l2
    mdotd = mag[0] * d[0] + mag[1] * d[1] + mag[2] * d[2]
    return np.array([1e9 * CM * np.sum(W * (3 * mdotd * d[c] / l5 - mag[c] / l3))
                     for c in range(3)])


def b_at(point, prism, mag, field="b"):
    out = oracle.prism_magnetic(tuple([c] for c in point), prism,
                                tuple([m] for m in mag), field)
    return np.array([o[0] for o in out]) if field == "b" else out[0]


def test_constant():
    """c_m = mu_0 / (4 pi) with mu_0 = 4 pi 1e-7 H/m."""
    b = oracle.prism_magnetic_pair((0.0, 0.0, 1e4), PRISM, (0.0, 0.0, 1.0))
    vol = 200.0 * 220.0 * 260.0
    r = 1e4 - (-170.0)
    npt.assert_allclose(b[2], 1e-7 * 2 * vol / r**3, rtol=1e-3)


@pytest.mark.parametrize("point", QUAD_POINTS)
@pytest.mark.parametrize("mag", MAGS)
def test_against_volume_quadrature(point, mag):
    expected = quadrature_b(point, PRISM, mag)
    got = b_at(point, PRISM, mag)
    scale = np.linalg.norm(expected)
    npt.assert_allclose(got, expected, rtol=0, atol=1e-10 * scale)


@pytest.mark.parametrize("mag", MAGS)
def test_far_field_dipole(mag):
    """At 200 prism sizes the field is the dipole's to O((a / r)^2)."""
    centre = np.array([(PRISM[0] + PRISM[1]) / 2, (PRISM[2] + PRISM[3]) / 2,
                       (PRISM[4] + PRISM[5]) / 2])
    vol = (PRISM[1] - PRISM[0]) * (PRISM[3] - PRISM[2]) * (PRISM[5] - PRISM[4])
    m = np.array(mag) * vol
    for direction in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, -2, 3)):
        d = np.array(direction, dtype=float)
        d *= 5e4 / np.linalg.norm(d)
        p = centre + d
        r = np.linalg.norm(d)
        dip = 1e9 * CM * (3 * np.dot(m, d) * d / r**5 - m / r**3)
        got = b_at(p, PRISM, mag)
        npt.assert_allclose(got, dip, rtol=0, atol=1e-4 * np.linalg.norm(dip))


@pytest.mark.parametrize("mag", MAGS)
def test_cube_centre_demagnetising_field(mag):
    cube = [-50.0, 50.0, -50.0, 50.0, -100.0, 0.0]
    got = b_at((0.0, 0.0, -50.0), cube, mag)
    npt.assert_allclose(got, -1e9 * MU0 / 3 * np.array(mag), rtol=0,
                        atol=1e-12 * 1e9 * MU0 * np.linalg.norm(mag))


@pytest.mark.parametrize("point", [(200.0, 30.0, 0.0), (-20.0, 260.0, -120.0),
                                   (10.0, 20.0, 50.0), (-300.0, -200.0, -500.0)])
def test_curl_and_divergence_free_outside(point):
    mag = (1.7, -0.6, 1.1)
    h = 1e-2
    grad = np.zeros((3, 3))  # grad[c, k] = d b_c / d x_k
    for k in range(3):
        dp = np.zeros(3)
        dp[k] = h
        grad[:, k] = (b_at(np.array(point) + dp, PRISM, mag) -
                      b_at(np.array(point) - dp, PRISM, mag)) / (2 * h)
    scale = np.abs(grad).max()
    npt.assert_allclose(grad, grad.T, rtol=0, atol=1e-6 * scale)
    assert abs(np.trace(grad)) < 1e-6 * scale


def test_superposition_and_symmetries():
    rng = np.random.default_rng(3)
    points = rng.uniform(-600, 600, size=(40, 3))
    mag = (0.8, -1.9, 2.2)
    whole = np.array(oracle.prism_magnetic(tuple(points.T), PRISM, tuple([m] for m in mag), "b"))
    w, e, s, n, b, t = PRISM
    for cut in ([[w, -10, s, n, b, t], [-10, e, s, n, b, t]],
                [[w, e, s, 33, b, t], [w, e, 33, n, b, t]],
                [[w, e, s, n, b, -123], [w, e, s, n, -123, t]]):
        parts = np.array(oracle.prism_magnetic(tuple(points.T), cut,
                                               tuple([m, m] for m in mag), "b"))
        npt.assert_allclose(parts, whole, rtol=0, atol=1e-11 * np.abs(whole).max())
    # reversing the magnetisation reverses the field
    neg = np.array(oracle.prism_magnetic(tuple(points.T), PRISM,
                                         tuple([-m] for m in mag), "b"))
    npt.assert_array_equal(neg, -whole)
    # exchanging east and north (prism, point and M) exchanges b_e and b_n;
    # the exchange mirrors the frame, so b_u keeps its sign
    swapped_prism = [s, n, w, e, b, t]
    sw = np.array(oracle.prism_magnetic((points[:, 1], points[:, 0], points[:, 2]),
                                        swapped_prism, ([mag[1]], [mag[0]], [mag[2]]), "b"))
    npt.assert_allclose(sw[[1, 0, 2]], whole, rtol=0, atol=1e-11 * np.abs(whole).max())


def test_relation_to_gravity_tensor():
    """B_i = c_m / (G rho) sum_j M_j T_ij with T the tensor of the potential
    in the east-north-up frame (harmonica's g_ez, g_nz point down)."""
    rho = 1000.0
    mag = np.array([1.2, 0.4, -2.0])
    for point in QUAD_POINTS:
        c = tuple([v] for v in point)
        g = {f: oracle.prism_gravity(c, PRISM, rho, f)[0] * 1e-9 for f in
             ("g_ee", "g_nn", "g_zz", "g_en", "g_ez", "g_nz")}
        T = np.array([[g["g_ee"], g["g_en"], -g["g_ez"]],
                      [g["g_en"], g["g_nn"], -g["g_nz"]],
                      [-g["g_ez"], -g["g_nz"], g["g_zz"]]])
        expected = 1e9 * CM / (G * rho) * T @ mag
        got = b_at(point, PRISM, mag)
        npt.assert_allclose(got, expected, rtol=0, atol=1e-12 * np.abs(expected).max())


def test_nan_on_edges_and_vertices_only():
    w, e, s, n, b, t = PRISM
    mag = (1.0, 2.0, 3.0)
    on_edges = [
        (0.0, s, b), (10.0, n, t), (w, 0.0, t), (e, 100.0, b),  # easting / northing edges
        (w, s, -100.0), (e, n, -200.0),                         # upward edges
        (w, s, b), (e, n, t), (w, n, t),                        # vertices
    ]
    for p in on_edges:
        got = b_at(p, PRISM, mag)
        assert np.all(np.isnan(got)), p
    finite = [
        (0.0, 0.0, t), (w, 0.0, -100.0), (0.0, s, -100.0),      # faces
        (w - 50.0, s, b), (w, s, t + 40.0), (e, n + 10.0, t),   # edge-line extensions
        (w + 1e-6, s + 1e-6, -100.0),                           # next to an edge
    ]
    for p in finite:
        assert np.all(np.isfinite(b_at(p, PRISM, mag))), p


def test_null_prisms_discarded():
    """A zero-volume or unmagnetised prism is dropped before the loop
    (magnetic.py:432-436), so a point on its edges stays finite."""
    prisms = [PRISM, [0.0, 0.0, 0.0, 10.0, 0.0, 10.0], [500.0, 600.0, 0.0, 10.0, 0.0, 10.0]]
    # east, north and up components of the three prisms' magnetisations:
    # (1, 0, 1), (1, 1, 1) and (0, 0, 0)
    mags = ([1.0, 1.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.0])
    for point in (([600.0], [10.0], [5.0]),   # on an edge of the unmagnetised prism
                  ([0.0], [5.0], [10.0])):    # on an edge of the flat prism
        got = oracle.prism_magnetic(point, prisms, mags, "b")
        alone = oracle.prism_magnetic(point, [PRISM], ([1.0], [0.0], [1.0]), "b")
        for g, a in zip(got, alone):
            assert np.isfinite(g[0])
            npt.assert_array_equal(g, a)


def test_single_components_match_vector():
    rng = np.random.default_rng(7)
    points = tuple(rng.uniform(-500, 500, size=(3, 30)))
    mags = (rng.normal(size=2), rng.normal(size=2), rng.normal(size=2))
    prisms = [PRISM, [200.0, 260.0, -40.0, 30.0, -80.0, -10.0]]
    vec = oracle.prism_magnetic(points, prisms, mags, "b")
    for c, field in enumerate(("b_e", "b_n", "b_u")):
        npt.assert_array_equal(oracle.prism_magnetic(points, prisms, mags, field), vec[c])
