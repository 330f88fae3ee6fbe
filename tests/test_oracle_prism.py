"""
Pin the CPU oracle's prism closed forms (oracle/oracle.c) against checks that
do not restate its formulas:

* the reference's own doctest values (gravity.py:182-193);
* Gauss-Legendre volume quadrature of the point-mass kernels, with the
  harmonica sign and unit conventions written from the physics
  (gravity.py:67-86 docstring: g_z and the z-tensor terms point downward);
* Laplace outside and Poisson inside the prism;
* finite differences of the potential and of the accelerations;
* the infinite-slab limit (test_prism.py:269-297, _gravity_corrections.py:77);
* superposition (a prism cut in pieces sums to the whole);
* continuity across the degenerate log/atan branches (edge-line extensions
  and face planes), one-sided limits on faces (gravity.py:153-158);
* the singular-point contract (gravity.py:141-152; test_prism.py:380-502);
* the inverted-prism sign (test_prism.py:139-158).
"""

import numpy as np
import numpy.testing as npt
import pytest

import oracle

G = 6.6743e-11
FIELDS = oracle.FIELDS
TENSORS = ("g_ee", "g_nn", "g_zz", "g_en", "g_ez", "g_nz")


def test_doctest_values_single_prism():
    """gravity.py:182-184."""
    coords = ([-40, 0, 40], [0, 0, 0], [30, 30, 30])
    gz = oracle.prism_gravity(coords, [-34, 5, -18, 14, -345, -146], 2670, "g_z")
    assert "({:.5f}, {:.5f}, {:.5f})".format(*gz) == "(0.06552, 0.06629, 0.06174)"


def test_doctest_values_two_prisms():
    """gravity.py:190-193."""
    coords = ([-40, 0, 40], [0, 0, 0], [30, 30, 30])
    prisms = [[-134, -5, -45, 45, -200, -50], [5, 134, -45, 45, -180, -30]]
    gz = oracle.prism_gravity(coords, prisms, [-300, 300], "g_z")
    assert "({:.5f}, {:.5f}, {:.5f})".format(*gz) == "(-0.05380, 0.02908, 0.11237)"


def _axis(a, b, nsub, nq):
    x, w = np.polynomial.legendre.leggauss(nq)
    edges = np.linspace(a, b, nsub + 1)
    nodes = [0.5 * (hi - lo) * x + 0.5 * (hi + lo) for lo, hi in zip(edges[:-1], edges[1:])]
    weights = [0.5 * (hi - lo) * w for lo, hi in zip(edges[:-1], edges[1:])]
    return np.concatenate(nodes), np.concatenate(weights)


def quadrature_fields(point, prism, rho, nsub=6, nq=16):
    """All ten fields of a homogeneous prism by tensor Gauss-Legendre
    quadrature of the point-mass kernels, in harmonica's units and signs."""
    ex, ew = _axis(prism[0], prism[1], nsub, nq)
    nx, nw = _axis(prism[2], prism[3], nsub, nq)
    ux, uw = _axis(prism[4], prism[5], nsub, nq)
    E, N, U = np.meshgrid(ex, nx, ux, indexing="ij")
    W = ew[:, None, None] * nw[None, :, None] * uw[None, None, :]
    de, dn, du = point[0] - E, point[1] - N, point[2] - U
    l2 = de**2 + dn**2 + du**2
    l = np.sqrt(l2)
    l3, l5 = l2 * l, l2 * l2 * l
    k = G * rho
    # V = G rho int 1/l; accelerations are grad V; g_z is the downward one.
    return {
        "potential": k * np.sum(W / l),
        "g_e": 1e5 * k * np.sum(W * -de / l3),
        "g_n": 1e5 * k * np.sum(W * -dn / l3),
        "g_z": 1e5 * k * np.sum(W * du / l3),
        "g_ee": 1e9 * k * np.sum(W * (3 * de**2 - l2) / l5),
        "g_nn": 1e9 * k * np.sum(W * (3 * dn**2 - l2) / l5),
        "g_zz": 1e9 * k * np.sum(W * (3 * du**2 - l2) / l5),
        "g_en": 1e9 * k * np.sum(W * 3 * de * dn / l5),
        "g_ez": 1e9 * k * np.sum(W * -3 * de * du / l5),
        "g_nz": 1e9 * k * np.sum(W * -3 * dn * du / l5),
    }


PRISM = np.array([-120.0, 80.0, -50.0, 170.0, -300.0, -40.0])
# Points in every octant around the prism and straddling its bounds, all at
# least ~100 m from its surface so the quadrature converges to ~1e-12.
QUAD_POINTS = [
    (0.0, 0.0, 100.0),
    (300.0, -250.0, 60.0),
    (-400.0, 350.0, -170.0),
    (-30.0, 420.0, -500.0),
    (250.0, 60.0, -600.0),
    (-350.0, -200.0, 200.0),
    (40.0, -260.0, -150.0),
]


@pytest.mark.parametrize("point", QUAD_POINTS)
def test_against_volume_quadrature(point):
    expected = quadrature_fields(point, PRISM, 2670.0)
    # Scale of each family at this point: |potential|, |g|, |tensor|.
    acc = np.sqrt(sum(expected[f] ** 2 for f in ("g_e", "g_n", "g_z")))
    ten = np.sqrt(sum(expected[f] ** 2 for f in TENSORS))
    for field in FIELDS:
        scale = abs(expected[field]) if field == "potential" else (acc if field in ("g_e", "g_n", "g_z") else ten)
        got = oracle.prism_gravity(tuple([c] for c in point), PRISM, 2670.0, field)[0]
        npt.assert_allclose(got, expected[field], rtol=0, atol=1e-10 * scale, err_msg=field)


def test_laplace_outside_poisson_inside():
    """Trace of the tensor: 0 outside; -4 pi G rho (in Eotvos) inside."""
    rho = 2900.0
    e, n = np.meshgrid(np.linspace(-500, 500, 9), np.linspace(-400, 450, 8))
    outside = (e, n, np.full_like(e, 90.0))
    trace = sum(oracle.prism_gravity(outside, PRISM, rho, f) for f in ("g_ee", "g_nn", "g_zz"))
    scale = np.abs(oracle.prism_gravity(outside, PRISM, rho, "g_zz")).max()
    npt.assert_allclose(trace, 0.0, atol=1e-12 * scale)
    inside = ([-100.0, 0.0, 75.0, 10.0], [-40.0, 60.0, 160.0, 0.0], [-290.0, -200.0, -50.0, -123.0])
    trace = sum(oracle.prism_gravity(inside, PRISM, rho, f) for f in ("g_ee", "g_nn", "g_zz"))
    npt.assert_allclose(trace, -4 * np.pi * G * rho * 1e9, rtol=1e-10)


@pytest.mark.parametrize("point", [(10.0, 20.0, 50.0), (-500.0, 330.0, -100.0), (0.0, 0.0, -170.0)])
def test_finite_differences(point):
    """Accelerations from the potential; tensor terms from accelerations."""
    rho, h = 2670.0, 1e-2
    axis = {"e": 0, "n": 1, "z": 2}

    def at(field, shift):
        p = np.array(point, dtype=float) + shift
        return oracle.prism_gravity(tuple([c] for c in p), PRISM, rho, field)[0]

    def d(field, direction):
        s = np.zeros(3)
        s[axis[direction]] = h
        return (at(field, s) - at(field, -s)) / (2 * h)

    # g_e, g_n are +dV/de, dV/dn; g_z is -dV/du (downward); mGal.
    gscale = sum(abs(at(f, 0)) for f in ("g_e", "g_n", "g_z"))
    npt.assert_allclose(at("g_e", 0), 1e5 * d("potential", "e"), rtol=1e-6, atol=1e-8 * gscale)
    npt.assert_allclose(at("g_n", 0), 1e5 * d("potential", "n"), rtol=1e-6, atol=1e-8 * gscale)
    npt.assert_allclose(at("g_z", 0), -1e5 * d("potential", "z"), rtol=1e-6, atol=1e-8 * gscale)
    scale = abs(at("g_zz", 0)) + abs(at("g_ee", 0))
    # Tensor in Eotvos from mGal accelerations; d/dz = -d/du (z is downward).
    for field, (acc, direction, sign) in {
        "g_ee": ("g_e", "e", 1), "g_nn": ("g_n", "n", 1), "g_zz": ("g_z", "z", -1),
        "g_en": ("g_e", "n", 1), "g_ez": ("g_e", "z", -1), "g_nz": ("g_n", "z", -1),
    }.items():
        npt.assert_allclose(at(field, 0), sign * 1e4 * d(acc, direction), rtol=1e-5,
                            atol=1e-6 * scale, err_msg=field)


def test_infinite_slab_limit():
    """test_prism.py:269-297: g_z converges to 2 pi G rho h (mGal)."""
    height, thickness, rho = 1.5, 10.5, 2670.0
    sizes = np.logspace(3, 9, 7)
    results = np.array([
        oracle.prism_gravity((0, 0, height), [-s / 2, s / 2, -s / 2, s / 2, height - thickness, height],
                             rho, "g_z")
        for s in sizes
    ]).ravel()
    analytical = 1e5 * 2 * np.pi * G * rho * thickness
    errors = np.abs(analytical - results)
    assert (errors[1:] < errors[:-1]).all()
    npt.assert_allclose(results[-1], analytical, rtol=1e-7)


def test_superposition():
    """A prism cut along all three axes sums to the whole, at points above
    the cut lines (zero shifted coordinates for the inner pieces)."""
    w, e, s, n, b, t = PRISM
    em, nm, um = -20.0, 60.0, -170.0
    pieces = [
        [x0, x1, y0, y1, z0, z1]
        for x0, x1 in ((w, em), (em, e))
        for y0, y1 in ((s, nm), (nm, n))
        for z0, z1 in ((b, um), (um, t))
    ]
    coords = ([em, 0.0, em, 300.0], [nm, nm, 0.0, -90.0], [100.0, 50.0, 10.0, -170.0])
    for field in FIELDS:
        whole = oracle.prism_gravity(coords, PRISM, 2000.0, field)
        parts = oracle.prism_gravity(coords, pieces, [2000.0] * 8, field)
        scale = np.abs(whole).max()
        npt.assert_allclose(parts, whole, atol=1e-11 * scale, err_msg=field)


def test_inverted_prism_sign():
    """test_prism.py:139-158: swapping bottom and top negates the field."""
    coords = ([1.0, 3.0], [2.0, -7.0], [5.0, 25.0])
    for field in FIELDS:
        good = oracle.prism_gravity(coords, [0, 10, 0, 10, -10, 0], 100, field)
        bad = oracle.prism_gravity(coords, [0, 10, 0, 10, 0, -10], 100, field)
        npt.assert_allclose(bad, -good, rtol=1e-10, err_msg=field)


def test_continuity_across_degenerate_branches():
    """On edge-line extensions and face planes (some shifted coordinates
    exactly zero) every non-singular field equals the limit of nearby points."""
    w, e, s, n, b, t = PRISM
    cases = [
        (w, s, 200.0),      # above a vertical edge: x = y = 0 for a vertex column
        (e, 0.0, 150.0),    # above the east face plane: x = 0
        (0.0, n, -700.0),   # below the north face plane
        (w - 300.0, s, t),  # on the line of a horizontal edge (y = z = 0)
        (e + 250.0, n, b),  # idem, other edge
    ]
    # On the line extension of an edge, outside the prism, one vertex log of
    # the non-diagonal tensor terms degenerates (ln 0) and the restated
    # convention sets it to 0; that term has no vanishing coefficient, so those
    # (point, field) pairs are not continuous.  The reference's choice there
    # comes from choclo and cannot be checked here: parity unpinned on that
    # measure-zero set (the HIP kernels follow the oracle's convention).
    skip = {0: ("g_en",), 4: ("g_nz",)}
    eps = 1e-6
    for ic, point in enumerate(cases):
        for field in FIELDS:
            if field in skip.get(ic, ()):
                continue
            p = np.array(point)
            v0 = oracle.prism_gravity(tuple([c] for c in p), PRISM, 2670.0, field)[0]
            nearby = []
            for shift in ((eps, eps, eps), (-eps, -eps, eps), (eps, -eps, -eps)):
                q = p + np.array(shift)
                nearby.append(oracle.prism_gravity(tuple([c] for c in q), PRISM, 2670.0, field)[0])
            scale = max(abs(v0), abs(oracle.prism_gravity(tuple([c] for c in p), PRISM, 2670.0, "g_zz")[0]))
            assert np.isfinite(v0), (point, field)
            npt.assert_allclose(nearby, v0, atol=1e-4 * scale, err_msg=f"{point} {field}")


def test_face_values_are_one_sided_limits():
    """Diagonal tensor terms on the faces normal to their direction
    (gravity.py:153-158) take a one-sided limit, and the two limits differ by
    the Poisson jump 4 pi G rho.  With shifted coordinates (bound - obs) and
    atan(a / 0) := sign(a) pi / 2, the value is the outside limit on the west,
    south and bottom faces and the inside limit on the east, north and top
    faces.  The reference docstring says "outside" for every face; which side
    choclo returns on the east/north/top faces cannot be checked here (choclo
    is absent), so on that measure-zero set parity is unpinned."""
    w, e, s, n, b, t = PRISM
    rho = 2670.0
    jump = 4 * np.pi * G * rho * 1e9
    eps = 1e-7
    faces = [
        ("g_zz", (0.0, 0.0, t), (0, 0, 1), "inside"),
        ("g_zz", (0.0, 0.0, b), (0, 0, -1), "outside"),
        ("g_ee", (e, 10.0, -100.0), (1, 0, 0), "inside"),
        ("g_ee", (w, 10.0, -100.0), (-1, 0, 0), "outside"),
        ("g_nn", (-30.0, n, -200.0), (0, 1, 0), "inside"),
        ("g_nn", (-30.0, s, -200.0), (0, -1, 0), "outside"),
    ]
    for field, point, normal, side in faces:
        p = np.array(point)
        on = oracle.prism_gravity(tuple([c] for c in p), PRISM, rho, field)[0]
        out = oracle.prism_gravity(tuple([c] for c in p + eps * np.array(normal)), PRISM, rho, field)[0]
        inn = oracle.prism_gravity(tuple([c] for c in p - eps * np.array(normal)), PRISM, rho, field)[0]
        npt.assert_allclose(out - inn, jump, rtol=1e-4, err_msg=field)
        npt.assert_allclose(on, out if side == "outside" else inn, atol=1e-5 * jump, err_msg=field)


SAMPLE = np.array([-10.3, 5.4, 8.6, 14.3, -30.3, 2.4])  # test_prism.py:387-390


def _vertices(p):
    return [c.ravel() for c in np.meshgrid(p[:2], p[2:4], p[4:6])]


def _edge_centres(p, direction):
    if direction == "e":
        n, u = (c.ravel() for c in np.meshgrid(p[2:4], p[4:6]))
        return np.full_like(n, (p[0] + p[1]) / 2), n, u
    if direction == "n":
        e, u = (c.ravel() for c in np.meshgrid(p[0:2], p[4:6]))
        return e, np.full_like(e, (p[2] + p[3]) / 2), u
    e, n = (c.ravel() for c in np.meshgrid(p[0:2], p[2:4]))
    return e, n, np.full_like(e, (p[4] + p[5]) / 2)


# Edge families on which each tensor component is singular (gravity.py:141-152).
SINGULAR_EDGES = {
    "g_ee": ("n", "z"), "g_nn": ("e", "z"), "g_zz": ("e", "n"),
    "g_en": ("z",), "g_ez": ("n",), "g_nz": ("e",),
}


def test_singular_contract():
    """NaN on every vertex for tensor terms, NaN on the edges of the
    component's singular families only, finite elsewhere and always finite
    for the potential and accelerations."""
    v = _vertices(SAMPLE)
    for field in FIELDS:
        res = oracle.prism_gravity(v, SAMPLE, 2900.0, field)
        assert np.isnan(res).all() == (field in TENSORS), field
        assert oracle.prism_any_singular(v, SAMPLE, field) == (field in TENSORS)
    for direction in ("e", "n", "z"):
        pts = _edge_centres(SAMPLE, direction)
        for field in FIELDS:
            singular = direction in SINGULAR_EDGES.get(field, ())
            res = oracle.prism_gravity(pts, SAMPLE, 2900.0, field)
            assert np.isnan(res).all() == singular, (field, direction)
            assert not singular == np.isfinite(res).all()
            assert oracle.prism_any_singular(pts, SAMPLE, field) == singular


def test_far_field_point_mass():
    """Far away the prism acts as a point mass of rho * volume.  At 100-400 km
    the closed form keeps only ~5 digits (the 8 vertex terms cancel), so the
    tolerance is 1e-4."""
    w, e, s, n, b, t = PRISM
    mass = 1000.0 * (e - w) * (n - s) * (t - b)
    centre = ((w + e) / 2, (s + n) / 2, (b + t) / 2)
    far = (np.array([3e5, -2e5]), np.array([1e5, 4e5]), np.array([5e4, -1e5]))
    for field in FIELDS:
        prism = oracle.prism_gravity(far, PRISM, 1000.0, field)
        point = oracle.point_gravity(far, centre, [mass], field)
        npt.assert_allclose(prism, point, rtol=1e-4, err_msg=field)


def test_dgu_dtop_finite_difference():
    """d(g_u)/d(top) closed form against central differences of the pair."""
    rho = -300.0
    for point in [(10.0, -5.0, 500.0), (300.0, 200.0, 0.0), (-50.0, 400.0, -20.0)]:
        p = PRISM.copy()
        h = 1e-3
        p_up, p_dn = p.copy(), p.copy()
        p_up[5] += h
        p_dn[5] -= h
        fd = (oracle.prism_pair("g_z", point, p_up, rho) - oracle.prism_pair("g_z", point, p_dn, rho)) / (2 * h)
        npt.assert_allclose(oracle.prism_dgu_dtop(point, p, rho), fd, rtol=1e-6)
