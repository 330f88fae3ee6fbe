"""
GPU parity of the equivalent-source path: jacobian and predict against the
CPU oracle, the GPU least-squares fit against the verde/sklearn restatement
(oracle.eqs_fit_coefs, itself pinned against scikit-learn), the
reference's accuracy tests (test_eq_sources_cartesian.py:111-212,255-404),
and the mixed-precision fit against the FP64 fit.
"""

import numpy as np
import numpy.testing as npt
import pytest

import oracle
from conftest import grid_coordinates

pytestmark = pytest.mark.gpu
REGION = (-3e3, -1e3, 5e3, 7e3)


@pytest.fixture(scope="module")
def hm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import harmonica_amd

    return harmonica_amd


def checkerboard(amplitude, region, easting, northing):
    """Synthetic checkerboard (verde.synthetic.CheckerBoard): wavelengths of
    half the region size in each direction."""
    w_e = (region[1] - region[0]) / 2
    w_n = (region[3] - region[2]) / 2
    return amplitude * np.sin(2 * np.pi * easting / w_e) * np.cos(2 * np.pi * northing / w_n)


@pytest.fixture(scope="module")
def synthetic():
    points = grid_coordinates(REGION, shape=(6, 6), extra=-1e3)
    masses = checkerboard(1e13, REGION, points[0], points[1])
    coords = grid_coordinates(REGION, shape=(40, 40), extra=0)
    data = oracle.point_gravity(coords, points, masses, "g_z")
    return points, masses, coords, data


def _scattered(seed, n_data, n_src, depth=1500.0):
    rng = np.random.default_rng(seed)
    coords = (rng.uniform(0, 5e3, n_data), rng.uniform(0, 5e3, n_data), rng.uniform(0, 100, n_data))
    pts = (rng.uniform(0, 5e3, n_src), rng.uniform(0, 5e3, n_src), np.full(n_src, -depth))
    return coords, pts, rng


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_jacobian_vs_oracle(hm, dtype):
    coords, pts, _ = _scattered(0, 333, 211)
    eqs = hm.EquivalentSources(dtype=dtype)
    jac = eqs.jacobian(coords, pts)
    assert jac.dtype == np.dtype(dtype)
    rtol = 1e-15 if dtype == "float64" else 1e-6
    npt.assert_allclose(jac, oracle.eqs_jacobian(coords, pts), rtol=rtol)


def test_jacobian_symmetry(hm):
    """test_eq_sources_cartesian.py:293-318."""
    e, n = grid_coordinates((-100, 100, -100, 100), shape=(2, 2))
    coords = (e.ravel(), n.ravel(), np.zeros(4))
    pts = (e.ravel(), n.ravel(), np.full(4, 100.0))
    jac = hm.EquivalentSources().jacobian(coords, pts)
    diag = np.diag_indices(4)
    npt.assert_allclose(jac[diag], jac[diag][0])
    anti = (diag[0], diag[1][::-1])
    npt.assert_allclose(jac[anti], jac[anti][0])
    nn = np.ones((4, 4), bool)
    nn[diag] = False
    nn[anti] = False
    npt.assert_allclose(jac[nn], jac[nn][0])


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_predict_vs_oracle(hm, dtype):
    from harmonica_amd._eqs import predict_device
    import torch

    coords, pts, rng = _scattered(1, 1000, 700)
    coefs = rng.normal(size=700)
    tdtype = torch.float64 if dtype == "float64" else torch.float32
    got = predict_device(tuple(c.astype(dtype) for c in coords), tuple(p.astype(dtype) for p in pts),
                         coefs.astype(dtype), dtype=tdtype).cpu().numpy()
    expected = oracle.eqs_predict(coords, pts, coefs)
    tol = 1e-12 if dtype == "float64" else 1e-5
    npt.assert_allclose(got, expected, atol=tol * np.abs(expected).max())


@pytest.mark.parametrize("damping", [None, 1e-3, 0.1])
@pytest.mark.parametrize("weighted", [False, True])
def test_fit_fp64_vs_least_squares_restatement(hm, damping, weighted):
    """GPU fit (scaled columns, weighted ridge normal equations, FP64 MFMA
    normal matrix and Cholesky) against the restatement of verde's
    least_squares checked against scikit-learn."""
    from harmonica_amd._eqs import fit_device

    coords, pts, rng = _scattered(2, 700, 300)
    m = pts[0].size
    J = oracle.eqs_jacobian(coords, pts)
    data = J @ rng.normal(size=m) + 1e-3 * rng.normal(size=700)
    w = rng.uniform(0.5, 2.0, 700) if weighted else None
    coefs, info = fit_device(coords, data, w, pts, damping)
    expected = oracle.eqs_fit_coefs(J, data, w, damping)
    pred, ref = J @ coefs, J @ expected
    if damping is None:
        # LinearRegression on a scaled jacobian of condition number 1.8e11
        # (the SVD route here, LAPACK's gelsd in the oracle): two backward-
        # stable solves differ in the fitted values by up to about
        # 2^-52 cond |residual| (~1e-6 of the data here)
        assert info["svd"] == 1.0
        npt.assert_allclose(pred, ref, atol=1e-6 * np.abs(ref).max())
    else:
        npt.assert_allclose(pred, ref, atol=1e-9 * np.abs(ref).max())
    if damping == 0.1:
        # The coefficient vector itself is only determined to about
        # eps * cond(A); cond(A) is moderate at this damping.
        npt.assert_allclose(coefs, expected, rtol=1e-7, atol=1e-9 * np.abs(expected).max())


def test_fit_multi_block_sizes(hm):
    """Sizes that are not multiples of the 128-wide Cholesky and MFMA tiles,
    spanning several blocks (exercises the blocked factorisation and the
    triangular solves end to end)."""
    from harmonica_amd._eqs import fit_device

    for n_data, n_src in [(520, 389), (300, 300), (1000, 129)]:
        coords, pts, rng = _scattered(3, n_data, n_src, depth=800.0)
        J = oracle.eqs_jacobian(coords, pts)
        data = J @ rng.normal(size=n_src)
        coefs, _ = fit_device(coords, data, None, pts, 1e-2)
        expected = oracle.eqs_fit_coefs(J, data, None, 1e-2)
        npt.assert_allclose(J @ coefs, J @ expected, atol=1e-10 * np.abs(J @ expected).max(),
                            err_msg=f"{n_data} x {n_src}")
        npt.assert_allclose(coefs, expected, rtol=1e-5, atol=1e-8 * np.abs(expected).max(),
                            err_msg=f"{n_data} x {n_src}")


@pytest.mark.parametrize("dtype", ["default", "float32"])
def test_equivalent_sources_cartesian(hm, synthetic, dtype):
    """test_eq_sources_cartesian.py:111-152."""
    points, masses, coords, data = synthetic
    atol = 1e-3 * np.abs(data).max()
    kwargs = {"dtype": dtype} if dtype != "default" else {}
    eqs = hm.EquivalentSources(damping=1e-12, **kwargs).fit(coords, data)
    npt.assert_allclose(data, eqs.predict(coords), atol=atol)
    grid_coords = grid_coordinates(REGION, shape=(60, 60), extra=0)
    true = oracle.point_gravity(grid_coords, points, masses, "g_z")
    npt.assert_allclose(true, eqs.predict(grid_coords), atol=atol)
    npt.assert_allclose(true, eqs.grid(grid_coords).scalars, atol=atol)
    profile = eqs.profile((REGION[0], REGION[2]), (REGION[0], REGION[3]), 0, 60)
    true = oracle.point_gravity((profile.easting, profile.northing, profile.upward), points,
                                masses, "g_z")
    npt.assert_allclose(true, profile.scalars, atol=atol)


def test_equivalent_sources_small_data(hm, synthetic):
    """test_eq_sources_cartesian.py:155-198 (damping None, depth 500)."""
    points, masses, _, _ = synthetic
    coords = grid_coordinates(REGION, shape=(8, 8), extra=0)
    data = oracle.point_gravity(coords, points, masses, "g_z")
    eqs = hm.EquivalentSources(depth=500).fit(coords, data)
    npt.assert_allclose(data, eqs.predict(coords), rtol=1e-5)
    tmp = [c.ravel() for c in coords]
    npt.assert_allclose(tmp[:2], eqs.points_[:2], rtol=1e-5)
    npt.assert_allclose(tmp[2] - 500, eqs.points_[2], rtol=1e-5)
    grid_coords = grid_coordinates(REGION, shape=(8, 8), extra=20)
    true = oracle.point_gravity(grid_coords, points, masses, "g_z")
    npt.assert_allclose(true, eqs.predict(grid_coords), rtol=0.08)
    npt.assert_allclose(true, eqs.grid(grid_coords).scalars, rtol=0.08)
    profile = eqs.profile((REGION[0], REGION[2]), (REGION[0], REGION[3]), 20, 10)
    true = oracle.point_gravity((profile.easting, profile.northing, profile.upward), points,
                                masses, "g_z")
    npt.assert_allclose(true, profile.scalars, rtol=0.05)


def test_custom_points_and_depth(hm, synthetic):
    """test_eq_sources_cartesian.py:255-281."""
    _, _, coords, data = synthetic
    custom = tuple(c.ravel() for c in grid_coordinates(REGION, shape=(3, 3), extra=-550))
    eqs = hm.EquivalentSources(points=custom).fit(coords, data)
    npt.assert_allclose(custom, eqs.points_, rtol=1e-5)
    assert eqs.depth_ is None
    small = grid_coordinates(REGION, shape=(5, 5))
    small = (small[0], small[1], np.arange(25, dtype=float).reshape(5, 5))
    d_small = oracle.point_gravity(small, custom, np.ones(9) * 1e9, "g_z")
    eqs = hm.EquivalentSources(depth=1.3e3).fit(small, d_small)
    npt.assert_allclose((small[0].ravel(), small[1].ravel(), small[2].ravel() - 1.3e3), eqs.points_)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("block_size", [None, 500])
@pytest.mark.parametrize("weights_none", [False, True])
def test_dtype(hm, synthetic, dtype, block_size, weights_none):
    """test_eq_sources_cartesian.py:339-378 (subset of the parameter grid)."""
    _, _, coords, data = synthetic
    weights = None if weights_none else np.ones_like(data)
    eqs = hm.EquivalentSources(damping=0.1, block_size=block_size, dtype=dtype)
    eqs.fit(coords, data, weights)
    prediction = eqs.predict(coords)
    for c in eqs.points_:
        assert c.dtype == np.dtype(dtype)
    assert prediction.dtype == np.dtype(dtype)


def test_grid_argument_errors(hm, synthetic):
    """test_eq_sources_cartesian.py:407-445."""
    _, _, coords, data = synthetic
    eqs = hm.EquivalentSources(damping=0.1).fit(coords, data)
    grid_coords = grid_coordinates(REGION, shape=(4, 4), extra=2e3)
    with pytest.raises(ValueError, match="The 'upward', 'region', 'shape' and 'spacing'"):
        eqs.grid(coordinates=grid_coords, upward=5e3, spacing=1)
    with pytest.warns(FutureWarning, match="The 'bla' arguments are being ignored."):
        eqs.grid(coordinates=grid_coords, bla="bla")


def test_default_depth_fit(hm, synthetic):
    _, _, coords, data = synthetic
    eqs = hm.EquivalentSources(damping=0.1).fit(coords, data)
    spacing = min(coords[0][1, 1] - coords[0][0, 0], coords[1][1, 1] - coords[1][0, 0])
    npt.assert_allclose(eqs.depth_, 4.5 * spacing)


@pytest.mark.parametrize("precision", ["mixed", "mixed-bf16", "mixed-fp32"])
@pytest.mark.parametrize("damping", [0.1, 1e-3])
def test_mixed_precision_fit_matches_fp64(hm, damping, precision):
    """Preconditioned mixed-precision fits (int8 or bf16 slice products, or
    FP32 MFMA with FP64 sums; FP64 Cholesky and FP64 residuals) against the
    FP64 fit:
    predictions on a grid agree to 1e-6 of their maximum, also at damping
    1e-3 where the normal matrix has cond ~ 1e11 (there the FP32 normal
    matrix needs a shift and the refinement takes tens of iterations)."""
    rng = np.random.default_rng(9)
    n = 3000
    L = np.sqrt(n * 0.16e6)
    coords = (rng.uniform(0, L, n), rng.uniform(0, L, n), rng.uniform(500, 1500, n))
    prisms = np.array([[0.3 * L, 0.5 * L, 0.2 * L, 0.6 * L, -4000.0, -1500.0]])
    data = oracle.prism_gravity(coords, prisms, [300.0], "g_z") + rng.normal(0, 0.05, n)
    grid = grid_coordinates((0, L, 0, L), shape=(40, 40), extra=1000.0)
    fp64 = hm.EquivalentSources(damping=damping, depth=4500).fit(coords, data)
    mixed = hm.EquivalentSources(damping=damping, depth=4500, fit_precision=precision).fit(coords, data)
    p64, pmx = fp64.predict(grid), mixed.predict(grid)
    err = np.abs(pmx - p64).max() / np.abs(p64).max()
    assert err <= 1e-6, (err, mixed.fit_info_)
    if precision != "mixed-fp32":  # the split normal matrix is close to the FP64 one
        assert mixed.fit_info_["iterations"] <= 30, mixed.fit_info_


@pytest.mark.parametrize("precision", ["mixed", "mixed-bf16"])
@pytest.mark.parametrize("n_data,n_src", [(5000, 700), (2049, 700), (6000, 2500)])
def test_split_normal_matrix_direct_solve(hm, precision, n_data, n_src):
    """The split normal matrix on its own: with no refinement (max_iter 0)
    the coefficients come from the direct solve with the low-precision
    matrix, which on a well-conditioned system (damping 1e3 on columns
    scaled to unit variance: cond below ~1e4) must agree with the FP64
    direct solve to about cond x 2^-31.  A wrong slice, digit or operand
    layout in the gram kernel gives differences of order 1.  700 sources
    (6 column tiles, the last partial); 5000 or 2049 data (several row
    groups of 1024, the last partial, odd chunk counts).  With 2500
    sources the factorisation makes two rank-1024 trailing updates (1476
    and 452 columns)."""
    from harmonica_amd._eqs import fit_device

    rng = np.random.default_rng(21)
    L = 40e3
    coords = (rng.uniform(0, L, n_data), rng.uniform(0, L, n_data), rng.uniform(500, 1500, n_data))
    pts = (rng.uniform(0, L, n_src), rng.uniform(0, L, n_src), rng.uniform(-3000, -1000, n_src))
    data = rng.normal(0, 1, n_data)
    w = rng.uniform(0.5, 2.0, n_data)
    c64, _ = fit_device(coords, data, w, pts, 1e3, "float64", max_iter=0)
    cmx, info = fit_device(coords, data, w, pts, 1e3, precision, max_iter=0)
    assert info["shift"] == 0.0, info
    err = np.abs(cmx - c64).max() / np.abs(c64).max()
    assert err <= 1e-5, err


@pytest.mark.parametrize("precision", ["mixed", "mixed-bf16", "mixed-fp32"])
@pytest.mark.parametrize("n_data,n_src", [(7, 5), (65, 33), (1025, 130)])
def test_mixed_fit_small_and_ragged(hm, precision, n_data, n_src):
    """Mixed fits where padding dominates: 7 data in one 64-row chunk (the
    rest zero digits), 65 and 1025 data one row past a chunk and a row
    group, 5 to 130 sources in partial 128-column tiles (with 130, the
    substitutions read the FP32 blocked copy of the factor's one block
    above the diagonal, 126 of its columns zero padding).  Refined to the
    FP64 normal equations, the coefficients match the FP64 fit's."""
    from harmonica_amd._eqs import fit_device

    rng = np.random.default_rng(n_data)
    L = 10e3
    coords = (rng.uniform(0, L, n_data), rng.uniform(0, L, n_data), rng.uniform(500, 1500, n_data))
    pts = (rng.uniform(0, L, n_src), rng.uniform(0, L, n_src), rng.uniform(-3000, -1000, n_src))
    data = rng.normal(0, 1, n_data)
    c64, _ = fit_device(coords, data, None, pts, 1.0, "float64")
    # refined past the default stop rule, so the coefficients themselves agree
    cmx, info = fit_device(coords, data, None, pts, 1.0, precision, rtol=1e-11)
    assert np.all(np.isfinite(cmx)), info
    npt.assert_allclose(cmx, c64, rtol=0, atol=1e-7 * np.abs(c64).max())


def _deep_grid_problem(depth, n_side=20, spacing=51.0, extra_sources=0):
    """Sources `depth` below a 20 x 20 grid of 51 m spacing (one per datum,
    plus `extra_sources` more, shifted and deeper); data from two deeper
    point masses."""
    e = np.arange(n_side) * spacing
    E, N = np.meshgrid(e, e)
    coords = (E.ravel(), N.ravel(), np.zeros(E.size))
    pts = (coords[0], coords[1], coords[2] - depth)
    if extra_sources:
        k = extra_sources
        pts = (np.concatenate([pts[0], coords[0][:k] + spacing / 3]),
               np.concatenate([pts[1], coords[1][:k] + spacing / 4]),
               np.concatenate([pts[2], coords[2][:k] - 1.3 * depth]))
    masses = (np.array([300.0, 700.0]), np.array([500.0, 200.0]), np.array([-800.0, -600.0]))
    data = oracle.eqs_jacobian(coords, masses) @ np.array([1e6, -2e6])
    return coords, pts, data


@pytest.mark.parametrize("depth,damping,route", [
    (500.0, 1e-18, "svd"), (1000.0, 1e-18, "svd"), (500.0, None, "lstsq"),
    (1000.0, None, "lstsq"), (200.0, 1e-18, "cholesky"), (1000.0, 1e-3, "cholesky"),
])
def test_fit_singular_follows_sklearn(hm, depth, damping, route):
    """Scaled jacobians with condition numbers 2e8-1e20 (400 sources under
    a 51 m grid): the fit takes scikit-learn's route (the SVD fallback of
    Ridge when the FP64 Cholesky factorisation fails, lstsq's cut-off for
    damping=None) and its predictions at the data match the oracle's
    restatement, itself pinned against scikit-learn.  At depth 200 m the
    normal matrix's condition number (5e16) is at the edge of FP64: LAPACK's
    Cholesky factorisation passes there and this one may not, so either
    route is accepted."""
    from harmonica_amd._eqs import fit_device

    coords, pts, data = _deep_grid_problem(depth)
    J = oracle.eqs_jacobian(coords, pts)
    expected, taken = oracle.eqs_fit_coefs_info(J, data, None, damping)
    assert taken == route
    coefs, info = fit_device(coords, data, None, pts, damping)
    if depth != 200.0:
        assert info["svd"] == (0.0 if route == "cholesky" else 1.0), info
    if info["svd"]:
        assert info["svd_sweeps"] > 0, info
    pred, ref = J @ coefs, J @ expected
    # the data are fitted as the reference fits them
    npt.assert_allclose(pred, data, rtol=0, atol=max(10 * np.abs(ref - data).max(),
                                                      1e-9 * np.abs(data).max()))
    npt.assert_allclose(pred, ref, rtol=0, atol=1e-6 * np.abs(ref).max())


def test_fit_no_damping_ill_conditioned_vs_lstsq(hm):
    """damping=None on the scattered problem with condition number 1.8e11
    and on a jacobian with exactly repeated columns: minimum-norm lstsq
    (repeated sources share the weight), as LinearRegression."""
    from harmonica_amd._eqs import fit_device

    coords, pts, rng = _scattered(2, 700, 300)
    J = oracle.eqs_jacobian(coords, pts)
    data = J @ rng.normal(size=300) + 1e-3 * rng.normal(size=700)
    coefs, info = fit_device(coords, data, None, pts, None)
    assert info["svd"] == 1.0 and info["svd_sweeps"] > 0
    expected = oracle.eqs_fit_coefs(J, data, None, None)
    # condition number 1.8e11: see test_fit_fp64_vs_least_squares_restatement
    npt.assert_allclose(J @ coefs, J @ expected, rtol=0, atol=1e-6 * np.abs(J @ expected).max())
    coords, pts, rng = _scattered(5, 500, 120, depth=600.0)
    pts = tuple(np.concatenate([x, x[:20]]) for x in pts)
    J = oracle.eqs_jacobian(coords, pts)
    data = J @ rng.normal(size=140) + 1e-3 * rng.normal(size=500)
    coefs, info = fit_device(coords, data, None, pts, None)
    assert info["svd"] == 1.0 and info["svd_rank"] == 120
    expected = oracle.eqs_fit_coefs(J, data, None, None)
    npt.assert_allclose(J @ coefs, J @ expected, rtol=0, atol=1e-9 * np.abs(J @ expected).max())
    npt.assert_allclose(coefs[:20], coefs[120:], rtol=1e-7)
    npt.assert_allclose(coefs, expected, rtol=0, atol=1e-6 * np.abs(expected).max())


@pytest.mark.parametrize("damping", [None, 1e-18, 1e-3])
def test_fit_more_sources_than_data(hm, damping):
    """144 data and 244 sources: LinearRegression's minimum-norm solution,
    Ridge's dual form and its lstsq fallback (damping 1e-18), against the
    oracle's restatement of each."""
    from harmonica_amd._eqs import fit_device

    coords, pts, data = _deep_grid_problem(400.0, n_side=12, extra_sources=100)
    J = oracle.eqs_jacobian(coords, pts)
    assert J.shape == (144, 244)
    expected, route = oracle.eqs_fit_coefs_info(J, data, None, damping)
    coefs, info = fit_device(coords, data, None, pts, damping)
    pred, ref = J @ coefs, J @ expected
    if route == "dual-lstsq":
        # The reference forms X X^T + alpha I and keeps its eigenvalues above
        # 2^-52 of the largest: those near the cut-off carry relative errors
        # of order one in the reference itself, so the fitted values agree to
        # about the data's components along them (1e-6 here); the fit here
        # matches the data at least as closely.
        assert info["svd"] == 1.0
        npt.assert_allclose(pred, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
        assert np.abs(pred - data).max() <= 1.5 * np.abs(ref - data).max() + 1e-9 * np.abs(data).max()
    else:
        npt.assert_allclose(pred, ref, rtol=0, atol=1e-7 * np.abs(ref).max())
