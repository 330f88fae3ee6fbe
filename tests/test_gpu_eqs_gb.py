"""
GPU tests of EquivalentSourcesGB (hmx_eqs_gb_fit): parity of the whole
gradient-boosting loop with the oracle's CPU restatement on the same
windows, and the reference's own checks
(test/test_gradient_boosted_eqs.py:114-128, 159-235, 309-361).
"""

import numpy as np
import numpy.testing as npt
import pytest

import oracle
from conftest import grid_coordinates

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import harmonica_amd

    return harmonica_amd


REGION = (-3e3, -1e3, 5e3, 7e3)


def checkerboard(amplitude, region, coordinates):
    """verde.synthetic.CheckerBoard with its default wavelengths (half the
    region's extent in each direction)."""
    w_e = (region[1] - region[0]) / 2
    w_n = (region[3] - region[2]) / 2
    e, n = coordinates[:2]
    return amplitude * np.sin(2 * np.pi / w_e * e) * np.cos(2 * np.pi / w_n * n)


@pytest.fixture(scope="module")
def model():
    """Point masses under a checkerboard (test_gradient_boosted_eqs.py:22-91);
    data from the oracle's point_gravity."""
    points = grid_coordinates(REGION, shape=(6, 6), extra=-1e3)
    masses = checkerboard(1e13, REGION, points)
    coordinates = grid_coordinates(REGION, shape=(40, 40), extra=0.0)
    data = oracle.point_gravity(coordinates, points, masses, "g_z")
    small = grid_coordinates(REGION, shape=(8, 8), extra=0.0)
    data_small = oracle.point_gravity(small, points, masses, "g_z")
    return {"points": points, "masses": masses, "coordinates": coordinates, "data": data,
            "small": small, "data_small": data_small}


def maxabs(*arrays):
    return max(np.abs(a).max() for a in arrays)


@pytest.mark.parametrize("damping", [1e-2, 1.0])
@pytest.mark.parametrize("weighted", [False, True])
def test_gb_vs_oracle_same_windows(hm, damping, weighted):
    """2000 scattered data, windows of ~250 data, 50% overlap, shuffled by
    random_state: coefficients, RMS history and grid predictions against the
    oracle's loop on the windows of its own direct-comparison restatement."""
    from sklearn import utils

    rng = np.random.default_rng(11)
    n = 2000
    coords = (rng.uniform(0, 8e3, n), rng.uniform(0, 8e3, n), rng.uniform(100, 300, n))
    prisms = [[2e3, 4e3, 2e3, 3e3, -1.5e3, -500], [5e3, 6e3, 4e3, 7e3, -2e3, -1e3]]
    data = oracle.prism_gravity(coords, prisms, [300.0, -200.0], "g_z")
    weights = rng.uniform(0.5, 2.0, n) if weighted else None
    eqs = hm.EquivalentSourcesGB(damping=damping, depth=1000.0, window_size=1500.0,
                                 random_state=7)
    eqs.fit(coords, data, weights)
    region = eqs.region_
    src_w = oracle.rolling_window_windows(eqs.points_, 1500.0, 0.5, region)
    dat_w = oracle.rolling_window_windows(coords, 1500.0, 0.5, region)
    src_w, dat_w = utils.shuffle(src_w, dat_w, random_state=7)
    keep = [(s, d) for s, d in zip(src_w, dat_w) if s.size > 0 and d.size > 0]
    coefs, rmse = oracle.eqs_gb_fit(coords, data, weights, eqs.points_, damping,
                                    [k[0] for k in keep], [k[1] for k in keep])
    assert eqs.fit_info_["windows"] == len(keep) > 20
    npt.assert_allclose(eqs.rmse_per_iteration_, rmse, rtol=1e-8)
    grid = grid_coordinates((0, 8e3, 0, 8e3), shape=(60, 60), extra=400.0)
    pred = eqs.predict(grid)
    expected = oracle.eqs_predict(grid, eqs.points_, coefs)
    npt.assert_allclose(pred, expected, rtol=0, atol=1e-8 * np.abs(expected).max())
    npt.assert_allclose(eqs.coefs_, coefs, rtol=0, atol=1e-7 * np.abs(coefs).max())


@pytest.mark.parametrize("weights", [None, np.ones((8, 8))])
def test_gb_eqs_small_data(hm, model, weights):
    """test_gradient_boosted_eqs.py:157-176."""
    eqs = hm.EquivalentSourcesGB(depth=1e3, damping=None, window_size=1e3, random_state=42)
    eqs.fit(model["small"], model["data_small"], weights=weights)
    npt.assert_allclose(model["data_small"], eqs.predict(model["small"]), rtol=0,
                        atol=0.05 * maxabs(model["data_small"]))


def test_gradient_boosted_eqs_single_window(hm, model):
    """test_gradient_boosted_eqs.py:178-196 at the reference's tolerances:
    one window over the whole region interpolates the data.  The scaled
    jacobian (1600 sources 500 m below a 51 m grid) has a condition number
    of about 2e20: the FP64 Cholesky factorisation of its normal matrix
    fails, and the window is solved by scikit-learn's SVD route, as the
    reference's Ridge does."""
    atol = 5e-8
    eqs = hm.EquivalentSourcesGB(depth=500, window_size=REGION[1] - REGION[0], damping=1e-18)
    eqs.fit(model["coordinates"], model["data"])
    assert eqs.fit_info_["windows"] == 1
    assert eqs.fit_info_["svd"] == 1.0 and eqs.fit_info_["shift"] == 0.0
    npt.assert_allclose(model["data"], eqs.predict(model["coordinates"]), rtol=1e-5, atol=atol)
    grid = grid_coordinates(REGION, shape=(60, 60), extra=0.0)
    true = oracle.point_gravity(grid, model["points"], model["masses"], "g_z")
    npt.assert_allclose(true, eqs.predict(grid), rtol=1e-3, atol=atol)


def test_gradient_boosted_eqs_predictions(hm, model):
    """test_gradient_boosted_eqs.py:198-221."""
    eqs = hm.EquivalentSourcesGB(window_size=1e3, depth=1e3, damping=1e-24, random_state=42)
    eqs.fit(model["coordinates"], model["data"])
    npt.assert_allclose(model["data"], eqs.predict(model["coordinates"]), rtol=0,
                        atol=0.02 * maxabs(model["data"]))
    grid = grid_coordinates(REGION, shape=(60, 60), extra=0.0)
    true = oracle.point_gravity(grid, model["points"], model["masses"], "g_z")
    npt.assert_allclose(true, eqs.predict(grid), rtol=0, atol=0.02 * maxabs(true))
    # the RMS of the residue falls from the data's
    assert eqs.rmse_per_iteration_[-1] < 0.05 * eqs.rmse_per_iteration_[0]


def test_gradient_boosted_eqs_random_state(hm, model):
    """test_gradient_boosted_eqs.py:223-235."""
    a = hm.EquivalentSourcesGB(window_size=500, random_state=0).fit(model["coordinates"],
                                                                    model["data"])
    b = hm.EquivalentSourcesGB(window_size=500, random_state=0).fit(model["coordinates"],
                                                                    model["data"])
    npt.assert_array_equal(a.coefs_, b.coefs_)


def test_custom_points(hm, model):
    """test_gradient_boosted_eqs.py:114-128."""
    points_custom = grid_coordinates(REGION, spacing=250, extra=-1e3)
    eqs = hm.EquivalentSourcesGB(points=points_custom, window_size=500, depth=500)
    eqs.fit(model["small"], model["data_small"])
    for p, q in zip(points_custom, eqs.points_):
        npt.assert_allclose(np.ravel(p), q, rtol=1e-5)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_dtype(hm, model, dtype):
    """test_gradient_boosted_eqs.py:303-345: float32 sources, coefficients
    and predictions; the predictions still fit the data (float32 tolerance
    of test_gradient_boosted_eqs.py:348-361)."""
    eqs = hm.EquivalentSourcesGB(damping=None, window_size=1e3, random_state=42,
                                 dtype=dtype, depth=500)
    eqs.fit(model["coordinates"], model["data"])
    for p in eqs.points_:
        assert p.dtype == np.dtype(dtype)
    assert eqs.coefs_.dtype == np.dtype(dtype)
    pred = eqs.predict(model["coordinates"])
    assert pred.dtype == np.dtype(dtype)
    npt.assert_allclose(model["data"], pred, rtol=0, atol=0.05 * maxabs(model["data"]))


def test_grid_arguments(hm, model):
    """test_gradient_boosted_eqs.py:364-401: deprecated and ignored grid
    arguments."""
    eqs = hm.EquivalentSourcesGB(window_size=500).fit(model["small"], model["data_small"])
    grid = grid_coordinates(REGION, shape=(4, 4), extra=2e3)
    with pytest.raises(ValueError):
        eqs.grid(coordinates=grid, upward=2e3)
    with pytest.warns(FutureWarning):
        eqs.grid(coordinates=grid, bla="bla")


@pytest.mark.parametrize("weighted", [False, True])
def test_batched_factorisation_matches_window_by_window(hm, weighted):
    """FP64 windows with damping: the normal matrices of batches of windows
    formed and factorised together (padded to a common size), each window
    then solved from its factor, against the window-by-window fit
    (hmx_eqs_fit per window): the padding adds exact zeros, so the
    coefficients agree to rounding.  12,000 scattered data, windows of
    about 1500 data, batches of 5 windows (the last partial)."""
    from harmonica_amd import _eqs_gb

    rng = np.random.default_rng(31)
    n = 12000
    coords = (rng.uniform(0, 3e4, n), rng.uniform(0, 3e4, n), rng.uniform(100, 600, n))
    prisms = np.array([[8e3, 14e3, 5e3, 20e3, -5e3, -2e3]])
    data = oracle.prism_gravity(coords, prisms, [250.0], "g_z") + rng.normal(0, 0.05, n)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    fits = {}
    try:
        for batch in (1, 5):
            _eqs_gb.FACTOR_BATCH = batch
            eqs = hm.EquivalentSourcesGB(damping=1e-2, depth=2e3, window_size=8e3,
                                         random_state=2)
            eqs.fit(coords, data, weights=w)
            fits[batch] = eqs
    finally:
        _eqs_gb.FACTOR_BATCH = None
    one, many = fits[1], fits[5]
    assert one.fit_info_["factor_batch"] == 1 and many.fit_info_["factor_batch"] == 5
    assert many.fit_info_["windows"] > 5 and many.fit_info_["windows"] % 5 != 0
    npt.assert_allclose(many.coefs_, one.coefs_, rtol=0, atol=1e-9 * np.abs(one.coefs_).max())
    npt.assert_allclose(many.rmse_per_iteration_, one.rmse_per_iteration_, rtol=1e-9)
    grid = grid_coordinates((0, 3e4, 0, 3e4), shape=(20, 20), extra=500.0)
    p1, p5 = one.predict(grid), many.predict(grid)
    npt.assert_allclose(p5, p1, rtol=0, atol=1e-10 * np.abs(p1).max())
