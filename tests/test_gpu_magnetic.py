"""
GPU parity of ``prism_magnetic`` (hmx_prism_magnetic) against the CPU
oracle, the reference's own magnetic test geometries
(test/test_prism_magnetic.py) and size-independent properties at the full
1M-prism size.  Tolerance: 1e-10 x max|component| (FP64; the same closed
forms summed in a different order with different roundings); NaN at the
same places.
"""

import ctypes

import numpy as np
import numpy.testing as npt
import pytest

import oracle

pytestmark = pytest.mark.gpu

RTOL_MAX = 1e-10
FIELDS = ("b", "b_e", "b_n", "b_u")


@pytest.fixture(scope="module")
def hm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import harmonica_amd

    return harmonica_amd


def assert_parity(got, expected, tol=RTOL_MAX, err_msg=""):
    got, expected = np.asarray(got), np.asarray(expected)
    assert got.shape == expected.shape, err_msg
    npt.assert_array_equal(np.isnan(got), np.isnan(expected), err_msg=err_msg)
    finite = np.isfinite(expected)
    scale = np.abs(expected[finite]).max() if finite.any() else 1.0
    npt.assert_allclose(got[finite], expected[finite], rtol=0, atol=tol * scale, err_msg=err_msg)


def _random_model(seed, n_prisms=57, n_obs=300):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-2000, 1500, n_prisms)
    s = rng.uniform(-2000, 1500, n_prisms)
    b = rng.uniform(-1500, -100, n_prisms)
    prisms = np.column_stack([w, w + rng.uniform(10, 600, n_prisms), s,
                              s + rng.uniform(10, 600, n_prisms), b, b + rng.uniform(10, 900, n_prisms)])
    mag = tuple(rng.normal(0, 2.0, n_prisms) for _ in range(3))
    coords = (rng.uniform(-2500, 2500, n_obs), rng.uniform(-2500, 2500, n_obs),
              rng.uniform(-2000, 600, n_obs))
    return coords, prisms, mag


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("field", FIELDS)
def test_prism_magnetic_vs_oracle(hm, field, seed):
    coords, prisms, mag = _random_model(seed)
    got = hm.prism_magnetic(coords, prisms, mag, field)
    expected = oracle.prism_magnetic(coords, prisms, mag, field)
    if field == "b":
        assert isinstance(got, tuple) and len(got) == 3
        for c, (g, x) in enumerate(zip(got, expected)):
            assert_parity(g, x, err_msg=f"b[{c}]")
    else:
        assert_parity(got, expected, err_msg=field)


def test_reference_sample_against_oracle(hm):
    """The four prisms, magnetisations and points of the reference's
    TestAgainstChoclo (test_prism_magnetic.py:231-269)."""
    prisms = np.array([[-10, 10, -10, 0, -10, 0], [-10, 0, -10, 10, -20, -10],
                       [5, 15, 5, 15, -15, -5], [-5, 5, -5, 5, -35, -20]], dtype=float)
    mag = (np.array([1.0, -1.0, 3.0, -1.5]), np.array([1.0, -0.5, -1.0, 2.0]),
           np.array([1.0, 2.0, -3.0, -2.0]))
    coords = (np.array([-5, 10, 0, 15], dtype=float), np.array([14, -4, 11, 0], dtype=float),
              np.array([9, 6, 6, 12], dtype=float))
    b = hm.prism_magnetic(coords, prisms, mag, "b")
    expected = oracle.prism_magnetic(coords, prisms, mag, "b")
    for g, x in zip(b, expected):
        assert_parity(g, x)
    for c, field in enumerate(("b_e", "b_n", "b_u")):
        npt.assert_allclose(hm.prism_magnetic(coords, prisms, mag, field), b[c], rtol=1e-14,
                            atol=0)


def test_points_on_prism_planes(hm):
    """Points snapped onto the face planes of one prism (faces, edges,
    vertices, their extensions, inside): every component against the
    oracle, NaN exactly on edges and vertices."""
    rng = np.random.default_rng(21)
    prism = np.array([-40.0, 60.0, -30.0, 50.0, -80.0, 20.0])
    n = 4000
    pts = rng.uniform(-150, 150, (3, n))
    for ax in range(3):
        lo, hi = prism[2 * ax], prism[2 * ax + 1]
        u = rng.uniform(size=n)
        pts[ax] = np.where(u < 0.2, lo, np.where(u < 0.4, hi, pts[ax]))
    coords = tuple(pts)
    mag = ([1.3], [-0.4], [2.2])
    got = hm.prism_magnetic(coords, prism, mag, "b")
    expected = oracle.prism_magnetic(coords, prism, mag, "b")
    for c in range(3):
        assert_parity(got[c], expected[c], err_msg=f"b[{c}]")
    assert np.isnan(got[0]).sum() > 100


def test_disable_checks_invalid_prisms(hm):
    """test_prism_magnetic.py:160-166: no error with checks disabled."""
    for bad in ([100, 0, -100, 0, -200, -100], [-100, 0, 200, 0, -200, -100],
                [-100, 0, -100, 0, -100, -200]):
        for field in FIELDS:
            hm.prism_magnetic([0, 0, 0], bad, [100, 10, -10], field, disable_checks=True)


def test_only_null_prisms_give_zeros(hm):
    coords = ([0.0, 5.0], [0.0, 1.0], [10.0, 10.0])
    b = hm.prism_magnetic(coords, [[0, 0, 0, 1, 0, 1], [0, 1, 0, 1, 0, 1]],
                          ([1.0, 0.0], [1.0, 0.0], [1.0, 0.0]), "b")
    for x in b:
        npt.assert_array_equal(x, 0.0)


def test_float32_dtype_and_shape(hm):
    coords = tuple(np.broadcast_to(c, (4, 5)) for c in np.meshgrid(np.linspace(-50, 50, 5),
                                                                   np.linspace(-50, 50, 4)))
    coords = (coords[0], coords[1], np.full((4, 5), 30.0))
    b = hm.prism_magnetic(coords, [-10, 10, -10, 10, -40, -20], [1.0, 2.0, 3.0], "b",
                          dtype=np.float32)
    assert all(x.dtype == np.float32 and x.shape == (4, 5) for x in b)
    ref = oracle.prism_magnetic(coords, [-10, 10, -10, 10, -40, -20], ([1.0], [2.0], [3.0]), "b")
    for x, r in zip(b, ref):
        npt.assert_allclose(x, r, rtol=1e-6)


def test_c_abi_matches_api(hm):
    """hmx_prism_magnetic called directly (component subsets), against the
    drop-in API."""
    import torch

    from harmonica_amd import _lib

    coords, prisms, mag = _random_model(3, n_prisms=200, n_obs=700)
    api = hm.prism_magnetic(coords, prisms, mag, "b")
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    e, n, u = (dev(c) for c in coords)
    P = dev(prisms)
    me, mn, mu = (dev(m) for m in mag)
    L = _lib.lib()
    for mask in range(1, 8):
        nbytes = ctypes.c_size_t(0)
        assert L.hmx_prism_magnetic_workspace_bytes(e.numel(), P.shape[0], mask,
                                                    ctypes.byref(nbytes)) == 0
        ws = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
        rows = [c for c in range(3) if mask >> c & 1]
        out = torch.empty((len(rows), e.numel()), dtype=torch.float64, device="cuda")
        st = L.hmx_prism_magnetic(_lib.ptr(e), _lib.ptr(n), _lib.ptr(u), e.numel(), _lib.ptr(P),
                                  _lib.ptr(me), _lib.ptr(mn), _lib.ptr(mu), P.shape[0], mask,
                                  _lib.ptr(out), _lib.ptr(ws), nbytes.value,
                                  _lib.stream_handle(None))
        assert st == 0, _lib.lib().hmx_last_error()
        got = out.cpu().numpy()
        for k, c in enumerate(rows):
            npt.assert_allclose(got[k], api[c], rtol=1e-14, atol=1e-14 * np.abs(api[c]).max())
    # error paths: bad mask, small workspace
    assert L.hmx_prism_magnetic(_lib.ptr(e), _lib.ptr(n), _lib.ptr(u), e.numel(), _lib.ptr(P),
                                _lib.ptr(me), _lib.ptr(mn), _lib.ptr(mu), P.shape[0], 8,
                                _lib.ptr(out), _lib.ptr(ws), nbytes.value,
                                _lib.stream_handle(None)) != 0
    assert L.hmx_prism_magnetic(_lib.ptr(e), _lib.ptr(n), _lib.ptr(u), e.numel(), _lib.ptr(P),
                                _lib.ptr(me), _lib.ptr(mn), _lib.ptr(mu), P.shape[0], 7,
                                _lib.ptr(out), _lib.ptr(ws), 8, _lib.stream_handle(None)) != 0


def test_terrain_layer_parity(hm):
    """100 x 100 prisms of a fractal terrain (shared edges and faces, grid
    points on edge planes) with induced + remanent magnetisation."""
    from harmonica_amd import workloads

    w = workloads.terrain_layer_large(n_prisms_side=100, n_obs_side=48)
    prisms = workloads.layer_to_prisms(w["easting"], w["northing"], w["surface"], w["reference"])
    mag = workloads.terrain_magnetization(n_prisms_side=100)
    coords = w["coordinates"]
    got = hm.prism_magnetic(coords, prisms, mag, "b")
    expected = oracle.prism_magnetic(coords, prisms, mag, "b")
    for c in range(3):
        assert_parity(got[c], expected[c], err_msg=f"b[{c}]")


def test_terrain_1m_sample_parity_and_linearity(hm):
    """The 1M-prism terrain with its magnetisation at 64 of the 256 x 256
    grid points at 2500 m against the oracle, and at all 65,536 points the
    linearity b(M1 + M2) = b(M1) + b(M2)."""
    from harmonica_amd import workloads

    w = workloads.terrain_layer_large()
    prisms = workloads.layer_to_prisms(w["easting"], w["northing"], w["surface"], w["reference"])
    mag = workloads.terrain_magnetization()
    idx = np.array([0, 31, 51, 102, 140, 204, 230, 255])
    sample = tuple(c[np.ix_(idx, idx)] for c in w["coordinates"])
    got = hm.prism_magnetic(sample, prisms, mag, "b")
    expected = oracle.prism_magnetic(sample, prisms, mag, "b")
    for c in range(3):
        assert_parity(got[c], expected[c], err_msg=f"b[{c}]")
    coords = w["coordinates"]
    rng = np.random.default_rng(9)
    m2 = tuple(rng.normal(0.0, 0.5, m.size) for m in mag)
    b1 = hm.prism_magnetic(coords, prisms, mag, "b")
    b2 = hm.prism_magnetic(coords, prisms, m2, "b")
    b12 = hm.prism_magnetic(coords, prisms, tuple(a + b for a, b in zip(mag, m2)), "b")
    for c in range(3):
        scale = np.abs(b12[c]).max()
        npt.assert_allclose(b12[c], b1[c] + b2[c], rtol=0, atol=1e-11 * scale)
