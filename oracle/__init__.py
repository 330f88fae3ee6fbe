"""
CPU oracle for the pairwise gravity path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; it is the checker, never the product.
The ``harmonica_amd`` package does not import it and has no CPU fallback.

The functions below are thin ctypes wrappers around ``oracle.c`` (a C99 +
OpenMP restatement of the reference's Numba loops and of the closed forms
they call; see the citations in that file) that reproduce the reference's
Python-level post-processing: sign flips and unit conversions of
``src/harmonica/_forward/prisms/gravity.py:228-235``,
``src/harmonica/_forward/prisms/layer.py:425-433`` and
``src/harmonica/_forward/point.py:253-260``, and two Python-level
restatements: the gradient-boosting window loop of
``src/harmonica/_equivalent_sources/gradient_boosted.py:231-283`` over the
oracle's jacobian, least squares and predict, with window membership by
direct comparison (``rolling_window_windows``), and the null-prism discard
and nT conversion of ``src/harmonica/_forward/prisms/magnetic.py``.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

#: Field names in the reference's FIELDS order (gravity.py:37-48).
FIELDS = ("potential", "g_e", "g_n", "g_z", "g_ee", "g_nn", "g_zz", "g_en", "g_ez", "g_nz")
FIELD_ID = {f: i for i, f in enumerate(FIELDS)}
GRAVITATIONAL_CONST = 6.6743e-11

_lib = None


def build():
    """Compile liboracle.so from oracle.c with the Makefile beside it."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    """Load (building first if absent) the oracle shared library."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        ci = ctypes.c_int
        cd = ctypes.c_double
        L.or_prism_gravity.argtypes = [dp, dp, dp, i64, dp, dp, i64, ci, dp, ci]
        L.or_prism_any_singular.argtypes = [dp, dp, dp, i64, dp, i64, ci]
        L.or_prism_any_singular.restype = ci
        L.or_prism_layer.argtypes = [dp, dp, dp, i64, dp, i64, dp, i64, dp, dp, dp, cd, ci, dp, ci]
        L.or_prism_pair.argtypes = [ci, cd, cd, cd, dp, cd]
        L.or_prism_pair.restype = cd
        L.or_prism_dgu_dtop.argtypes = [cd, cd, cd, dp, cd]
        L.or_prism_dgu_dtop.restype = cd
        L.or_point_pair.argtypes = [ci, cd, cd, cd, cd, cd, cd, cd]
        L.or_point_pair.restype = cd
        L.or_point_cartesian.argtypes = [dp, dp, dp, i64, dp, dp, dp, dp, i64, ci, dp, ci]
        L.or_point_spherical.argtypes = [dp, dp, dp, i64, dp, dp, dp, dp, i64, ci, dp, ci]
        L.or_eqs_jacobian.argtypes = [dp, dp, dp, i64, dp, dp, dp, i64, dp, ci]
        L.or_eqs_predict.argtypes = [dp, dp, dp, i64, dp, dp, dp, dp, i64, dp, ci]
        L.or_prism_magnetic.argtypes = [dp, dp, dp, i64, dp, dp, i64, ci, dp, ci]
        L.or_prism_magnetic_pair.argtypes = [cd, cd, cd, dp, cd, cd, cd, dp]
        L.or_eqs_sph_jacobian.argtypes = [dp, dp, dp, i64, dp, dp, dp, i64, dp, ci]
        L.or_eqs_sph_predict.argtypes = [dp, dp, dp, i64, dp, dp, dp, dp, i64, dp, ci]
        L.or_tesseroid_gravity.argtypes = [dp, dp, dp, i64, dp, dp, i64, ci, cd, ci, dp, ci]
        L.or_tesseroid_gravity.restype = ci
        L.or_tesseroid_split_list.argtypes = [cd, cd, cd, dp, cd, ci, ci, dp, ctypes.c_long]
        L.or_tesseroid_split_list.restype = ctypes.c_long
        L.or_num_threads.restype = ci
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _units(field, result):
    """Sign flip and unit conversion (gravity.py:228-235, point.py:253-260)."""
    if field in ("g_z", "g_ez", "g_nz", "g_ze", "g_zn"):
        result *= -1
    if field in ("g_e", "g_n", "g_z"):
        result *= 1e5
    if field in ("g_ee", "g_nn", "g_zz", "g_en", "g_ez", "g_nz", "g_ne", "g_ze", "g_zn"):
        result *= 1e9
    return result


def _coords(coordinates):
    cast = np.broadcast(*coordinates[:3])
    arrs = tuple(_f64(np.broadcast_to(c, cast.shape).ravel()) for c in coordinates[:3])
    return cast.shape, arrs


def prism_gravity(coordinates, prisms, density, field, n_threads=0):
    """Oracle of ``harmonica.prism_gravity`` (gravity.py:51-236) without the
    input checks; null prisms are kept (they contribute exactly zero)."""
    shape, (e, n, u) = _coords(coordinates)
    prisms = _f64(np.atleast_2d(prisms))
    density = _f64(np.atleast_1d(density).ravel())
    out = np.zeros(e.size)
    lib().or_prism_gravity(_p(e), _p(n), _p(u), e.size, _p(prisms), _p(density),
                           prisms.shape[0], FIELD_ID[field], _p(out), n_threads)
    return _units(field, out).reshape(shape)


def prism_any_singular(coordinates, prisms, field):
    """Oracle of ``_check_singular_points`` (gravity.py:239-449)."""
    _, (e, n, u) = _coords(coordinates)
    prisms = _f64(np.atleast_2d(prisms))
    return bool(lib().or_prism_any_singular(_p(e), _p(n), _p(u), e.size, _p(prisms),
                                            prisms.shape[0], FIELD_ID[field]))


def prism_pair(field, point, prism, density):
    """One prism at one point, SI units, before sign flip/units."""
    prism = _f64(prism)
    return lib().or_prism_pair(FIELD_ID[field], float(point[0]), float(point[1]),
                               float(point[2]), _p(prism), float(density))


def prism_dgu_dtop(point, prism, density):
    """d(g_u SI)/d(top) of one prism at one point."""
    prism = _f64(prism)
    return lib().or_prism_dgu_dtop(float(point[0]), float(point[1]), float(point[2]),
                                   _p(prism), float(density))


def prism_layer_gravity(coordinates, easting, northing, bottom, top, density, field,
                        thickness_threshold=None, n_threads=0):
    """Oracle of ``DatasetAccessorPrismLayer.gravity`` (layer.py:314-433)."""
    shape, (e, n, u) = _coords(coordinates)
    ce, cn = _f64(easting), _f64(northing)
    b, t, d = _f64(bottom), _f64(top), _f64(density)
    thr = 0.0 if thickness_threshold is None else float(thickness_threshold)
    out = np.zeros(e.size)
    lib().or_prism_layer(_p(e), _p(n), _p(u), e.size, _p(ce), ce.size, _p(cn), cn.size,
                         _p(b), _p(t), _p(d), thr, FIELD_ID[field], _p(out), n_threads)
    return _units(field, out).reshape(shape)


_ALIASES = {"g_ne": "g_en", "g_ze": "g_ez", "g_zn": "g_nz"}


def point_gravity(coordinates, points, masses, field, coordinate_system="cartesian",
                  n_threads=0):
    """Oracle of ``harmonica.point_gravity`` (point.py:30-261)."""
    shape, (e, n, u) = _coords(coordinates)
    pe, pn, pu = (_f64(np.atleast_1d(p).ravel()) for p in points[:3])
    m = _f64(np.atleast_1d(masses).ravel())
    out = np.zeros(e.size)
    fid = FIELD_ID[_ALIASES.get(field, field)]
    if coordinate_system == "cartesian":
        lib().or_point_cartesian(_p(e), _p(n), _p(u), e.size, _p(pe), _p(pn), _p(pu),
                                 _p(m), m.size, fid, _p(out), n_threads)
    else:
        lib().or_point_spherical(_p(e), _p(n), _p(u), e.size, _p(pe), _p(pn), _p(pu),
                                 _p(m), m.size, fid, _p(out), n_threads)
    return _units(field, out).reshape(shape)


def point_pair(field, point, source, mass):
    return lib().or_point_pair(FIELD_ID[field], *(float(v) for v in point),
                               *(float(v) for v in source), float(mass))


def eqs_jacobian(coordinates, points, n_threads=0):
    """Oracle of ``EquivalentSources.jacobian`` (cartesian.py:385-415)."""
    e, n, u = (_f64(np.atleast_1d(c).ravel()) for c in coordinates[:3])
    pe, pn, pu = (_f64(np.atleast_1d(c).ravel()) for c in points[:3])
    jac = np.zeros((e.size, pe.size))
    lib().or_eqs_jacobian(_p(e), _p(n), _p(u), e.size, _p(pe), _p(pn), _p(pu), pe.size,
                          _p(jac), n_threads)
    return jac


def eqs_predict(coordinates, points, coefs, n_threads=0):
    """Oracle of ``EquivalentSources.predict``'s loop (utils.py:77-97)."""
    shape, (e, n, u) = _coords(coordinates)
    pe, pn, pu = (_f64(np.atleast_1d(c).ravel()) for c in points[:3])
    c = _f64(coefs)
    out = np.zeros(e.size)
    lib().or_eqs_predict(_p(e), _p(n), _p(u), e.size, _p(pe), _p(pn), _p(pu), _p(c),
                         pe.size, _p(out), n_threads)
    return out.reshape(shape)


def eqs_sph_jacobian(coordinates, points, n_threads=0):
    """Oracle of ``EquivalentSourcesSph.jacobian`` (spherical.py:250-282)."""
    c = tuple(_f64(np.atleast_1d(x).ravel()) for x in coordinates[:3])
    p = tuple(_f64(np.atleast_1d(x).ravel()) for x in points[:3])
    jac = np.zeros((c[0].size, p[0].size))
    lib().or_eqs_sph_jacobian(*(_p(x) for x in c), c[0].size, *(_p(x) for x in p), p[0].size,
                              _p(jac), n_threads)
    return jac


def eqs_sph_predict(coordinates, points, coefs, n_threads=0):
    """Oracle of ``EquivalentSourcesSph.predict``'s loop (utils.py:77-97)."""
    shape, c = _coords(coordinates)
    p = tuple(_f64(np.atleast_1d(x).ravel()) for x in points[:3])
    cf = _f64(coefs)
    out = np.zeros(c[0].size)
    lib().or_eqs_sph_predict(*(_p(x) for x in c), c[0].size, *(_p(x) for x in p), _p(cf),
                             p[0].size, _p(out), n_threads)
    return out.reshape(shape)


def eqs_fit_coefs(jacobian, data, weights=None, damping=None):
    """Restatement of verde's ``least_squares`` as called at cartesian.py:280
    (verde 1.8, not vendored): columns scaled by their population standard
    deviation (zero -> 1; StandardScaler(with_mean=False)), rows and data
    weighted by sqrt(w), then scikit-learn's solver (1.7, not vendored):

    * damping None: LinearRegression -> scipy.linalg.lstsq with the cut-off
      cond = max(n, m) eps (minimum-norm solution, singular values
      s <= cond s_max dropped);
    * damping alpha, m <= n: Ridge's Cholesky solve of X^T X + alpha I
      (scipy solve, assume_a="pos"); when that raises LinAlgError, the SVD
      solve c = V diag(s / (s^2 + alpha)) U^T y over s > 1e-15;
    * damping alpha, m > n: the dual form X^T (X X^T + alpha I)^-1 y, and
      when its Cholesky solve raises, the dual from scipy.linalg.lstsq of
      X X^T + alpha I (default cut-off eps).

    The result is unscaled (c / scale).  Pinned in the tests against
    scikit-learn's StandardScaler + Ridge / LinearRegression, including
    singular problems."""
    return eqs_fit_coefs_info(jacobian, data, weights, damping)[0]


def eqs_fit_coefs_info(jacobian, data, weights=None, damping=None):
    """``eqs_fit_coefs`` plus the route taken: "lstsq", "cholesky", "svd",
    "dual" or "dual-lstsq"."""
    import scipy.linalg

    J = np.asarray(jacobian, dtype=np.float64)
    d = np.asarray(data, dtype=np.float64).ravel()
    scale = J.std(axis=0)
    scale[scale == 0] = 1.0
    Js = J / scale
    sw = None if weights is None else np.sqrt(np.asarray(weights, dtype=np.float64).ravel())
    X = Js if sw is None else Js * sw[:, None]
    y = d if sw is None else d * sw
    n, m = X.shape
    if damping is None:
        cond = max(n, m) * np.finfo(np.float64).eps
        sol = scipy.linalg.lstsq(X, y, cond=cond)[0]
        route = "lstsq"
    elif m > n:
        K = X @ X.T + damping * np.eye(n)
        try:
            dual = scipy.linalg.solve(K, y, assume_a="pos")
            route = "dual"
        except np.linalg.LinAlgError:
            dual = scipy.linalg.lstsq(K, y)[0]
            route = "dual-lstsq"
        sol = X.T @ dual
    else:
        A = X.T @ X + damping * np.eye(m)
        try:
            sol = scipy.linalg.solve(A, X.T @ y, assume_a="pos")
            route = "cholesky"
        except np.linalg.LinAlgError:
            U, s, Vt = np.linalg.svd(X, full_matrices=False)
            keep = s > 1e-15
            f = np.zeros_like(s)
            f[keep] = s[keep] / (s[keep] ** 2 + damping)
            sol = Vt.T @ (f * (U.T @ y))
            route = "svd"
    return sol / scale, route


MAGNETIC_FIELDS = ("b", "b_e", "b_n", "b_u")


def prism_magnetic(coordinates, prisms, magnetization, field, n_threads=0):
    """Oracle of ``harmonica.prism_magnetic`` (magnetic.py:28-134) without the
    input checks: null prisms (zero volume or zero magnetisation) discarded
    as magnetic.py:403-440 does, nT (magnetic.py:195-197,271), the tuple
    (b_e, b_n, b_u) for ``field="b"``."""
    shape, (e, n, u) = _coords(coordinates)
    prisms = _f64(np.atleast_2d(prisms))
    mag = np.stack([_f64(np.atleast_1d(m).ravel()) for m in magnetization])
    w, ea, s_, no, b, t = (prisms[:, i] for i in range(6))
    null = (w == ea) | (s_ == no) | (b == t) | np.all(mag == 0, axis=0)
    prisms = _f64(prisms[~null])
    mag = _f64(mag[:, ~null])
    comp = 7 if field == "b" else 1 << ("b_e", "b_n", "b_u").index(field)
    rows = 3 if field == "b" else 1
    out = np.zeros((rows, e.size))
    lib().or_prism_magnetic(_p(e), _p(n), _p(u), e.size, _p(prisms), _p(mag),
                            prisms.shape[0], comp, _p(out), n_threads)
    out *= 1e9
    if field == "b":
        return tuple(o.reshape(shape) for o in out)
    return out[0].reshape(shape)


def prism_magnetic_pair(point, prism, magnetization):
    """(b_e, b_n, b_u) in T of one prism at one point."""
    prism = _f64(prism)
    b = np.zeros(3)
    lib().or_prism_magnetic_pair(*(float(v) for v in point), _p(prism),
                                 *(float(v) for v in magnetization), _p(b))
    return b


def rolling_window_windows(coordinates, window_size, overlap, region):
    """Window membership by direct comparison (no KD-tree): the points with
    |easting - c_e| <= size / 2 and |northing - c_n| <= size / 2 for every
    window centre, centres on a northing-major grid spanning the region
    shrunk by half a window, (1 - overlap) * size apart rounded to a whole
    number of steps (bordado.rolling_window as gradient_boosted.py:343-354
    calls it)."""
    w, e, s, n = region
    half = window_size / 2
    spacing = (1 - overlap) * window_size
    axes = []
    for lo, hi in ((w + half, e - half), (s + half, n - half)):
        axes.append(np.linspace(lo, hi, int(round((hi - lo) / spacing)) + 1))
    east, north = (np.asarray(c, dtype=np.float64).ravel() for c in coordinates[:2])
    windows = []
    for cn in axes[1]:
        for ce in axes[0]:
            inside = (np.abs(east - ce) <= half) & (np.abs(north - cn) <= half)
            windows.append(np.nonzero(inside)[0])
    return windows


def eqs_gb_fit(coordinates, data, weights, points, damping, source_windows, data_windows):
    """EquivalentSourcesGB._gradient_boosting (gradient_boosted.py:231-283)
    on the CPU for the given windows: per window the jacobian, verde's least
    squares on the current residue, the prediction on every datum
    subtracted, the coefficients accumulated.  Returns (coefs, rmse per
    iteration)."""
    coords = tuple(_f64(np.asarray(c).ravel()) for c in coordinates[:3])
    pts = tuple(_f64(np.asarray(p).ravel()) for p in points[:3])
    residue = _f64(np.asarray(data).ravel()).copy()
    coefs = np.zeros(pts[0].size)
    errors = [np.sqrt(np.mean(residue**2))]
    for sw, dw in zip(source_windows, data_windows):
        pchunk = tuple(p[sw] for p in pts)
        cchunk = tuple(c[dw] for c in coords)
        wchunk = None if weights is None else np.asarray(weights).ravel()[dw]
        jac = eqs_jacobian(cchunk, pchunk)
        cw = eqs_fit_coefs(jac, residue[dw], wchunk, damping)
        residue -= eqs_predict(coords, pchunk, cw)
        errors.append(np.sqrt(np.mean(residue**2)))
        coefs[sw] += cw
    return coefs, np.array(errors)


#: Distance-size ratios of the adaptive discretisation (tesseroid_gravity.py:32).
TESSEROID_RATIOS = {"potential": 1.0, "g_z": 2.5}
#: Mean radius of the WGS84 ellipsoid (boule.WGS84.mean_radius), used by the
#: reference's tesseroid tests.
WGS84_MEAN_RADIUS = 6371008.771415059


def tesseroid_gravity(coordinates, tesseroids, density, field, radial=False, n_threads=0,
                      continuity=True):
    """Oracle of ``harmonica.tesseroid_gravity`` (tesseroid_gravity.py:35-213)
    for constant densities, without the input checks: west > east boundaries
    moved to [-180, 180) (_tesseroid_utils.py:_longitude_continuity, part of
    the checks; ``continuity=False`` leaves them as they are, as the
    reference does with disable_checks=True), null tesseroids discarded, g_z
    flipped and in mGal.  ``density`` may also be
    an (m, 2) array of densities at the lower and upper radial quadrature
    nodes of each tesseroid (variable density without radial
    discretisation).  Raises OverflowError as the reference does."""
    shape, (lon, lat, rad) = _coords(coordinates)
    t = _f64(np.atleast_2d(tesseroids)).copy()
    swap = (t[:, 0] > t[:, 1]) & bool(continuity)
    t[swap, 1] = ((t[swap, 1] + 180) % 360) - 180
    t[swap, 0] = ((t[swap, 0] + 180) % 360) - 180
    rho = np.asarray(density, dtype=np.float64)
    rho = np.repeat(np.atleast_1d(rho).ravel()[:, None], 2, axis=1) if rho.ndim < 2 else rho
    null = ((t[:, 0] == t[:, 1]) | (t[:, 2] == t[:, 3]) | (t[:, 4] == t[:, 5])
            | np.all(rho == 0, axis=1))
    t, rho = _f64(t[~null]), _f64(rho[~null])
    out = np.zeros(lon.size)
    st = lib().or_tesseroid_gravity(_p(lon), _p(lat), _p(rad), lon.size, _p(t), _p(rho),
                                    t.shape[0], OR_FIELD_TESS[field], TESSEROID_RATIOS[field],
                                    int(bool(radial)), _p(out), n_threads)
    if st == 1:
        raise OverflowError("Stack Overflow. Try to increase the stack size.")
    if st == 2:
        raise OverflowError("Exceeded maximum discretizations. Please increase the "
                            "MAX_DISCRETIZATIONS.")
    if field == "g_z":
        out *= -1e5
    return out.reshape(shape)


OR_FIELD_TESS = {"potential": 0, "g_z": 3}


def tesseroid_split_list(point, tesseroid, ratio, radial=False, stack_size=100, max_out=100000):
    """Small tesseroids of one pair in production order
    (_tesseroid_utils.py:_adaptive_discretization); raises OverflowError."""
    out = np.zeros((max_out, 6))
    t = _f64(tesseroid)
    n = lib().or_tesseroid_split_list(float(point[0]), float(point[1]), float(point[2]), _p(t),
                                      float(ratio), int(bool(radial)), int(stack_size),
                                      _p(out), int(max_out))
    if n < 0:
        raise OverflowError("discretisation overflow")
    return out[:n]


def tesseroid_gravity_variable(coordinates, tesseroids, density, field, radial=False):
    """Oracle of the variable-density loop (tesseroid_gravity.py:339-452)
    for small cases, in Python: per (point, tesseroid) the small tesseroids
    of the adaptive discretisation (tesseroid_split_list, the C
    restatement), each integrated by 2 x 2 x 2 Gauss-Legendre nodes with the
    density FUNCTION evaluated at its radial nodes
    (_tesseroid_variable_density.py gauss_legendre_quadrature_variable_density)
    and the spherical point-mass kernels (point.py:324-354) at the distance
    of _forward/utils.py:163-201.  ``tesseroids`` must already be split by
    the density-based discretisation (or need none).  g_z flipped, mGal."""
    node = np.sqrt(1.0 / 3.0) * np.array([-1.0, 1.0])
    ratio = TESSEROID_RATIOS[field]
    lon, lat, rad = (np.atleast_1d(np.asarray(c, dtype=np.float64)).ravel() for c in coordinates[:3])
    out = np.zeros(lon.size)
    deg = np.pi / 180
    for i in range(lon.size):
        lon_r, sinphi, cosphi, r = lon[i] * deg, np.sin(lat[i] * deg), np.cos(lat[i] * deg), rad[i]
        for t in np.atleast_2d(tesseroids):
            for w, e, s, n, bottom, top in tesseroid_split_list((lon[i], lat[i], r), t, ratio, radial):
                a = 1.0 / 8.0 * ((e - w) * deg) * ((n - s) * deg) * (top - bottom)
                for lat_n in 0.5 * (n - s) * node + 0.5 * (n + s):
                    sp, cp = np.sin(lat_n * deg), np.cos(lat_n * deg)
                    for r_n in 0.5 * (top - bottom) * node + 0.5 * (top + bottom):
                        mass = density(r_n) * a * r_n * r_n * cp
                        for lon_n in 0.5 * (e - w) * node + 0.5 * (e + w):
                            cospsi = sp * sinphi + cp * cosphi * np.cos(lon_n * deg - lon_r)
                            dist = np.sqrt((r - r_n) ** 2 + 2.0 * r * r_n * (1.0 - cospsi))
                            if field == "potential":
                                out[i] += mass * GRAVITATIONAL_CONST / dist
                            else:
                                out[i] += mass * (-GRAVITATIONAL_CONST) * (r - r_n * cospsi) / dist**3
    if field == "g_z":
        out *= -1e5
    return out


def num_threads():
    return lib().or_num_threads()
