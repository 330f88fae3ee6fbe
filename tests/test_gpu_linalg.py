"""
The MFMA and Cholesky building blocks of the equivalent-source fit:
operand/result lane layout of the 16x16x4 MFMA builtins, the MFMA GEMM
(C = beta C + sign A^T B + shift I, full and upper-triangular) against
numpy, and the blocked FP64 Cholesky solve against numpy.
"""

import ctypes

import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from harmonica_amd import _lib

    return torch, _lib, _lib.lib()


def _layouts(a, b, dtype):
    """Result D (16 x 16) of one MFMA under the layout the kernels assume:
    lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; element v of
    lane l holds D[row][l % 16] with row = l / 16 + 4 v (f64) or
    4 (l / 16) + v (f32)."""
    A = np.zeros((16, 4))
    B = np.zeros((4, 16))
    for l in range(64):
        A[l % 16, l // 16] = a[l]
        B[l // 16, l % 16] = b[l]
    D = A @ B
    out = np.zeros(256)
    for l in range(64):
        for v in range(4):
            row = l // 16 + 4 * v if dtype == 0 else 4 * (l // 16) + v
            out[4 * l + v] = D[row, l % 16]
    return out


@pytest.mark.parametrize("dtype", [0, 1], ids=["f64", "f32"])
def test_mfma_layout(env, dtype):
    torch, _lib, L = env
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 9, 64).astype(float)
    b = rng.integers(-8, 9, 64).astype(float)
    A, B = torch.tensor(a).cuda(), torch.tensor(b).cuda()
    out = torch.empty(256, dtype=torch.float64, device="cuda")
    _lib.check(L.hmx_mfma_probe(dtype, _lib.ptr(A), _lib.ptr(B), _lib.ptr(out), None))
    got = out.cpu().numpy()
    expected = _layouts(a, b, dtype)
    assert np.array_equal(got, expected), f"matching layouts: {_which(a, b, got)}"


def test_mfma_layout_bf16(env):
    """16x16x32 bf16 MFMA: lane l supplies row l % 16 of A and column l % 16
    of B, 8 values each of the k block l / 16; element v of its result is
    D[4 (l / 16) + v][l % 16].  The order of k inside the 32 is invisible
    in the product as long as A and B use the same one (the Gram kernel
    does), so both orders are accepted."""
    torch, _lib, L = env
    rng = np.random.default_rng(1)
    a = rng.integers(-8, 9, 512).astype(float)
    b = rng.integers(-8, 9, 512).astype(float)
    A, B = torch.tensor(a).cuda(), torch.tensor(b).cuda()
    out = torch.empty(256, dtype=torch.float64, device="cuda")
    _lib.check(L.hmx_mfma_probe(2, _lib.ptr(A), _lib.ptr(B), _lib.ptr(out), None))
    got = out.cpu().numpy()
    layouts = {}
    for kname, kmap in (("k=8(l/16)+e", lambda l, e: 8 * (l // 16) + e),
                        ("k=(l/16)+4e", lambda l, e: l // 16 + 4 * e)):
        Am, Bm = np.zeros((16, 32)), np.zeros((32, 16))
        for l in range(64):
            for e in range(8):
                Am[l % 16, kmap(l, e)] = a[8 * l + e]
                Bm[kmap(l, e), l % 16] = b[8 * l + e]
        D = Am @ Bm
        for dname, dmap in (("D[4(l/16)+v][l%16]", lambda l, v: (4 * (l // 16) + v, l % 16)),
                            ("D[(l/16)+4v][l%16]", lambda l, v: (l // 16 + 4 * v, l % 16))):
            layouts[(kname, dname)] = np.array([D[dmap(l, v)] for l in range(64) for v in range(4)])
    matches = [k for k, v in layouts.items() if np.array_equal(v, got)]
    assert ("k=8(l/16)+e", "D[4(l/16)+v][l%16]") in matches, matches
    assert all(d == "D[4(l/16)+v][l%16]" for _, d in matches), matches


def _which(a, b, got):
    """Diagnose: which (A map, B map, D map) combinations reproduce `got`."""
    amaps = {"A[l%16][l/16]": lambda l: (l % 16, l // 16), "A[l/4][l%4]": lambda l: (l // 4, l % 4)}
    bmaps = {"B[l/16][l%16]": lambda l: (l // 16, l % 16), "B[l%4][l/4]": lambda l: (l % 4, l // 4)}
    dmaps = {
        "D[4(l/16)+v][l%16]": lambda l, v: (4 * (l // 16) + v, l % 16),
        "D[l%16][4(l/16)+v]": lambda l, v: (l % 16, 4 * (l // 16) + v),
        "D[(l/16)+4v][l%16]": lambda l, v: (l // 16 + 4 * v, l % 16),
    }
    found = []
    for an, am in amaps.items():
        for bn, bm in bmaps.items():
            A = np.zeros((16, 4))
            B = np.zeros((4, 16))
            for l in range(64):
                A[am(l)] = a[l]
                B[bm(l)] = b[l]
            D = A @ B
            for dn, dmap in dmaps.items():
                out = np.array([D[dmap(l, v)] for l in range(64) for v in range(4)])
                if np.array_equal(out, got):
                    found.append((an, bn, dn))
    return found


def _gemm(env, dtype, A, B, C, beta, sign, shift, upper):
    torch, _lib, L = env
    tdt = torch.float64 if dtype == 0 else torch.float32
    dA = torch.tensor(A, dtype=tdt).cuda()
    dB = torch.tensor(B, dtype=tdt).cuda()
    dC = torch.tensor(C, dtype=torch.float64).cuda()
    K, M = A.shape
    N = B.shape[1]
    _lib.check(L.hmx_gemm_tn(dtype, _lib.ptr(dA), M, _lib.ptr(dB), N, K, M, N, _lib.ptr(dC), N,
                             beta, sign, shift, int(upper), None))
    return dC.cpu().numpy()


@pytest.mark.parametrize("shape", [(37, 19, 23), (200, 130, 260), (16, 128, 128), (5, 300, 7)])
@pytest.mark.parametrize("dtype", [0, 1], ids=["f64", "f32"])
def test_gemm_tn_full(env, shape, dtype):
    K, M, N = shape
    rng = np.random.default_rng(K + M + N)
    A, B, C = rng.normal(size=(K, M)), rng.normal(size=(K, N)), rng.normal(size=(M, N))
    if dtype == 1:
        A, B = A.astype(np.float32).astype(float), B.astype(np.float32).astype(float)
    got = _gemm(env, dtype, A, B, C, 0.5, -1.0, 0.0, False)
    expected = 0.5 * C - A.T @ B
    tol = 1e-12 if dtype == 0 else 1e-5
    npt.assert_allclose(got, expected, atol=tol * np.abs(expected).max())


@pytest.mark.parametrize("dtype", [0, 1], ids=["f64", "f32"])
@pytest.mark.parametrize("m", [1, 77, 128, 300])
def test_syrk_upper(env, dtype, m):
    rng = np.random.default_rng(m)
    X = rng.normal(size=(150, m))
    if dtype == 1:
        X = X.astype(np.float32).astype(float)
    C = np.full((m, m), 7.0)
    got = _gemm(env, dtype, X, X, C, 0.0, 1.0, 2.5, True)
    expected = X.T @ X + 2.5 * np.eye(m)
    iu = np.triu_indices(m)
    tol = 1e-12 if dtype == 0 else 1e-5
    npt.assert_allclose(got[iu], expected[iu], atol=tol * np.abs(expected).max())
    il = np.tril_indices(m, -1)
    assert np.all(got[il] == 7.0)  # lower triangle untouched


@pytest.mark.parametrize("m", [5, 128, 129, 300, 700, 1500, 2100, 4500])
def test_cholesky_solve(env, m):
    """Sizes below, across and beyond the 1024-row outer block (the rank-1024
    trailing update runs for m > 1024; beyond two outer blocks the update of
    the rows past the next outer block runs on a second stream: once at
    2100, three times at 4500)."""
    torch, _lib, L = env
    rng = np.random.default_rng(m)
    X = rng.normal(size=(m + 20, m))
    A = X.T @ X + 0.5 * np.eye(m)
    rhs = rng.normal(size=m)
    nbytes = ctypes.c_size_t(0)
    _lib.check(L.hmx_cholesky_workspace_bytes(m, ctypes.byref(nbytes)))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    dA = torch.tensor(A).cuda()
    dx = torch.tensor(rhs).cuda()
    info = ctypes.c_int(-1)
    _lib.check(L.hmx_cholesky_solve(_lib.ptr(dA), m, _lib.ptr(dx), ctypes.byref(info), _lib.ptr(ws),
                                    nbytes.value, None))
    assert info.value == 0
    expected = np.linalg.solve(A, rhs)
    npt.assert_allclose(dx.cpu().numpy(), expected, rtol=1e-9, atol=1e-11 * np.abs(expected).max())
    U = np.triu(dA.cpu().numpy())
    npt.assert_allclose(U.T @ U, A, atol=1e-10 * np.abs(A).max())


@pytest.mark.parametrize("m,bad", [(200, 150), (1300, 1200), (4500, 3000)])
def test_cholesky_reports_indefinite(env, m, bad):
    """The failed pivot's column; at 4500 the factorisation stops early while
    an update on the second stream may still run (it is waited for)."""
    torch, _lib, L = env
    A = np.eye(m)
    A[bad, bad] = -1.0
    nbytes = ctypes.c_size_t(0)
    _lib.check(L.hmx_cholesky_workspace_bytes(m, ctypes.byref(nbytes)))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    dA, dx = torch.tensor(A).cuda(), torch.ones(m, dtype=torch.float64, device="cuda")
    info = ctypes.c_int(-1)
    _lib.check(L.hmx_cholesky_solve(_lib.ptr(dA), m, _lib.ptr(dx), ctypes.byref(info), _lib.ptr(ws),
                                    nbytes.value, None))
    assert info.value == bad + 1


def _lsq_svd(env, G, y, damping, filt, cutoff):
    """hmx_lsq_svd on a host matrix: (coef, singular values, info)."""
    torch, _lib, L = env
    n, m = G.shape
    g = torch.tensor(np.asfortranarray(G).ravel(order="F")).cuda()
    yd = torch.tensor(y).cuda()
    coef = torch.empty(m, dtype=torch.float64, device="cuda")
    sig = torch.empty(m, dtype=torch.float64, device="cuda")
    nbytes = ctypes.c_size_t(0)
    _lib.check(L.hmx_lsq_svd_workspace_bytes(n, m, ctypes.byref(nbytes)))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    info = (ctypes.c_double * 2)()
    _lib.check(L.hmx_lsq_svd(_lib.ptr(g), n, m, _lib.ptr(yd), damping, filt, cutoff,
                             _lib.ptr(coef), _lib.ptr(sig), info, _lib.ptr(ws), nbytes.value,
                             None))
    return coef.cpu().numpy(), sig.cpu().numpy(), list(info)


def _filtered_solution(G, y, damping, filt, cutoff):
    U, s, Vt = np.linalg.svd(G, full_matrices=False)
    f = np.zeros_like(s)
    if filt == 0:
        keep = s > cutoff
        f[keep] = s[keep] / (s[keep] ** 2 + damping)
    elif filt == 1:
        keep = s > cutoff * s[0]
        f[keep] = 1 / s[keep]
    else:
        keep = s ** 2 + damping > cutoff * (s[0] ** 2 + damping)
        f[keep] = s[keep] / (s[keep] ** 2 + damping)
    return Vt.T @ (f * (U.T @ y)), s, int(keep.sum())


@pytest.mark.parametrize("shape", [(40, 40), (257, 130), (90, 201), (1, 3), (5, 1)])
@pytest.mark.parametrize("filt", [0, 1, 2])
def test_lsq_svd_random(env, shape, filt):
    """Well-conditioned random matrices (tall, square, wide, degenerate
    shapes): singular values against LAPACK's (sorted) and the filtered
    solution of each filter against the same formula on numpy's SVD."""
    rng = np.random.default_rng(shape[0] * 1000 + shape[1] + filt)
    G = rng.normal(size=shape)
    y = rng.normal(size=shape[0])
    damping, cutoff = {0: (0.3, 1e-15), 1: (0.0, 1e-12), 2: (0.3, 2.0 ** -52)}[filt]
    coef, sig, info = _lsq_svd(env, G, y, damping, filt, cutoff)
    expected, s, rank = _filtered_solution(G, y, damping, filt, cutoff)
    assert info[0] > 0 and int(info[1]) == rank
    k = min(shape)
    npt.assert_allclose(np.sort(sig)[::-1][:k], s, rtol=1e-12, atol=1e-13 * s[0])
    if shape[1] > shape[0]:  # the extra columns are rotated to zero
        assert np.sort(sig)[::-1][k:].max(initial=0) <= 1e-13 * s[0]
    npt.assert_allclose(coef, expected, rtol=1e-10, atol=1e-11 * np.abs(expected).max())


def test_lsq_svd_graded_and_rank_deficient(env):
    """Singular values spread over 12 decades (recovered to roundoff of the
    largest) and three exactly repeated columns: the lstsq filter drops the
    repeated directions to give the minimum-norm solution (repeated columns
    share the weight)."""
    rng = np.random.default_rng(3)
    n, m = 150, 60
    Q1, _ = np.linalg.qr(rng.normal(size=(n, m)))
    Q2, _ = np.linalg.qr(rng.normal(size=(m, m)))
    s_true = np.logspace(0, -12, m)
    G = (Q1 * s_true) @ Q2.T
    coef, sig, info = _lsq_svd(env, G.copy(), rng.normal(size=n), 0.0, 1, 1e-15)
    npt.assert_allclose(np.sort(sig)[::-1], s_true, rtol=1e-9, atol=1e-14)
    G = (Q1 * np.logspace(0, -3, m)) @ Q2.T
    H = np.concatenate([G[:, :40], G[:, :3]], axis=1)
    y = rng.normal(size=n)
    coef, sig, info = _lsq_svd(env, H, y, 0.0, 1, max(H.shape) * 2.0 ** -52)
    expected = np.linalg.lstsq(H, y, rcond=max(H.shape) * 2.0 ** -52)[0]
    assert int(info[1]) == 40
    npt.assert_allclose(H @ coef, H @ expected, rtol=0, atol=1e-10 * np.abs(H @ expected).max())
    npt.assert_allclose(coef[:3], coef[40:], rtol=1e-8)
    npt.assert_allclose(coef, expected, rtol=1e-6, atol=1e-8 * np.abs(expected).max())
