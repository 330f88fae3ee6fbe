"""
Parity at the full sizes of the benchmark workloads that the other GPU
tests cover only at small sizes:

* 512 batched uneven models (log-uniform 1e3-2e5 prisms and 1e3-5e4
  points, g_z): every model of the one persistent launch against the
  single-model path (prism_gravity, src/harmonica/_forward/prisms/
  gravity.py:51-236) to 1e-12 of its max|g_z|, every 32nd model (every 8th
  point) against the oracle to 1e-10.
* the 60,000-datum equivalent-source fit (damping 1e-3, sources 4500 m
  below the data; cartesian.py:236-281): the mixed-precision fit against the
  FP64 fit on the 512 x 512 grid to the 1e-6 prediction bar, and the FP64
  fit of the first 20,000 data against the oracle's LAPACK solve of the same
  least-squares problem (predictions on a 128 x 128 grid).
* the 200 x 200 basement layer with 128 x 128 points: autograd gradients
  against central differences (steps 1e-3 m and 1e-4 kg/m^3, FP64) on
  sampled prisms, the adjoint identity, and 20 Gauss-Newton iterations.
"""

import numpy as np
import numpy.testing as npt
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_batched_512_models_full_size(torch_gpu):
    import harmonica_amd as hm
    from harmonica_amd._batched import BatchedPrismGz
    from harmonica_amd.workloads import batched_models

    models = batched_models()
    assert len(models) == 512
    batch = BatchedPrismGz(models)
    batch.run()
    res = batch.split()
    worst_single = 0.0
    for b, (c, p, d) in enumerate(models):
        ref = hm.prism_gravity(c, p, d, field="g_z")
        err = float(np.abs(res[b] - ref).max() / np.abs(ref).max())
        worst_single = max(worst_single, err)
        assert err <= 1e-12, (b, err)
    worst_oracle = 0.0
    for b in range(0, 512, 32):
        c, p, d = models[b]
        sub = tuple(x[::8] for x in c)
        ref = oracle.prism_gravity(sub, p, d, "g_z")
        err = float(np.abs(res[b][::8] - ref).max() / np.abs(ref).max())
        worst_oracle = max(worst_oracle, err)
        assert err <= 1e-10, (b, err)
    print(f"batched 512 models: {batch.pairs:.3e} pairs, vs single model {worst_single:.2e}, "
          f"vs oracle {worst_oracle:.2e}", flush=True)


@pytest.fixture(scope="module")
def eqs_60k(torch_gpu):
    import harmonica_amd as hm
    from harmonica_amd.workloads import eqs_problem

    w = eqs_problem()
    data = hm.prism_gravity(w["coordinates"], w["prisms"], w["density"], field="g_z") + w["noise"]
    return w, data


def test_eqs_60k_mixed_vs_fp64(eqs_60k):
    import harmonica_amd as hm

    w, data = eqs_60k
    preds = {}
    for mode in ("float64", "mixed"):
        eqs = hm.EquivalentSources(damping=w["damping"], depth=w["depth"], fit_precision=mode)
        eqs.fit(w["coordinates"], data)
        preds[mode] = eqs.predict(w["grid"])
        print(f"eqs 60k {mode}: {eqs.fit_info_}", flush=True)
    ref = preds["float64"]
    err = float(np.abs(preds["mixed"] - ref).max() / np.abs(ref).max())
    print(f"eqs 60k mixed vs FP64 on the 512 x 512 grid: {err:.2e}", flush=True)
    assert err <= 1e-6


def test_eqs_20k_fp64_vs_oracle(eqs_60k):
    import harmonica_amd as hm

    w, data = eqs_60k
    n = 20000
    sub = tuple(c[:n] for c in w["coordinates"])
    eqs = hm.EquivalentSources(damping=w["damping"], depth=w["depth"]).fit(sub, data[:n])
    J = oracle.eqs_jacobian(sub, eqs.points_)
    coefs = oracle.eqs_fit_coefs(J, data[:n], None, w["damping"])
    del J
    grid = tuple(c[::4, ::4] for c in w["grid"])
    p_gpu = eqs.predict(grid)
    p_cpu = oracle.eqs_predict(grid, eqs.points_, coefs)
    err = float(np.abs(p_gpu - p_cpu).max() / np.abs(p_cpu).max())
    # Both are FP64 solves of a least-squares problem whose scaled jacobian
    # has a condition number near 1e5 (normal matrix ~1e11): they agree to
    # about cond(A) x 2^-53 of the predictions, far below the 1e-6 bar.
    print(f"eqs 20k FP64 vs oracle: {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.fixture(scope="module")
def inversion(torch_gpu):
    from harmonica_amd import _autograd
    from harmonica_amd.workloads import inversion_layer

    w = inversion_layer()
    geom = _autograd.LayerGeometry(w["easting"], w["northing"], w["bottom"], w["coordinates"])
    dev = lambda a: torch_gpu.tensor(a, dtype=torch_gpu.float64, device="cuda")
    return torch_gpu, _autograd, w, geom, dev


@pytest.mark.parametrize("weights", ["ones", "normal"])
def test_inversion_layer_gradients_full_size(inversion, weights):
    """Gradients of sum_i v_i g_z(i) against central differences of the GPU
    forward model, FP64, steps 1e-4 kg/m^3 and 1e-3 m, on six prisms.  With
    v = 1 the sum does not cancel and the bars are relative to the
    derivative; with v ~ N(0, 1) it cancels, and the differences' error is
    bounded against the scale sum_i |v_i| |d g_z(i)| (the rounding of g_z
    at points up to 50 km away, divided by the step)."""
    import harmonica_amd as hm

    torch, ag, w, geom, dev = inversion
    assert geom.n_prisms == 40000 and geom.n_obs == 16384
    rng = np.random.default_rng(7)
    v = np.ones(geom.n_obs) if weights == "ones" else rng.normal(size=geom.n_obs)
    rho_t = dev(w["density"]).requires_grad_(True)
    top_t = dev(w["top"]).requires_grad_(True)
    (ag.layer_gz(rho_t, top_t, geom) * dev(v)).sum().backward()
    pts = [c.ravel() for c in w["coordinates"]]
    he = (w["easting"][1] - w["easting"][0]) / 2
    worst = 0.0
    for k, j in ((0, 0), (10, 20), (100, 100), (150, 37), (190, 5), (199, 199)):
        p = np.array([w["easting"][j] - he, w["easting"][j] + he, w["northing"][k] - he,
                      w["northing"][k] + he, w["bottom"][k, j], w["top"][k, j]])
        d = w["density"][k, j]
        dr = (hm.prism_gravity(pts, p, d + 1e-4, "g_z") - hm.prism_gravity(pts, p, d - 1e-4, "g_z")) / 2e-4
        pu, pd = p.copy(), p.copy()
        pu[5] += 1e-3
        pd[5] -= 1e-3
        dt = (hm.prism_gravity(pts, pu, d, "g_z") - hm.prism_gravity(pts, pd, d, "g_z")) / 2e-3
        fr, ft = v @ dr, v @ dt
        g_r, g_t = rho_t.grad[k, j].item(), top_t.grad[k, j].item()
        if weights == "ones":
            # density: g_z is linear in it, the differences are exact to
            # rounding; top: the second-order term h^2/6 times the third
            # derivative plus rounding over the step
            npt.assert_allclose(g_r, fr, rtol=1e-8)
            npt.assert_allclose(g_t, ft, rtol=2e-6)
            err = max(abs(g_r - fr) / abs(fr), abs(g_t - ft) / abs(ft))
        else:
            scale_r, scale_t = np.abs(v) @ np.abs(dr), np.abs(v) @ np.abs(dt)
            assert abs(g_r - fr) <= 1e-8 * scale_r
            assert abs(g_t - ft) <= 2e-6 * scale_t
            err = max(abs(g_r - fr) / scale_r, abs(g_t - ft) / scale_t)
        worst = max(worst, err)
    print(f"inversion layer ({weights}): worst gradient vs central differences {worst:.2e}",
          flush=True)


def test_inversion_layer_adjoint_identity(inversion):
    torch, ag, w, geom, dev = inversion
    rng = np.random.default_rng(8)
    x_rho, x_top = rng.normal(size=w["density"].shape), rng.normal(size=w["top"].shape)
    y = rng.normal(size=geom.n_obs)
    rho, top = dev(w["density"]), dev(w["top"])
    jx = geom.jvp(rho, top, dev(x_rho), dev(x_top)).cpu().numpy()
    a_rho, a_top = (t.cpu().numpy() for t in geom.vjp(rho, top, dev(y)))
    lhs, rhs = jx @ y, (x_rho * a_rho).sum() + (x_top * a_top).sum()
    npt.assert_allclose(lhs, rhs, rtol=1e-10)


def test_inversion_layer_gauss_newton_20_iterations(inversion):
    torch, ag, w, geom, dev = inversion
    rho, top = dev(w["density"]), dev(w["top"])
    data = geom.forward(rho, top)
    # the forward model of this layer against the oracle's layer loop on a
    # sample of the points
    pts = tuple(c.ravel()[::97] for c in w["coordinates"])
    ref = oracle.prism_layer_gravity(pts, w["easting"], w["northing"], w["bottom"], w["top"],
                                     w["density"], "g_z")
    got = data.cpu().numpy()[::97]
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
    rho0 = torch.full_like(rho, -300.0)
    top0 = torch.full_like(top, -3500.0)
    r, t, hist = ag.gauss_newton(geom, data, rho0, top0, n_iterations=20)
    print(f"Gauss-Newton misfit {hist[0]:.3g} -> {hist[-1]:.3g} mGal", flush=True)
    assert len(hist) == 21 and np.isfinite(hist).all()
    assert hist[-1] < 0.01 * hist[0]
    assert all(b <= a * 1.05 for a, b in zip(hist, hist[1:]))
