"""
Batched uneven prism models in one persistent launch: per-model results
equal single-model prism_gravity (<= 1e-12 of max|g_z|) and the CPU oracle
(1e-10), with the points uploaded in strip order or in the caller's order;
empty models and null prisms; HIP-graph replays identical to the
eager run with eager work in between; autotuning keeps the results.
"""

import numpy as np
import numpy.testing as npt
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import harmonica_amd

    return harmonica_amd


def _models(seed=3, n_models=14):
    from harmonica_amd.workloads import batched_models

    models = batched_models(seed=seed, n_models=n_models, m_range=(1, 3000), n_range=(1, 2500))
    # an empty model and one with null prisms
    models.append(((np.zeros(0), np.zeros(0), np.zeros(0)), np.zeros((0, 6)), np.zeros(0)))
    c, p, d = models[0]
    p = np.vstack([p, [[0, 0, 0, 10, -10, 0]]])
    models.append((c, p, np.r_[d, 100.0]))
    return models


@pytest.mark.parametrize("strip_ratio", [0.25, None], ids=["strip_order", "caller_order"])
def test_batched_matches_single_models(hm, strip_ratio):
    from harmonica_amd._batched import BatchedPrismGz

    models = _models()
    batch = BatchedPrismGz(models, prisms_per_item=512, blocks_per_cu=4, strip_ratio=strip_ratio)
    batch.run()
    results = batch.split()
    for (c, p, d), got in zip(models, results):
        if np.size(c[0]) == 0:
            assert got.size == 0
            continue
        single = hm.prism_gravity(c, p, d, field="g_z")
        scale = max(np.abs(single).max(), 1e-300)
        npt.assert_allclose(got, single, rtol=0, atol=1e-12 * scale)
    for (c, p, d), got in list(zip(models, results))[:4]:
        ref = oracle.prism_gravity(c, p, d, "g_z")
        npt.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_graph_replay_is_stable(hm):
    import torch

    from harmonica_amd._batched import BatchedPrismGz

    batch = BatchedPrismGz(_models(seed=5, n_models=10), prisms_per_item=1024, blocks_per_cu=2)
    eager = batch.run().clone()
    batch.capture()
    for k in range(4):
        # eager work between replays (including torch reductions)
        junk = torch.randn(1 << 16, device="cuda", dtype=torch.float64).sum()
        torch.cuda.synchronize()
        replay = batch.replay().clone()
        torch.cuda.synchronize()
        npt.assert_array_equal(replay.cpu().numpy(), eager.cpu().numpy(), err_msg=f"replay {k}")
        assert np.isfinite(float(junk))


def test_autotune_keeps_results(hm, tmp_path, monkeypatch):
    from harmonica_amd._batched import BatchedPrismGz

    monkeypatch.setenv("HMX_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    batch = BatchedPrismGz(_models(seed=7, n_models=8))
    ref = batch.run().clone().cpu().numpy()
    timings = batch.autotune(prisms_per_item=(256, 2048), blocks_per_cu=(2, 4))
    assert len(timings) == 4
    npt.assert_allclose(batch.run().cpu().numpy(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    again = BatchedPrismGz(_models(seed=7, n_models=8))
    assert "cached" in again.autotune()
