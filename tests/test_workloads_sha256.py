"""
The seeded benchmark and parity inputs are the ones whose SHA-256 is
committed in tests/golden/workload_sha256.json (written by
tests/golden/make_workload_sha256.py).
"""

import json
import os

import numpy as np
import pytest

from harmonica_amd import workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

with open(os.path.join(GOLDEN, "workload_sha256.json")) as f:
    EXPECTED = json.load(f)

MAKERS = {
    "terrain_layer_small": workloads.terrain_layer_small,
    "terrain_layer_large": workloads.terrain_layer_large,
    "terrain_magnetization": workloads.terrain_magnetization,
    "global_tesseroid_layer": workloads.global_tesseroid_layer,
    "batched_models_first32": lambda: workloads.batched_models()[:32],
    "eqs_problem": workloads.eqs_problem,
    "inversion_layer": workloads.inversion_layer,
}


def test_every_input_set_has_a_hash():
    assert set(EXPECTED) == set(MAKERS)


@pytest.mark.parametrize("name", sorted(MAKERS))
def test_workload_sha256(name):
    assert workloads.sha256(MAKERS[name]()) == EXPECTED[name]


def test_sha256_sees_one_ulp():
    w = workloads.terrain_layer_small()
    h = workloads.sha256(w)
    w["density"][37, 11] = np.nextafter(w["density"][37, 11], np.inf)
    assert workloads.sha256(w) != h
    w = workloads.terrain_layer_small()
    w["coordinates"] = w["coordinates"][::-1]
    assert workloads.sha256(w) != h
