"""
Write workload_sha256.json: the SHA-256 of every seeded input set the
benchmark and parity workloads generate (harmonica_amd.workloads.sha256),
so that a change of a generator, of its seed or of numpy's random streams
shows up as a failing test rather than as a silent change of the measured
workload.  Run from the repository root:

    python tests/golden/make_workload_sha256.py
"""

import json
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))

from harmonica_amd import workloads as w  # noqa: E402

INPUT_SETS = {
    "terrain_layer_small": lambda: w.terrain_layer_small(),
    "terrain_layer_large": lambda: w.terrain_layer_large(),
    "terrain_magnetization": lambda: w.terrain_magnetization(),
    "global_tesseroid_layer": lambda: w.global_tesseroid_layer(),
    "batched_models_first32": lambda: w.batched_models()[:32],
    "eqs_problem": lambda: w.eqs_problem(),
    "inversion_layer": lambda: w.inversion_layer(),
}


def main():
    out = {name: w.sha256(make()) for name, make in INPUT_SETS.items()}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workload_sha256.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
