"""
Benchmark: prism-observation pair evaluations per second of the FP64 g_z
forward model on one MI355X, on the 1,000,000-prism terrain workload
(1000 x 1000 prisms of 100 m on a 100 km square, fractal tops in
[0, 2000] m, bottoms at -5000 m, densities U(2600, 2800); a 256 x 256 grid
of observation points at 2500 m: 6.5536e10 pairs per step).

One step = one hmx_prism_gravity call (the drop-in path of
prism_gravity(coordinates, prisms, density, field="g_z")) with inputs
resident in HBM.  Prints one JSON line (see the fields below), including
the roofline of the dominant kernel and the CPU oracle timed on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--extra]
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X FP64 vector peak (AMD published: 256 CU x 128 FP64 flop/clk x 2.4 GHz).
FP64_PEAK = 78.6e12

# Algorithmic FP64 operations per prism-observation pair of the g_z closed
# form, frozen from rocprofv3 FP64 VALU counters of the first (ocml-based)
# kernel version: 8 vertices x (sqrt + 2 log + atan + division + the
# arithmetic of the kernel), FMA = 2.  Kept fixed so later, leaner
# implementations are credited in pair-evals/s, not in flop count.
FLOPS_PER_PAIR = {"g_z": None, "all": None}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, prisms, density):
    """The CPU oracle (C + OpenMP restatement of the reference's parallel
    Numba loop) on a bounded sample: every 8th observation row and column
    (32 x 32 = 1024 points) against all 1,000,000 prisms."""
    import oracle

    E, N, U = (c[::8, ::8] for c in w["coordinates"])
    cores = len(os.sched_getaffinity(0))
    threads = min(cores, 16)
    pairs = E.size * prisms.shape[0]
    t0 = time.perf_counter()
    oracle.prism_gravity((E, N, U), prisms, density, "g_z", n_threads=threads)
    dt = time.perf_counter() - t0
    return {"value": pairs / dt, "unit": "pair-evals/s", "cores": threads, "kind": "port",
            "sample": f"{E.size} of {w['coordinates'][0].size} observation points (every 8th row/column) x "
                      f"{prisms.shape[0]} prisms = {pairs:.3e} pairs, g_z FP64, {dt:.1f} s",
            "host_cores_visible": cores}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--extra", action="store_true", help="also time the fused ten-field launch")
    ap.add_argument("--obs-side", type=int, default=256)
    ap.add_argument("--prism-side", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path's "
              "multi-GPU sharding is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch

    from harmonica_amd import _lib
    from harmonica_amd._forward import ctypes_size, to_device
    from harmonica_amd.workloads import layer_to_prisms, terrain_layer_large

    assert torch.cuda.is_available(), "no GPU"
    log("generating workload")
    w = terrain_layer_large(n_prisms_side=args.prism_side, n_obs_side=args.obs_side)
    prisms = layer_to_prisms(w["easting"], w["northing"], w["surface"], w["reference"])
    density = w["density"].ravel()
    E, N, U = (to_device(c.ravel()) for c in w["coordinates"])
    P, D = to_device(prisms), to_device(density)
    n_obs, n_prisms = E.numel(), P.shape[0]
    pairs = n_obs * n_prisms
    L = _lib.lib()
    stream = torch.cuda.current_stream()
    sh = _lib.stream_handle(stream)

    def make_step(mask, nf):
        nbytes = ctypes_size(n_obs, n_prisms, mask)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=E.device)
        out = torch.empty((nf, n_obs), dtype=torch.float64, device=E.device)

        def step():
            _lib.check(L.hmx_prism_gravity(_lib.ptr(E), _lib.ptr(N), _lib.ptr(U), n_obs,
                                           _lib.ptr(P), _lib.ptr(D), n_prisms, mask,
                                           _lib.ptr(out), None, _lib.ptr(ws), nbytes, sh))
        return step, out

    def timed(step, steps, warmup):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return ev0.elapsed_time(ev1) / steps, wall

    log(f"g_z: {n_obs} points x {n_prisms} prisms, warmup {args.warmup}, steps {args.steps}")
    step, out = make_step(_lib.FIELD_BIT["g_z"], 1)
    ms, wall = timed(step, args.steps, args.warmup)
    value = pairs / (ms * 1e-3)
    log(f"g_z: {ms:.2f} ms/step, {value:.4e} pair-evals/s (wall {wall:.2f} s)")

    fpp = FLOPS_PER_PAIR["g_z"]
    roofline = None
    if fpp:
        achieved = value * fpp
        roofline = {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12,
                    "unit": "TFLOP/s", "frac": achieved / FP64_PEAK, "traffic": None,
                    "flops_per_pair": fpp}

    extra = {}
    if args.extra:
        mask = (1 << 10) - 1
        step10, _ = make_step(mask, 10)
        ms10, _ = timed(step10, 1, 1)
        extra["all_ten_fields"] = {"ms_per_step": ms10, "pair_evals_per_s": pairs / (ms10 * 1e-3)}
        log(f"ten fields fused: {ms10:.2f} ms/step")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(w, prisms, density)
        log(f"cpu: {cpu['value']:.4e} pair-evals/s on {cpu['cores']} threads")

    line = {
        "metric": "pair-evals/s g_z FP64 (1M prisms, % FP64 peak)",
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"terrain layer {args.prism_side}x{args.prism_side} prisms "
                               f"(100 km square, bottoms -5000 m) x {args.obs_side}x{args.obs_side} "
                               "grid at 2500 m, g_z, prism_gravity path",
                   "n_prisms": n_prisms, "n_obs": n_obs, "pairs_per_step": pairs},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if extra:
        line["extra"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
