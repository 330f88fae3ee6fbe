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
The second half of the metric, the equivalent-source fit of 60,000 data
(FP64 vs mixed precision at the 1e-6 prediction bar), is timed after the
headline steps and reported under "extra".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-eqs]
       [--eqs-all] [--eqs-oracle] [--extra] [--layer] [--point] [--batched] [--inversion]
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

# Frozen per-pair operation counts of the forward kernels (rocprofv3
# counters on these workloads; provenance in the JSON file and the profiles
# it names).
with open(os.path.join(ROOT, "bench", "flop_model.json")) as _fh:
    FLOP_MODEL = json.load(_fh)
FP64_PEAK = FLOP_MODEL["fp64_peak_flops"]
SIMD_CYCLES_PER_S = FLOP_MODEL["simd_cycles_per_s"]
_GZ, _ALL = FLOP_MODEL["prism_gz"], FLOP_MODEL["prism_all_ten"]
_B, _BU = FLOP_MODEL["magnetic_b"], FLOP_MODEL["magnetic_b_u"]
FLOPS_PER_PAIR_GZ, VALU_INSTR_PER_PAIR_GZ, TRANS_PER_PAIR_GZ = (
    _GZ["flops_per_pair"], _GZ["valu_instr_per_pair"], _GZ["trans_per_pair"])
PER_VERTEX_FLOPS_PER_PAIR_GZ = _GZ["per_vertex_flops_per_pair"]
FLOPS_PER_PAIR_ALL, VALU_INSTR_PER_PAIR_ALL, TRANS_PER_PAIR_ALL = (
    _ALL["flops_per_pair"], _ALL["valu_instr_per_pair"], _ALL["trans_per_pair"])
FLOPS_PER_PAIR_B, VALU_INSTR_PER_PAIR_B, TRANS_PER_PAIR_B = (
    _B["flops_per_pair"], _B["valu_instr_per_pair"], _B["trans_per_pair"])
FLOPS_PER_PAIR_BU, VALU_INSTR_PER_PAIR_BU, TRANS_PER_PAIR_BU = (
    _BU["flops_per_pair"], _BU["valu_instr_per_pair"], _BU["trans_per_pair"])
TRAFFIC_BYTES_GZ = _GZ["traffic_bytes"]


def valu_issue_frac(pair_rate, valu, trans):
    """Fraction of the SIMDs' issue cycles the pair rate needs."""
    cycles_per_pair_wave = 4.0 * (valu - trans) + 16.0 * trans
    return pair_rate * cycles_per_pair_wave / (64.0 * SIMD_CYCLES_PER_S)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, prisms, density, gpu_gz):
    """The CPU oracle (C + OpenMP restatement of the reference's parallel
    Numba loop) on a bounded sample: every 8th observation row and column
    (32 x 32 = 1024 points) against all 1,000,000 prisms.  The sample's
    values are also compared with the GPU result of the timed run
    (max |gpu - cpu| / max |cpu|)."""
    import oracle

    E, N, U = (c[::8, ::8] for c in w["coordinates"])
    cores = len(os.sched_getaffinity(0))
    threads = min(cores, 16)
    pairs = E.size * prisms.shape[0]
    t0 = time.perf_counter()
    ref = oracle.prism_gravity((E, N, U), prisms, density, "g_z", n_threads=threads)
    dt = time.perf_counter() - t0
    got = gpu_gz.reshape(w["coordinates"][0].shape)[::8, ::8]
    return {"value": pairs / dt, "unit": "pair-evals/s", "cores": threads, "kind": "port",
            "sample": f"{E.size} of {w['coordinates'][0].size} observation points (every 8th row/column) x "
                      f"{prisms.shape[0]} prisms = {pairs:.3e} pairs, g_z FP64, {dt:.1f} s",
            "host_cores_visible": cores,
            "max_rel_diff_vs_gpu": float(np.abs(got - ref).max() / np.abs(ref).max())}


def ten_field_parity(w, prisms, density, gpu_all):
    """All ten fields of the timed fused launch against the oracle at 81 of
    the 256 x 256 points (rows and columns 0, 17, 51, 102, 130, 153, 204,
    240, 255), max |gpu - cpu| / max |cpu| per field, over all 81 points
    and over the 56 that do not lie on a vertical line through prism edges
    (x and y both on edge planes: rows/columns 0, 51, 102, 153, 204, 255 of
    this grid, where the tensor terms take their measure-zero conventions)."""
    import oracle
    from harmonica_amd._lib import FIELDS

    idx = np.array([0, 17, 51, 102, 130, 153, 204, 240, 255])
    coords = tuple(c[np.ix_(idx, idx)] for c in w["coordinates"])
    edges_e = np.concatenate([w["easting"] - 50.0, w["easting"] + 50.0])
    edges_n = np.concatenate([w["northing"] - 50.0, w["northing"] + 50.0])
    on_line = np.isin(coords[0], edges_e) & np.isin(coords[1], edges_n)
    shape = w["coordinates"][0].shape
    t0 = time.perf_counter()
    out = {"points": int(on_line.size), "on_edge_lines": int(on_line.sum())}
    worst_all = worst_off = 0.0
    for k, field in enumerate(FIELDS):
        ref = oracle.prism_gravity(coords, prisms, density, field)
        got = gpu_all[k].reshape(shape)[np.ix_(idx, idx)]
        scale = np.abs(ref).max()
        e_all = float(np.abs(got - ref).max() / scale)
        e_off = float(np.abs(got - ref)[~on_line].max() / scale)
        out[field] = {"all_points": e_all, "off_edge_lines": e_off}
        worst_all, worst_off = max(worst_all, e_all), max(worst_off, e_off)
    out["max_all_points"], out["max_off_edge_lines"] = worst_all, worst_off
    out["cpu_s"] = time.perf_counter() - t0
    log(f"ten fields vs oracle at 81 points: {worst_all:.2e} (all), {worst_off:.2e} "
        f"(off edge lines)")
    return out


def _sync_time(fn):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = fn()
    torch.cuda.synchronize()
    return res, time.perf_counter() - t0


def bench_layer_fused(args):
    """Ten fields of the 1000 x 1000 terrain layer in one fused launch
    through the prism-layer accessor path (hmx_prism_layer_gravity)."""
    import torch

    from harmonica_amd._forward import layer_fields_device, to_device
    from harmonica_amd._lib import FIELDS
    from harmonica_amd.workloads import terrain_layer_large

    w = terrain_layer_large(n_prisms_side=args.prism_side, n_obs_side=args.obs_side)
    E, N, U = (to_device(c.ravel()) for c in w["coordinates"])
    top = np.maximum(w["surface"], w["reference"])
    bottom = np.minimum(w["surface"], w["reference"]) + 0 * top
    args_dev = (to_device(w["easting"]), to_device(w["northing"]), to_device(bottom),
                to_device(top), to_device(w["density"]))
    pairs = E.numel() * w["density"].size
    out = {}
    for name, fields in (("g_z", ("g_z",)), ("all_ten", FIELDS)):
        layer_fields_device(E, N, U, *args_dev, fields)
        _, dt = _sync_time(lambda: layer_fields_device(E, N, U, *args_dev, fields))
        out[name] = {"s": dt, "pair_evals_per_s": pairs / dt}
        log(f"layer {name}: {dt:.3f} s, {pairs / dt:.4e} pair-evals/s")
    # The drop-in Python API from host numpy arrays (PCIe copies both ways,
    # validation and null-prism discard included).
    import harmonica_amd as hm
    from harmonica_amd.workloads import layer_to_prisms

    prisms = layer_to_prisms(w["easting"], w["northing"], w["surface"], w["reference"])
    dens = w["density"].ravel()
    t0 = time.perf_counter()
    hm.prism_gravity(w["coordinates"], prisms, dens, field="g_z")
    dt = time.perf_counter() - t0
    out["api_prism_gravity_host_arrays"] = {"s": dt, "pair_evals_per_s": pairs / dt}
    log(f"prism_gravity API from host arrays: {dt:.3f} s, {pairs / dt:.4e} pair-evals/s")
    return out


def bench_point(args):
    """point_gravity on 1,000,000 point masses (the terrain prisms' centres
    of mass and masses) at the 256 x 256 grid at 2500 m: g_z and the ten
    fields fused (hmx_point_gravity, inputs in HBM)."""
    from harmonica_amd._forward import point_fields_device, to_device
    from harmonica_amd._lib import FIELDS
    from harmonica_amd import workloads
    from harmonica_amd.workloads import layer_to_prisms, terrain_layer_large

    w = terrain_layer_large(n_prisms_side=args.prism_side, n_obs_side=args.obs_side)
    prisms = layer_to_prisms(w["easting"], w["northing"], w["surface"], w["reference"])
    centres = [0.5 * (prisms[:, 2 * a] + prisms[:, 2 * a + 1]) for a in range(3)]
    vol = np.prod([prisms[:, 2 * a + 1] - prisms[:, 2 * a] for a in range(3)], axis=0)
    masses = w["density"].ravel() * vol
    E, N, U = (to_device(c.ravel()) for c in w["coordinates"])
    P = [to_device(c) for c in centres]
    M = to_device(masses)
    pairs = E.numel() * M.numel()
    out = {}
    for name, fields in (("g_z", ("g_z",)), ("all_ten", FIELDS)):
        point_fields_device(E, N, U, *P, M, fields)
        _, dt = _sync_time(lambda: point_fields_device(E, N, U, *P, M, fields))
        out[name] = {"s": dt, "pair_evals_per_s": pairs / dt}
        log(f"point {name}: {dt:.4f} s, {pairs / dt:.4e} pair-evals/s")
    return out


def bench_magnetic(args):
    """prism_magnetic field "b" (three components in one launch) of the 1M
    terrain prisms with induced + remanent magnetisation at the 256 x 256
    grid at 2500 m (hmx_prism_magnetic, inputs in HBM)."""
    from harmonica_amd._forward import to_device
    from harmonica_amd._magnetic import magnetic_fields_device
    from harmonica_amd.workloads import layer_to_prisms, terrain_layer_large, terrain_magnetization

    w = terrain_layer_large(n_prisms_side=args.prism_side, n_obs_side=args.obs_side)
    prisms = layer_to_prisms(w["easting"], w["northing"], w["surface"], w["reference"])
    mag = terrain_magnetization(n_prisms_side=args.prism_side)
    E, N, U = (to_device(c.ravel()) for c in w["coordinates"])
    P = to_device(prisms)
    M = [to_device(m) for m in mag]
    pairs = E.numel() * P.shape[0]
    out = {}
    counts = {"b": (FLOPS_PER_PAIR_B, VALU_INSTR_PER_PAIR_B, TRANS_PER_PAIR_B),
              "b_u": (FLOPS_PER_PAIR_BU, VALU_INSTR_PER_PAIR_BU, TRANS_PER_PAIR_BU)}
    full_size = args.prism_side == 1000 and args.obs_side == 256
    for name, comps in (("b", ("b_e", "b_n", "b_u")), ("b_u", ("b_u",))):
        magnetic_fields_device(E, N, U, P, *M, comps)
        _, dt = _sync_time(lambda: magnetic_fields_device(E, N, U, P, *M, comps))
        rate = pairs / dt
        out[name] = {"s": dt, "pair_evals_per_s": rate}
        if full_size:  # the counters were taken on this workload
            flops, valu, trans = counts[name]
            out[name].update({"fp64_tflops": rate * flops / 1e12,
                              "fp64_peak_frac": rate * flops / FP64_PEAK,
                              "valu_issue_frac": valu_issue_frac(rate, valu, trans)})
        log(f"magnetic {name}: {dt:.4f} s, {rate:.4e} pair-evals/s")
    return out


def bench_eqs_gb(args):
    """Gradient-boosted equivalent sources (EquivalentSourcesGB) on the
    scattered g_z data of the EQS workload with --gb-data points (default
    200,000), default windows (about 5000 data each, 50% overlap), damping 10
    and depth 9 km (the values of the reference's gradient-boosting example,
    doc/user_guide/equivalent_sources/gradient-boosted-eqs.rst:92-94), FP64
    window fits; predictions on the 512 x 512 grid.  With --gb-oracle the
    oracle's CPU loop runs the first 3 windows for a seconds-per-window
    comparison."""
    import harmonica_amd as hm
    from harmonica_amd.workloads import eqs_problem

    damping, depth = 10.0, 9e3
    w = eqs_problem(n_data=args.gb_data)
    coords = w["coordinates"]
    data = hm.prism_gravity(coords, w["prisms"], w["density"], field="g_z") + w["noise"]
    small = tuple(c[:20000] for c in coords)
    hm.EquivalentSourcesGB(damping=damping, depth=depth, random_state=0).fit(
        small, data[:20000])  # untimed warm-up of the code path
    eqs = hm.EquivalentSourcesGB(damping=damping, depth=depth, random_state=0)
    _, dt = _sync_time(lambda: eqs.fit(coords, data))
    pred, dtp = _sync_time(lambda: eqs.predict(w["grid"]))
    rms = eqs.rmse_per_iteration_
    out = {"n_data": coords[0].size, "windows": eqs.fit_info_["windows"],
           "window_size_m": float(eqs.window_size_), "fit_s": dt,
           "s_per_window": dt / eqs.fit_info_["windows"], "predict_s": dtp,
           "rmse_first_last": [float(rms[0]), float(rms[-1])], "info": eqs.fit_info_}
    log(f"eqs gb: {out['n_data']} data, {out['windows']} windows, fit {dt:.2f} s "
        f"({1e3 * out['s_per_window']:.1f} ms/window), rms {rms[0]:.3g} -> {rms[-1]:.3g}")
    if args.gb_oracle:
        import oracle

        src_w, dat_w = eqs._create_windows(coords)
        t0 = time.perf_counter()
        _, rms_cpu = oracle.eqs_gb_fit(coords, data, None, eqs.points_, damping, src_w[:3],
                                       dat_w[:3])
        cpu_s = (time.perf_counter() - t0) / 3
        out["oracle_s_per_window"] = cpu_s
        out["oracle_rmse_after_3"] = float(rms_cpu[-1])
        out["gpu_rmse_after_3"] = float(rms[3])
        log(f"eqs gb oracle: {cpu_s:.2f} s/window on {oracle.num_threads()} threads, rms after "
            f"3 windows {rms_cpu[-1]:.6g} (GPU {rms[3]:.6g})")
    return out


def bench_tesseroid(args):
    """tesseroid_gravity g_z of a global 1 x 1 degree tesseroid layer
    (64,800 tesseroids, tops from a fractal surface in [-4, 4] km about the
    mean Earth radius, bottoms 30 km below the mean radius) on a 1 x 1
    degree grid of 65,160 points 10 km above the mean radius (2D adaptive
    discretisation, distance-size ratio 2.5)."""
    import harmonica_amd as hm
    from harmonica_amd.workloads import global_tesseroid_layer

    w = global_tesseroid_layer()
    tess, dens, coords = w["tesseroids"], w["density"], w["coordinates"]
    pairs = coords[0].size * tess.shape[0]
    hm.tesseroid_gravity(tuple(c[:2, :] for c in coords), tess, dens, field="g_z",
                         disable_checks=True)  # untimed warm-up
    _, dt = _sync_time(lambda: hm.tesseroid_gravity(coords, tess, dens, field="g_z",
                                                    disable_checks=True))
    out = {"n_points": coords[0].size, "n_tesseroids": tess.shape[0], "s": dt,
           "pair_evals_per_s": pairs / dt}
    log(f"tesseroid g_z: {coords[0].size} points x {tess.shape[0]} tesseroids, {dt:.3f} s, "
        f"{pairs / dt:.4e} pair-evals/s")
    return out


def bench_batched(args):
    """512 independent uneven models (log-uniform 1e3-2e5 prisms, 1e3-5e4
    points), g_z, one persistent graph-captured launch, autotuned."""
    from harmonica_amd._batched import BatchedPrismGz
    from harmonica_amd.workloads import batched_models

    models = batched_models(n_models=args.batched_models)
    batch = BatchedPrismGz(models)
    timings = batch.autotune(prisms_per_item=(1024, 4096), blocks_per_cu=(3, 6), use_cache=False)
    batch.capture()
    batch.replay()
    _, dt = _sync_time(batch.replay)
    log(f"batched: {batch.n_models} models, {batch.pairs:.3e} pairs, {dt:.3f} s")
    # spot check against single-model launches
    import harmonica_amd as hm

    res = batch.split()
    worst = 0.0
    for b in range(0, batch.n_models, max(batch.n_models // 8, 1)):
        c, p, d = models[b]
        ref = hm.prism_gravity(c, p, d, field="g_z")
        worst = max(worst, float(np.abs(res[b] - ref).max() / np.abs(ref).max()))
    return {"models": batch.n_models, "pairs": batch.pairs, "s": dt, "pair_evals_per_s": batch.pairs / dt,
            "prisms_per_item": batch.prisms_per_item, "blocks_per_cu": batch.blocks_per_cu,
            "autotune_ms": {str(k): v for k, v in timings.items()},
            "max_rel_diff_vs_single_model": worst}


def bench_eqs(args, modes=None):
    """Equivalent-source fit of 60,000 scattered g_z data (3 prisms + 1 mGal
    noise), damping 1e-3, depth 4500 m: FP64 fit vs mixed-precision fit,
    predictions on a 512 x 512 grid at 1000 m."""
    import harmonica_amd as hm
    from harmonica_amd.workloads import eqs_problem

    w = eqs_problem(n_data=args.eqs_data)
    coords = w["coordinates"]
    data = hm.prism_gravity(coords, w["prisms"], w["density"], field="g_z") + w["noise"]
    out = {"n_data": coords[0].size}
    preds = {}
    small = tuple(c[:2000] for c in coords)
    if modes is None:
        modes = ("float64", "mixed", "mixed-bf16")
        if args.eqs_all:
            modes += ("mixed-fp32",)
    for mode in modes:  # untimed warm-up of every code path
        hm.EquivalentSources(damping=w["damping"], depth=w["depth"], fit_precision=mode).fit(
            small, data[:2000])
    for mode in modes:
        eqs = hm.EquivalentSources(damping=w["damping"], depth=w["depth"], fit_precision=mode)
        _, dt = _sync_time(lambda: eqs.fit(coords, data))
        pred, dtp = _sync_time(lambda: eqs.predict(w["grid"]))
        preds[mode] = pred
        out[mode] = {"fit_s": dt, "predict_s": dtp, "info": eqs.fit_info_}
        log(f"eqs {mode}: fit {dt:.2f} s, predict {dtp:.3f} s, {eqs.fit_info_}")
    ref = preds.get("float64")
    from harmonica_amd._eqs import MIXED_PREDICTION_BAR, MIXED_RTOL

    for mode in modes[1:] if ref is not None else ():
        out[mode]["max_rel_pred_diff_vs_fp64"] = float(np.abs(preds[mode] - ref).max() /
                                                       np.abs(ref).max())
        out[mode]["prediction_bar"] = MIXED_PREDICTION_BAR
        out[mode]["cg_rtol"] = MIXED_RTOL
        out[mode]["speedup_vs_fp64"] = out["float64"]["fit_s"] / out[mode]["fit_s"]
        # Both fits factorise the same size of matrix in FP64: with the rest
        # of the mixed fit free, the speedup would be the FP64 fit's time over
        # its factorisation's.
        fact = out["float64"]["info"].get("factorisation_s", 0.0)
        out[mode]["speedup_ceiling"] = out["float64"]["fit_s"] / fact if fact > 0 else None
    if args.eqs_oracle:
        import oracle

        n = min(20000, coords[0].size)
        sub = tuple(c[:n] for c in coords)
        eqs = hm.EquivalentSources(damping=w["damping"], depth=w["depth"]).fit(sub, data[:n])
        J = oracle.eqs_jacobian(sub, eqs.points_)
        t0 = time.perf_counter()
        coefs = oracle.eqs_fit_coefs(J, data[:n], None, w["damping"])
        cpu_s = time.perf_counter() - t0
        del J
        grid = tuple(c[::4, ::4] for c in w["grid"])
        p_gpu = eqs.predict(grid)
        p_cpu = oracle.eqs_predict(grid, eqs.points_, coefs)
        out["oracle_subset"] = {"n": n, "cpu_fit_s": cpu_s,
                                "max_rel_pred_diff": float(np.abs(p_gpu - p_cpu).max() / np.abs(p_cpu).max())}
        log(f"eqs oracle subset: {out['oracle_subset']}")
    return out


def bench_inversion(args):
    """Differentiable g_z of a 200 x 200 basement layer on a 128 x 128 grid:
    forward, backward (densities and tops), finite-difference check, 20
    Gauss-Newton iterations."""
    import torch

    import harmonica_amd as hm
    from harmonica_amd import _autograd
    from harmonica_amd.workloads import inversion_layer

    w = inversion_layer()
    geom = _autograd.LayerGeometry(w["easting"], w["northing"], w["bottom"], w["coordinates"])
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    rho, top = dev(w["density"]), dev(w["top"])
    data = geom.forward(rho, top)
    pairs = geom.n_obs * geom.n_prisms
    _, t_fwd = _sync_time(lambda: geom.forward(rho, top))
    v = torch.ones_like(data)
    g_rho, g_top = geom.vjp(rho, top, v)  # first call also warms the backward kernels
    _, t_bwd = _sync_time(lambda: geom.vjp(rho, top, v))
    # central differences on three prisms, FP64, steps 1e-4 kg/m^3 and 1e-3 m
    pts = [c.ravel() for c in w["coordinates"]]
    he = (w["easting"][1] - w["easting"][0]) / 2
    errs = []
    for k, j in ((10, 20), (100, 100), (190, 5)):
        p = np.array([w["easting"][j] - he, w["easting"][j] + he, w["northing"][k] - he,
                      w["northing"][k] + he, w["bottom"][k, j], w["top"][k, j]])
        d = w["density"][k, j]
        fr = (hm.prism_gravity(pts, p, d + 1e-4, "g_z").sum()
              - hm.prism_gravity(pts, p, d - 1e-4, "g_z").sum()) / 2e-4
        pu, pd = p.copy(), p.copy()
        pu[5] += 1e-3
        pd[5] -= 1e-3
        ft = (hm.prism_gravity(pts, pu, d, "g_z").sum() - hm.prism_gravity(pts, pd, d, "g_z").sum()) / 2e-3
        errs.append(abs(g_rho[k, j].item() - fr) / abs(fr))
        errs.append(abs(g_top[k, j].item() - ft) / abs(ft))
    rho0 = torch.full_like(rho, -300.0)
    top0 = torch.full_like(top, -3500.0)
    (r, t, hist), t_gn = _sync_time(lambda: _autograd.gauss_newton(geom, data, rho0, top0, n_iterations=20))
    log(f"inversion: fwd {t_fwd * 1e3:.1f} ms, bwd {t_bwd * 1e3:.1f} ms, GN {t_gn:.2f} s, misfit {hist[0]:.3g} -> {hist[-1]:.3g}")
    return {"pairs": pairs, "forward_ms": t_fwd * 1e3, "backward_ms": t_bwd * 1e3,
            "fd_max_rel_err": max(errs), "gauss_newton_s": t_gn, "misfit_mgal": [hist[0], hist[-1]],
            "iterations": len(hist) - 1}


def bench_cpu_configs(args):
    """The oracle's CPU loop beside the GPU on the other workloads: the
    100 x 100 terrain layer with 64 x 64 points (all 4.1e7 pairs, layer
    path, g_z), a 16-model subset of the batched workload, and one forward
    model of the inversion layer; each with the GPU time of the same call
    and the largest difference relative to max|g_z|."""
    import harmonica_amd as hm
    import oracle
    from harmonica_amd.workloads import batched_models, inversion_layer, terrain_layer_small

    threads = min(len(os.sched_getaffinity(0)), 16)
    out = {"threads": threads}
    w = terrain_layer_small()
    layer = hm.prism_layer((w["easting"], w["northing"]), w["surface"], w["reference"],
                           properties={"density": w["density"]})
    bottom, top = layer.bottom.values, layer.top.values
    pairs = w["coordinates"][0].size * w["density"].size
    layer.prism_layer.gravity(w["coordinates"], field="g_z")  # warm-up
    gpu, t_gpu = _sync_time(lambda: layer.prism_layer.gravity(w["coordinates"], field="g_z"))
    t0 = time.perf_counter()
    cpu = oracle.prism_layer_gravity(w["coordinates"], w["easting"], w["northing"], bottom, top,
                                     w["density"], "g_z", n_threads=threads)
    t_cpu = time.perf_counter() - t0
    out["terrain_small_layer"] = {
        "pairs": pairs, "gpu_s": t_gpu, "cpu_s": t_cpu, "cpu_pair_evals_per_s": pairs / t_cpu,
        "gpu_pair_evals_per_s": pairs / t_gpu,
        "max_rel_diff": float(np.abs(gpu - cpu).max() / np.abs(cpu).max())}
    # 16 models spread over the batch (every 32nd of 512); the oracle runs
    # on every 8th observation point of each, the GPU on all of them
    models = batched_models(n_models=args.batched_models)
    models = models[:: max(len(models) // 16, 1)][:16]
    pairs = pairs_cpu = 0
    t_gpu = t_cpu = 0.0
    diff = 0.0
    for coords, prisms, density in models:
        hm.prism_gravity(coords, prisms, density, field="g_z")
        g, dt = _sync_time(lambda: hm.prism_gravity(coords, prisms, density, field="g_z"))
        t_gpu += dt
        pairs += coords[0].size * prisms.shape[0]
        sub = tuple(c[::8] for c in coords)
        pairs_cpu += sub[0].size * prisms.shape[0]
        t0 = time.perf_counter()
        c = oracle.prism_gravity(sub, prisms, density, "g_z", n_threads=threads)
        t_cpu += time.perf_counter() - t0
        diff = max(diff, float(np.abs(g[::8] - c).max() / np.abs(c).max()))
    out["batched_16_models"] = {
        "pairs": pairs, "gpu_s": t_gpu, "gpu_pair_evals_per_s": pairs / t_gpu,
        "cpu_pairs": pairs_cpu, "cpu_s": t_cpu, "cpu_pair_evals_per_s": pairs_cpu / t_cpu,
        "max_rel_diff": diff, "note": "GPU: one launch per model, not the batched graph"}
    w = inversion_layer()
    layer = hm.prism_layer((w["easting"], w["northing"]), w["top"], w["bottom"],
                           properties={"density": w["density"]})
    pairs = w["coordinates"][0].size * w["density"].size
    layer.prism_layer.gravity(w["coordinates"], field="g_z")
    gpu, t_gpu = _sync_time(lambda: layer.prism_layer.gravity(w["coordinates"], field="g_z"))
    t0 = time.perf_counter()
    cpu = oracle.prism_layer_gravity(w["coordinates"], w["easting"], w["northing"],
                                     layer.bottom.values, layer.top.values, w["density"], "g_z",
                                     n_threads=threads)
    t_cpu = time.perf_counter() - t0
    out["inversion_forward"] = {
        "pairs": pairs, "gpu_s": t_gpu, "cpu_s": t_cpu, "cpu_pair_evals_per_s": pairs / t_cpu,
        "gpu_pair_evals_per_s": pairs / t_gpu,
        "max_rel_diff": float(np.abs(gpu - cpu).max() / np.abs(cpu).max())}
    for k, v in out.items():
        if isinstance(v, dict):
            log(f"cpu configs {k}: GPU {v['gpu_s']:.4f} s, "
                f"CPU {v['cpu_s']:.2f} s on {threads} threads, max rel diff {v['max_rel_diff']:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ten", action="store_true",
                    help="skip the fused ten-field launch (timed by default)")
    ap.add_argument("--extra", action="store_true", help="accepted; the ten-field leg is default")
    ap.add_argument("--obs-side", type=int, default=256)
    ap.add_argument("--prism-side", type=int, default=1000)
    ap.add_argument("--layer", action="store_true", help="also time the layer accessor path")
    ap.add_argument("--point", action="store_true", help="also time point_gravity (1M masses)")
    ap.add_argument("--magnetic", action="store_true",
                    help="also time prism_magnetic (field b) on the 1M terrain prisms")
    ap.add_argument("--tesseroid", action="store_true",
                    help="also time tesseroid_gravity on a global 1-degree layer")
    ap.add_argument("--batched", action="store_true", help="also run the 512-model batch")
    ap.add_argument("--batched-models", type=int, default=512)
    ap.add_argument("--no-eqs", action="store_true",
                    help="skip the equivalent-source fits (FP64 vs mixed, run by default)")
    ap.add_argument("--eqs-all", action="store_true", help="also time the FP32-MFMA mixed fit")
    ap.add_argument("--eqs-data", type=int, default=60000)
    ap.add_argument("--eqs-gb", action="store_true",
                    help="also run the gradient-boosted equivalent-source fit")
    ap.add_argument("--gb-data", type=int, default=200000)
    ap.add_argument("--gb-oracle", action="store_true",
                    help="time the oracle's CPU loop on the first 3 windows")
    ap.add_argument("--eqs-oracle", action="store_true", help="check the FP64 fit on a 20k subset")
    ap.add_argument("--inversion", action="store_true", help="also run the autograd inversion")
    ap.add_argument("--cpu-configs", action="store_true",
                    help="also time the oracle beside the GPU on the other workloads")
    ap.add_argument("--eqs-only", choices=("float64", "mixed", "mixed-bf16", "mixed-fp32"),
                    help="only run the equivalent-source fit in this precision (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": 1 << 30}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the "
              "multi-GPU sharding of this path (observation tiles per GPU) is not built",
              file=sys.stderr)
        sys.exit(2)

    if args.eqs_only:
        print(json.dumps({"eqs_fit": bench_eqs(args, (args.eqs_only,))}))
        return

    import torch

    from harmonica_amd import _lib
    from harmonica_amd._forward import ctypes_size, to_device
    from harmonica_amd import workloads
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

    input_sha256 = workloads.sha256(w)
    log(f"g_z: {n_obs} points x {n_prisms} prisms, warmup {args.warmup}, steps {args.steps}, "
        f"input sha256 {input_sha256[:16]}")
    step, out = make_step(_lib.FIELD_BIT["g_z"], 1)
    ms, wall = timed(step, args.steps, args.warmup)
    value = pairs / (ms * 1e-3)
    log(f"g_z: {ms:.2f} ms/step, {value:.4e} pair-evals/s (wall {wall:.2f} s)")

    achieved = value * FLOPS_PER_PAIR_GZ
    full_size = args.prism_side == 1000 and args.obs_side == 256
    roofline = {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK,
                "traffic": TRAFFIC_BYTES_GZ if full_size else None,
                "flops_per_pair": FLOPS_PER_PAIR_GZ,
                "valu_issue_frac": valu_issue_frac(value, VALU_INSTR_PER_PAIR_GZ, TRANS_PER_PAIR_GZ),
                "per_vertex_equivalent_tflops": value * PER_VERTEX_FLOPS_PER_PAIR_GZ / 1e12}

    extra = {}
    if not args.no_ten:
        # The full-tensor workload: all ten fields in one fused launch.
        mask = (1 << 10) - 1
        step10, out10 = make_step(mask, 10)
        steps10 = max(1, min(args.steps, 3))
        ms10, _ = timed(step10, steps10, 1)
        rate10 = pairs / (ms10 * 1e-3)
        ach10 = rate10 * FLOPS_PER_PAIR_ALL
        extra["all_ten_fields"] = {
            "ms_per_step": ms10, "steps": steps10, "pair_evals_per_s": rate10,
            "roofline": {"bound": "fp64-valu", "achieved": ach10 / 1e12, "peak": FP64_PEAK / 1e12,
                         "unit": "TFLOP/s", "frac": ach10 / FP64_PEAK, "traffic": None,
                         "flops_per_pair": FLOPS_PER_PAIR_ALL,
                         "valu_issue_frac": valu_issue_frac(rate10, VALU_INSTR_PER_PAIR_ALL,
                                                            TRANS_PER_PAIR_ALL)}}
        log(f"ten fields fused: {ms10:.2f} ms/step, {ach10 / FP64_PEAK:.3f} of FP64 peak")
        if not args.no_cpu and full_size:
            extra["all_ten_fields"]["parity"] = ten_field_parity(w, prisms, density,
                                                                 out10.cpu().numpy())

    if args.layer:
        extra["layer_path"] = bench_layer_fused(args)
    if args.point:
        extra["point_masses"] = bench_point(args)
    if args.magnetic:
        extra["magnetic"] = bench_magnetic(args)
    if args.tesseroid:
        extra["tesseroid"] = bench_tesseroid(args)
    if args.batched:
        extra["batched"] = bench_batched(args)
    if not args.no_eqs:
        extra["eqs_fit"] = bench_eqs(args)
    if args.eqs_gb:
        extra["eqs_gb"] = bench_eqs_gb(args)
    if args.inversion:
        extra["inversion"] = bench_inversion(args)
    if args.cpu_configs:
        extra["cpu_configs"] = bench_cpu_configs(args)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(w, prisms, density, out[0].cpu().numpy())
        log(f"cpu: {cpu['value']:.4e} pair-evals/s on {cpu['cores']} threads, "
            f"max rel diff vs GPU {cpu['max_rel_diff_vs_gpu']:.2e}")

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
                   "n_prisms": n_prisms, "n_obs": n_obs, "pairs_per_step": pairs,
                   "input_sha256": input_sha256},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if extra:
        line["extra"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
