"""
Many independent g_z prism models in one persistent, graph-captured launch
(no counterpart in the reference, where each model is one prism_gravity
call, src/harmonica/_forward/prisms/gravity.py:51-236).

``BatchedPrismGz(models)`` concatenates the models on the GPU, plans the
work list in libhmx (hmx_batched_plan: items of one 256-point tile x one
prism chunk, longest first), optionally autotunes the chunk size and the
persistent grid size (timed with HIP events on the real workload, cached as
JSON), and runs hmx_batched_prism_gz; ``capture()`` records the launch
sequence in a HIP graph (torch.cuda.CUDAGraph) for replay.

Each model's points are uploaded in boustrophedon strip order
(``_eqs.strip_order``): a wavefront's 64 points then cover a compact patch,
so that for most prisms no point of the wave lies between the prism's
y-planes (or x-planes) and the wave takes the lean form of the g_z kernel
instead of the general one.  On the 512-model workload (random points)
unsorted waves needed the general form for 86% of the (wave, prism) pairs,
sorted ones for 8%.  ``split()`` returns each model's values in the
caller's order.
"""

import ctypes
import json
import os
import time

import numpy as np
import torch

from . import _lib
from ._eqs import strip_order
from ._forward import check_prisms, discard_null_prisms, to_device
from ._lib import check, lib, ptr, stream_handle

PRISMS_PER_ITEM = (512, 1024, 2048, 4096, 8192)
BLOCKS_PER_CU = (2, 4, 8)


def _cache_path():
    return os.environ.get(
        "HMX_AUTOTUNE_CACHE",
        os.path.join(os.path.expanduser("~"), ".cache", "harmonica_amd", "autotune.json"),
    )


def _i64p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class BatchedPrismGz:
    """g_z (mGal) of a list of models ``[(coordinates, prisms, density), ...]``.

    ``strip_ratio``: northing-to-easting aspect of the 64-point runs of the
    strip order (None uploads the points in the caller's order).  ``out``
    holds the values in the upload order; ``split()`` in the caller's."""

    def __init__(self, models, prisms_per_item=2048, blocks_per_cu=4, disable_checks=False,
                 strip_ratio=0.25):
        _lib.require_device()
        obs, prs, dens, obs_off, pr_off = [], [], [], [0], [0]
        self.shapes = []
        self.orders = []
        self.strip_ratio = strip_ratio
        for coordinates, prisms, density in models:
            cast = np.broadcast(*coordinates[:3])
            self.shapes.append(cast.shape)
            c = [np.broadcast_to(np.asarray(x, dtype=float), cast.shape).ravel() for x in coordinates[:3]]
            order = None if strip_ratio is None else strip_order(c[0], c[1], strip_ratio)
            self.orders.append(order)
            if order is not None:
                c = [x[order] for x in c]
            p = np.atleast_2d(np.asarray(prisms, dtype=float)).reshape(-1, 6)
            d = np.atleast_1d(np.asarray(density, dtype=float)).ravel()
            if not disable_checks:
                if d.size != p.shape[0]:
                    raise ValueError(
                        f"Number of elements in density ({d.size}) "
                        + f"mismatch the number of prisms ({p.shape[0]})"
                    )
                check_prisms(p)
            p, d = discard_null_prisms(p, d)
            obs.append(np.column_stack(c) if c[0].size else np.zeros((0, 3)))
            prs.append(p)
            dens.append(d)
            obs_off.append(obs_off[-1] + c[0].size)
            pr_off.append(pr_off[-1] + p.shape[0])
        allobs = np.concatenate(obs) if obs else np.zeros((0, 3))
        self.n_models = len(models)
        self.n_obs = int(obs_off[-1])
        self.obs_offsets = np.asarray(obs_off, dtype=np.int64)
        self.prism_offsets = np.asarray(pr_off, dtype=np.int64)
        self.pairs = int(sum((obs_off[b + 1] - obs_off[b]) * (pr_off[b + 1] - pr_off[b])
                             for b in range(self.n_models)))
        self.e, self.n, self.u = (to_device(np.ascontiguousarray(allobs[:, k])) for k in range(3))
        self.prisms = to_device(np.concatenate(prs) if prs else np.zeros((0, 6)))
        self.density = to_device(np.concatenate(dens) if dens else np.zeros(0))
        self.d_obs_off = torch.as_tensor(self.obs_offsets).cuda()
        self.out = torch.empty(max(self.n_obs, 1), dtype=torch.float64, device=self.e.device)
        self.graph = None
        self.plan(prisms_per_item, blocks_per_cu)

    def plan(self, prisms_per_item, blocks_per_cu):
        """Build the work list for a chunk size and persistent grid size."""
        L = lib()
        n_items = ctypes.c_int64(0)
        check(L.hmx_batched_plan_size(_i64p(self.obs_offsets), _i64p(self.prism_offsets),
                                      self.n_models, int(prisms_per_item), ctypes.byref(n_items)))
        items = np.zeros(max(n_items.value, 1) * 5, dtype=np.int64)
        base = np.zeros(self.n_models, dtype=np.int64)
        chunks = np.zeros(self.n_models, dtype=np.int64)
        got = ctypes.c_int64(0)
        check(L.hmx_batched_plan(_i64p(self.obs_offsets), _i64p(self.prism_offsets), self.n_models,
                                 int(prisms_per_item), _i64p(items), n_items.value,
                                 ctypes.byref(got), _i64p(base), _i64p(chunks)))
        self.n_items = got.value
        self.items = torch.as_tensor(items).cuda()
        self.item_base = torch.as_tensor(base).cuda()
        self.chunks = torch.as_tensor(chunks).cuda()
        nbytes = ctypes.c_size_t(0)
        check(L.hmx_batched_workspace_bytes(self.n_items, ctypes.byref(nbytes)))
        self.ws = torch.empty(nbytes.value, dtype=torch.uint8, device=self.e.device)
        self.ws_bytes = nbytes.value
        self.prisms_per_item = int(prisms_per_item)
        self.blocks_per_cu = int(blocks_per_cu)
        self.graph = None

    def run(self, stream=None):
        """Launch on the current (or given) stream; returns the output tensor."""
        check(lib().hmx_batched_prism_gz(
            ptr(self.e), ptr(self.n), ptr(self.u), ptr(self.d_obs_off), ptr(self.prisms),
            ptr(self.density), ptr(self.items), self.n_items, ptr(self.item_base), ptr(self.chunks),
            self.n_models, self.n_obs, self.blocks_per_cu, ptr(self.out), ptr(self.ws),
            self.ws_bytes, stream_handle(stream)))
        return self.out

    def capture(self):
        """Record the launch sequence in a HIP graph; replay() re-runs it."""
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.run()
        torch.cuda.synchronize()
        return self.graph

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.out

    def split(self, out=None):
        """Per-model numpy results in the caller's point order, reshaped to
        each model's coordinates."""
        vals = (self.out if out is None else out).cpu().numpy()
        res = []
        for b in range(self.n_models):
            v = vals[self.obs_offsets[b]:self.obs_offsets[b + 1]]
            if self.orders[b] is not None:
                unsorted = np.empty_like(v)
                unsorted[self.orders[b]] = v
                v = unsorted
            res.append(v.reshape(self.shapes[b]))
        return res

    def time(self, reps=2):
        """Mean milliseconds per run (HIP events on the current stream)."""
        self.run()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            self.run()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps

    def autotune(self, prisms_per_item=PRISMS_PER_ITEM, blocks_per_cu=BLOCKS_PER_CU, reps=1,
                 use_cache=True):
        """Time every (chunk size, grid size) pair on this workload, keep the
        fastest, and cache it by workload size (number of models and log2 of
        the pair count).  Returns {config: ms}."""
        key = f"batched_gz:{self.n_models}:{int(np.log2(max(self.pairs, 1)))}:{self.strip_ratio}"
        path = _cache_path()
        cache = {}
        if use_cache and os.path.exists(path):
            try:
                with open(path) as f:
                    cache = json.load(f)
            except (OSError, ValueError):
                cache = {}
        if use_cache and key in cache:
            best = cache[key]
            self.plan(best["prisms_per_item"], best["blocks_per_cu"])
            return {"cached": best}
        timings = {}
        for ppi in prisms_per_item:
            for bpc in blocks_per_cu:
                self.plan(ppi, bpc)
                timings[(ppi, bpc)] = self.time(reps)
        (ppi, bpc), ms = min(timings.items(), key=lambda kv: kv[1])
        self.plan(ppi, bpc)
        if use_cache:
            cache[key] = {"prisms_per_item": ppi, "blocks_per_cu": bpc, "ms": ms,
                          "when": time.strftime("%Y-%m-%d %H:%M:%S")}
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        return {f"{k[0]}x{k[1]}": v for k, v in timings.items()}
