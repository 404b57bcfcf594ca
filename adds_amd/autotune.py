"""Persisted autotuning of the solver's parameters (delta, bucket ring, open buckets, workgroups per CU,
claim size) per workload.

The search itself runs in the native library (adds_autotune: coordinate descent timed with HIP events,
with every tuning point's distances checked identical on the device).  This module keys the result by
graph statistics -- the ones the reference's delta heuristic uses (n, m, mean weight; wl.h:395-449,
kernel.cu:238-257) plus the weight type and the number of sources -- and stores it in a JSON cache.
"""
import ctypes
import json
import os

import numpy as np

from . import _lib as L

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_cache.json")
TUNED_FIELDS = ("delta", "wg_per_cu", "chunk", "nbuckets", "parts", "concurrency", "task_edges", "flags",
                "light_delta", "edge_windows", "tail_items", "hub_degree")


def workload_key(host, k):
    sample = host.w[:: max(1, host.m // 65536)] if host.m else np.zeros(1)
    mean_w = float(np.mean(sample.astype(np.float64)))
    kind = "f32" if host.is_float else "u32"
    return f"n={host.n};m={host.m};w={kind};meanw={mean_w:.4g};k={k}"


def _load(path):
    if path is None:  # no cache
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache_path, key, params, ms):
    if cache_path is None:
        return
    cache = _load(cache_path)
    cache[key] = {"params": params, "kernel_ms": ms}
    try:
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass


def tune(graph, host, srcs, trials=3, cache_path=DEFAULT_CACHE, refresh=False):
    """Returns the tuned parameter dict for solving `srcs` on `graph` (cached by workload key)."""
    srcs = np.ascontiguousarray(np.atleast_1d(srcs), dtype=np.uint32)
    key = workload_key(host, len(srcs))
    cache = _load(cache_path)
    if not refresh and key in cache:
        return dict(cache[key]["params"])
    best = L.Params()
    ms = ctypes.c_double()
    L.check(L._lib().adds_autotune(graph.handle, srcs.ctypes.data_as(L.c_u32p), len(srcs), trials,
                                   ctypes.byref(best), ctypes.byref(ms)), "autotune")
    params = {f: getattr(best, f) for f in TUNED_FIELDS}
    _store(cache_path, key, params, ms.value)
    return params


def tune_batched(graphs, hosts, srcs, trials=3, cache_path=DEFAULT_CACHE, refresh=False):
    """Tuned parameters for adds_sssp_batched over these graphs and sources (the block-diagonal union
    is what gets tuned); cached under the list of per-graph keys."""
    srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
    key = "batched:" + "|".join(workload_key(h, 1) for h in hosts)
    cache = _load(cache_path)
    if not refresh and key in cache:
        return dict(cache[key]["params"])
    hs = (ctypes.c_void_p * len(graphs))(*[g.handle.value for g in graphs])
    best = L.Params()
    ms = ctypes.c_double()
    L.check(L._lib().adds_autotune_batched(hs, srcs.ctypes.data_as(L.c_u32p), len(graphs), trials,
                                           ctypes.byref(best), ctypes.byref(ms)), "autotune_batched")
    params = {f: getattr(best, f) for f in TUNED_FIELDS}
    _store(cache_path, key, params, ms.value)
    return params
