"""Persisted autotuning of the solver's parameters (delta, light threshold, bucket ring, sub-queues and
open buckets, workgroups per CU, claim size, range split, tail mode, flags) per workload.

The search itself runs in the native library (adds_autotune: coordinate descent timed with HIP events,
with every tuning point's distances checked identical on the device).  This module keys the result by
graph statistics -- the ones the reference's delta heuristic uses (n, m, mean weight; wl.h:395-449,
kernel.cu:238-257) plus the weight type and the number of sources -- and looks it up in two JSON files of
the same format ({key: {"params": {...}, "kernel_ms": ms}}):

* tuned_params.json (committed): the parameters pinned for the benchmark workloads, consulted first, so
  that every run and every profile of those workloads measures the same point.  The drop-in binaries
  (sssp_int / sssp_float) read it too.
* tune_cache.json (not committed): results of tuning runs on this machine.

workload_key() must stay in step with the key the binaries compute (sssp_main.cpp, tuned_key).
"""
import ctypes
import json
import os

import numpy as np

from . import _lib as L

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(_HERE, "tune_cache.json")
PINNED = os.path.join(_HERE, "tuned_params.json")
TUNED_FIELDS = ("delta", "wg_per_cu", "chunk", "nbuckets", "parts", "concurrency", "task_edges", "flags",
                "light_delta", "edge_windows", "tail_items", "hub_degree", "pull_delta",
                "order_degree", "pull_early")


def workload_key(host, k):
    """n, m, weight type, the mean of every max(1, m // 65536)-th weight (4 significant digits) and the
    number of sources."""
    sample = host.w[:: max(1, host.m // 65536)] if host.m else np.zeros(1)
    mean_w = float(np.mean(sample.astype(np.float64)))
    kind = "f32" if host.is_float else "u32"
    return f"n={host.n};m={host.m};w={kind};meanw={mean_w:.4g};k={k}"


def batched_key(hosts):
    return "batched:" + "|".join(workload_key(h, 1) for h in hosts)


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


def lookup(key, cache_path=DEFAULT_CACHE, pinned=True):
    """(params, source file) for a workload key, or (None, None)."""
    for path in ((PINNED,) if pinned else ()) + (cache_path,):
        entry = _load(path).get(key)
        if entry:
            return dict(entry["params"]), path
    return None, None


def tune(graph, host, srcs, trials=3, cache_path=DEFAULT_CACHE, refresh=False, pinned=True):
    """Returns the tuned parameter dict for solving `srcs` on `graph`: the pinned or cached entry for its
    workload key, else the result of a new search (stored in cache_path).  refresh=True always searches."""
    srcs = np.ascontiguousarray(np.atleast_1d(srcs), dtype=np.uint32)
    key = workload_key(host, len(srcs))
    if not refresh:
        params, _ = lookup(key, cache_path, pinned)
        if params is not None:
            return params
    best = L.Params()
    ms = ctypes.c_double()
    L.check(L._lib().adds_autotune(graph.handle, srcs.ctypes.data_as(L.c_u32p), len(srcs), trials,
                                   ctypes.byref(best), ctypes.byref(ms)), "autotune")
    params = {f: getattr(best, f) for f in TUNED_FIELDS}
    _store(cache_path, key, params, ms.value)
    return params


def tune_batched(graphs, hosts, srcs, trials=3, cache_path=DEFAULT_CACHE, refresh=False, pinned=True):
    """Tuned parameters for adds_sssp_batched over these graphs and sources (the block-diagonal union
    is what gets tuned); keyed by the list of per-graph keys."""
    srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
    key = batched_key(hosts)
    if not refresh:
        params, _ = lookup(key, cache_path, pinned)
        if params is not None:
            return params
    hs = (ctypes.c_void_p * len(graphs))(*[g.handle.value for g in graphs])
    best = L.Params()
    ms = ctypes.c_double()
    L.check(L._lib().adds_autotune_batched(hs, srcs.ctypes.data_as(L.c_u32p), len(graphs), trials,
                                           ctypes.byref(best), ctypes.byref(ms)), "autotune_batched")
    params = {f: getattr(best, f) for f in TUNED_FIELDS}
    _store(cache_path, key, params, ms.value)
    return params
