"""CPU baselines of the reference's evaluation (run_all.sh:77-107): single-thread Dijkstra (Galois
-algo=dijkstra -t 1) and parallel delta-stepping (Galois -algo=deltaStep -t NUM_THREADS), as a shared
library built by this directory's Makefile.  They are comparison legs, timed by run_all.py and the
benchmark; the GPU library never calls them."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None


def _lib():
    global _L
    if _L is None:
        path = os.path.join(HERE, "libcpu_sssp.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C {HERE})")
        L = ctypes.CDLL(path)
        p = ctypes.c_void_p
        u32 = ctypes.c_uint32
        L.cpu_dijkstra.restype = ctypes.c_int
        L.cpu_dijkstra.argtypes = [ctypes.c_int, u32, p, p, p, u32, p, p]
        L.cpu_delta_step.restype = ctypes.c_int
        L.cpu_delta_step.argtypes = [ctypes.c_int, u32, p, p, p, u32, ctypes.c_double, ctypes.c_uint, p, p]
        _L = L
    return _L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _csr(row_ptr, col, w, is_float):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    wb = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32) if is_float else \
        np.ascontiguousarray(w, dtype=np.uint32)
    return row_ptr, col, wb


def dijkstra(row_ptr, col, w, src, is_float):
    """Distances (u32 words: int values, or float bit patterns) and the settled-vertex count."""
    row_ptr, col, wb = _csr(row_ptr, col, w, is_float)
    n = len(row_ptr) - 1
    d = np.empty(n, dtype=np.uint32)
    work = ctypes.c_uint64()
    if _lib().cpu_dijkstra(int(is_float), n, _p(row_ptr), _p(col), _p(wb), int(src), _p(d), ctypes.byref(work)):
        raise ValueError(f"source {src} out of range")
    return d, work.value


def delta_step(row_ptr, col, w, src, is_float, delta, threads=None):
    """Parallel delta-stepping with bucket width `delta` on `threads` threads (default: every CPU this
    process may run on).  Returns the distances as u32 words and the non-stale requests processed."""
    row_ptr, col, wb = _csr(row_ptr, col, w, is_float)
    n = len(row_ptr) - 1
    if not 0 <= int(src) < n:
        raise ValueError(f"source {src} out of range")
    threads = threads or len(os.sched_getaffinity(0))
    d = np.empty(n, dtype=np.uint32)
    work = ctypes.c_uint64()
    _lib().cpu_delta_step(int(is_float), n, _p(row_ptr), _p(col), _p(wb), int(src), float(delta), int(threads),
                          _p(d), ctypes.byref(work))
    return d, work.value
