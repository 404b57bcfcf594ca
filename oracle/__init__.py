"""ctypes wrapper of oracle/liboracle.so -- the CPU oracle.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, as the checker; the product (adds_amd/) never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_L = None


def _lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        p = ctypes.c_void_p
        u32 = ctypes.c_uint32
        for name in ("oracle_dijkstra", "oracle_bellman_ford", "oracle_certify"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, u32, p, p, p, p, u32, p]
        L.oracle_tree_grad.restype = ctypes.c_int
        L.oracle_tree_grad.argtypes = [u32, p, p, p, u32, p, p, p, p, p, p]
        L.oracle_grad_scale.restype = ctypes.c_int
        L.oracle_grad_scale.argtypes = [u32, u32, p, p, p]
        _L = L
    return _L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _prep(row_ptr, col, w):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    w = np.ascontiguousarray(w)
    wb = w.view(np.uint32) if w.dtype in (np.float32, np.uint32, np.int32) else w.astype(np.uint32)
    return row_ptr, col, np.ascontiguousarray(wb)


def dijkstra(row_ptr, col, w, srcs, is_float, saturate=False):
    """Distance words (uint32 view) from the sources.  An integer path sum at or above 2^32-1 saturates
    at 2^32-2 (oracle.c relax_add); that raises OverflowError unless saturate=True, which returns the
    saturated solution (the one the device solvers compute)."""
    row_ptr, col, wb = _prep(row_ptr, col, w)
    srcs = np.ascontiguousarray(np.atleast_1d(srcs), dtype=np.uint32)
    n = len(row_ptr) - 1
    d = np.empty(n, dtype=np.uint32)
    if _lib().oracle_dijkstra(int(is_float), n, _p(row_ptr), _p(col), _p(wb), _p(srcs), len(srcs), _p(d)) \
            and not saturate:
        raise OverflowError("integer distance overflow")
    return d


def bellman_ford(row_ptr, col, w, srcs, is_float, saturate=False):
    row_ptr, col, wb = _prep(row_ptr, col, w)
    srcs = np.ascontiguousarray(np.atleast_1d(srcs), dtype=np.uint32)
    n = len(row_ptr) - 1
    d = np.empty(n, dtype=np.uint32)
    if _lib().oracle_bellman_ford(int(is_float), n, _p(row_ptr), _p(col), _p(wb), _p(srcs), len(srcs), _p(d)) \
            and not saturate:
        raise OverflowError("integer distance overflow")
    return d


def certify(row_ptr, col, w, srcs, dist, is_float):
    """0 when dist is exactly the shortest-path solution; 1/2/3 name the violated condition."""
    row_ptr, col, wb = _prep(row_ptr, col, w)
    srcs = np.ascontiguousarray(np.atleast_1d(srcs), dtype=np.uint32)
    d = np.ascontiguousarray(dist).view(np.uint32)
    return int(_lib().oracle_certify(int(is_float), len(row_ptr) - 1, _p(row_ptr), _p(col), _p(wb), _p(srcs),
                                     len(srcs), _p(d)))


def tree_grad(row_ptr, col, w, srcs, dist, gdist):
    """dL/dw (float32[m]), the parent edges ([k][n], 0xFFFFFFFF for roots and unreached vertices) and the
    fixed-point scale S, for forward distances dist [k][n] from srcs and dL/d(dist) gdist [k][n]
    (oracle.c: tree rule with equal-distance rounds, exact 64-bit sums of rint(g * 2^S), one rounding).
    Raises ValueError when dist is not a shortest-path solution from srcs or a reached g is not finite."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    dist = np.ascontiguousarray(np.atleast_2d(dist), dtype=np.float32)
    gdist = np.ascontiguousarray(np.atleast_2d(gdist), dtype=np.float32)
    k, n = dist.shape
    srcs = np.ascontiguousarray(np.atleast_1d(srcs), dtype=np.uint32)
    assert len(srcs) == k
    gw = np.empty(len(col), dtype=np.float32)
    pe = np.empty((k, n), dtype=np.uint32)
    scale = ctypes.c_int()
    rc = _lib().oracle_tree_grad(n, _p(row_ptr), _p(col), _p(w), k, _p(srcs), _p(dist), _p(gdist), _p(gw), _p(pe),
                                 ctypes.byref(scale))
    if rc != 0:
        raise ValueError("dist is not a shortest-path solution from srcs, or a reached gradient is not finite")
    return gw, pe, scale.value


def grad_scale(dist, gdist):
    """The fixed-point scale S of a batch and its number of nonzero reached gradients."""
    dist = np.ascontiguousarray(np.atleast_2d(dist), dtype=np.float32)
    gdist = np.ascontiguousarray(np.atleast_2d(gdist), dtype=np.float32)
    k, n = dist.shape
    nnz = ctypes.c_uint64()
    s = _lib().oracle_grad_scale(n, k, _p(dist), _p(gdist), ctypes.byref(nnz))
    return s, nnz.value
