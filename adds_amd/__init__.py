"""adds_amd: MI355X-native asynchronous dynamic delta-stepping SSSP (ADDS) behind a C ABI.

The solver, graph I/O and generators live in the native library adds_amd/lib/libadds.so
(include/adds.h); this package is a thin ctypes layer over it for tests, benchmarks and the autograd
binding.
"""
from . import _lib
from ._lib import AddsError, Params, Stats, INF_U32, INF_F32_BITS, LIB_PATH
from .graph import HostGraph, Graph, SolveGraph, sssp_batched, write_dist_text

__all__ = [
    "AddsError", "Params", "Stats", "INF_U32", "INF_F32_BITS", "LIB_PATH",
    "HostGraph", "Graph", "SolveGraph", "sssp_batched", "write_dist_text", "load",
]


def load():
    """Loads the native library (raises ImportError when it has not been built)."""
    return _lib._lib()
