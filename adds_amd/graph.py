"""Host graphs (generators, .gr files, distance text) and device-resident graph handles.

This is the host-side mirror of the reference's CSRGraph (ads_int/csr_graph.h:20-97): ``HostGraph``
is the host CSR that CSRGraph::read produces (csr_graph.cu:203-298) and ``Graph`` is the
device-resident copy that copy_to_gpu makes (csr_graph.cu:309-328), solved through the C ABI.
"""
import ctypes

import numpy as np

from . import _lib as L

WKINDS = {"u32": L.WK_U32_UNIFORM, "f32": L.WK_F32_UNIFORM, "lognormal": L.WK_F32_LOGNORMAL}


def _wtype(w):
    if w in (L.W_U32, "int", "u32", np.uint32):
        return L.W_U32
    if w in (L.W_F32, "float", "f32", np.float32):
        return L.W_F32
    raise ValueError(f"unknown weight type {w!r}")


class HostGraph:
    """A CSR graph in host memory (row_ptr[n+1], col[m], 4-byte weights w[m])."""

    def __init__(self, csr: L.HostCSR, symmetric=False):
        self._csr = csr
        # every edge (u, v, w) has its reverse (v, u, w): the generators build such graphs (one weight per
        # undirected pair); files and arrays are not assumed to be (see check_symmetric)
        self.symmetric = bool(symmetric)
        n, m = csr.n, csr.m
        self.n, self.m = int(n), int(m)
        self.wtype = int(csr.wtype)
        self.row_ptr = np.ctypeslib.as_array(csr.row_ptr, shape=(n + 1,))
        self.col = np.ctypeslib.as_array(csr.col, shape=(max(m, 1),))[:m]
        wbits = np.ctypeslib.as_array(csr.w, shape=(max(m, 1),))[:m]
        self.w_bits = wbits
        self.w = wbits.view(np.float32) if self.wtype == L.W_F32 else wbits

    def __del__(self):
        csr = getattr(self, "_csr", None)
        if csr is None or not csr.row_ptr:
            return
        try:
            L._lib().adds_host_csr_free(ctypes.byref(csr))
        except (TypeError, AttributeError):  # interpreter shutdown: module globals already cleared
            return
        self._csr = None

    # ---- construction ----
    @classmethod
    def _from_call(cls, fn, *args, what="generator", symmetric=False):
        csr = L.HostCSR()
        L.check(fn(*args, ctypes.byref(csr)), what)
        return cls(csr, symmetric=symmetric)

    @classmethod
    def rmat(cls, scale, edge_factor, seed, wkind="u32", lo=1, hi=100, a=0.57, b=0.19, c=0.19):
        return cls._from_call(L._lib().adds_gen_rmat, scale, edge_factor, a, b, c, seed, WKINDS[wkind],
                              float(lo), float(hi), what="rmat", symmetric=True)

    @classmethod
    def grid(cls, rows, cols, shortcuts, seed, wkind="f32", lo=1.0, hi=2.0):
        return cls._from_call(L._lib().adds_gen_grid, rows, cols, shortcuts, seed, WKINDS[wkind], float(lo),
                              float(hi), what="grid", symmetric=True)

    @classmethod
    def er(cls, n, avg_degree, seed, wkind="u32", lo=1, hi=1000):
        return cls._from_call(L._lib().adds_gen_er, n, float(avg_degree), seed, WKINDS[wkind], float(lo),
                              float(hi), what="er", symmetric=True)

    @classmethod
    def read_gr(cls, path, wtype="int"):
        return cls._from_call(L._lib().adds_read_gr, str(path).encode(), _wtype(wtype), what=str(path))

    @classmethod
    def read_mtx(cls, path, wtype="float"):
        """MatrixMarket coordinate file (SuiteSparse); weights |value| (see adds_read_mtx)."""
        return cls._from_call(L._lib().adds_read_mtx, str(path).encode(), _wtype(wtype), what=str(path))

    @classmethod
    def read_dimacs(cls, path, wtype="int"):
        """DIMACS shortest-path file ("p sp n m" / "a u v w")."""
        return cls._from_call(L._lib().adds_read_dimacs, str(path).encode(), _wtype(wtype), what=str(path))

    @classmethod
    def from_arrays(cls, row_ptr, col, w, wtype, symmetric=False):
        """A host graph holding its own copies of the given arrays (they may be views into another graph,
        which can then go away); `symmetric` declares that every edge has its reverse with the same weight."""
        wt = _wtype(wtype)
        row_ptr = np.array(row_ptr, dtype=np.uint32, copy=True, order="C")
        col = np.array(col, dtype=np.uint32, copy=True, order="C")
        wb = np.array(w, dtype=np.float32 if wt == L.W_F32 else np.uint32, copy=True, order="C").view(np.uint32)
        obj = cls.__new__(cls)
        obj._csr = None
        obj.n, obj.m, obj.wtype = len(row_ptr) - 1, len(col), wt
        obj.row_ptr, obj.col, obj.w_bits = row_ptr, col, wb
        obj.w = wb.view(np.float32) if wt == L.W_F32 else wb
        obj.symmetric = bool(symmetric)
        return obj

    def _csr_view(self):
        c = L.HostCSR()
        c.n, c.m = self.n, self.m
        c.row_ptr = self.row_ptr.ctypes.data_as(L.c_u32p)
        c.col = self.col.ctypes.data_as(L.c_u32p)
        c.w = self.w_bits.ctypes.data_as(L.c_u32p)
        c.wtype = self.wtype
        return c

    def reweight(self, seed, wkind="f32", lo=0.0, hi=1.0):
        """Re-draws all weights in place (same rule as the generators) and updates the weight type."""
        if self._csr is None:
            raise ValueError("reweight needs a library-owned graph")
        L.check(L._lib().adds_gen_weights(ctypes.byref(self._csr), seed, WKINDS[wkind], float(lo), float(hi)),
                "reweight")
        self.wtype = int(self._csr.wtype)
        self.w = self.w_bits.view(np.float32) if self.wtype == L.W_F32 else self.w_bits
        return self

    # ---- utilities ----
    def check_symmetric(self):
        """True when every edge (u, v, w) has its reverse (v, u, w) with the same weight bits, counting
        multiplicities (an exact test: the sorted (u, v, w) triples equal the sorted (v, u, w) triples)."""
        u = np.repeat(np.arange(self.n, dtype=np.uint64), np.diff(self.row_ptr.astype(np.int64)))
        v = self.col.astype(np.uint64)
        wb = self.w_bits.astype(np.uint64)
        a = np.lexsort((wb, v, u))
        b = np.lexsort((wb, u, v))
        return bool(np.array_equal(u[a], v[b]) and np.array_equal(v[a], u[b]) and np.array_equal(wb[a], wb[b]))

    @property
    def is_float(self):
        return self.wtype == L.W_F32

    def degrees(self):
        return np.diff(self.row_ptr.astype(np.int64))

    def write_gr(self, path):
        c = self._csr_view()
        L.check(L._lib().adds_write_gr(str(path).encode(), ctypes.byref(c)), str(path))

    def max_degree_vertex(self):
        c = self._csr_view()
        return int(L._lib().adds_max_degree_vertex(ctypes.byref(c)))

    def sample_component(self, root, k, seed):
        out = np.zeros(k, dtype=np.uint32)
        c = self._csr_view()
        got = L._lib().adds_sample_component(ctypes.byref(c), root, k, seed, out.ctypes.data_as(L.c_u32p))
        return out[:got]

    def reached_counts(self, dist_bits):
        d = np.ascontiguousarray(dist_bits).view(np.uint32)
        c = self._csr_view()
        nr, mr = ctypes.c_uint64(), ctypes.c_uint64()
        L._lib().adds_reached_counts(ctypes.byref(c), d.ctypes.data_as(L.c_u32p), ctypes.byref(nr),
                                     ctypes.byref(mr))
        return int(nr.value), int(mr.value)


def write_dist_text(path, dist, wtype):
    d = np.ascontiguousarray(dist).view(np.uint32)
    L.check(L._lib().adds_write_dist_text(str(path).encode(), d.ctypes.data_as(L.c_u32p), len(d),
                                          _wtype(wtype)), str(path))


def _stream_ptr(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


class Graph:
    """A device-resident graph handle (adds_graph) with the solver entry points."""

    def __init__(self, host: HostGraph, device=0):
        L.check(L._lib().adds_set_device(device), "device")
        h = ctypes.c_void_p()
        L.check(L._lib().adds_graph_create(host.row_ptr.ctypes.data_as(L.c_u32p),
                                           host.col.ctypes.data_as(L.c_u32p),
                                           host.w_bits.ctypes.data_as(ctypes.c_void_p),
                                           host.n, host.m, host.wtype, ctypes.byref(h)), "graph")
        self._h = h
        self.n, self.m, self.wtype = host.n, host.m, host.wtype
        self.host = host
        if getattr(host, "symmetric", False):
            L.check(L._lib().adds_graph_set_symmetric(h, 1), "graph")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            L._lib().adds_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except (TypeError, AttributeError):  # interpreter shutdown: module globals already cleared
            pass

    @property
    def handle(self):
        return self._h

    def check_symmetric(self):
        """Tests on the device whether the graph is symmetric with strictly ascending rows and declares it
        accordingly (pull phases follow the declaration); returns the outcome."""
        out = ctypes.c_int()
        L.check(L._lib().adds_graph_check_symmetric(self._h, ctypes.byref(out)), "check_symmetric")
        return bool(out.value)

    @property
    def dist_dtype(self):
        return np.float32 if self.wtype == L.W_F32 else np.uint32

    # ---- solves with host output ----
    def sssp(self, src, **params):
        """Single-source solve; returns (dist as numpy, stats dict)."""
        out = np.empty(self.n, dtype=np.uint32)
        p = L.Params.make(**params)
        p.flags |= L.F_HOST_OUT
        st = L.Stats()
        L.check(L._lib().adds_sssp(self._h, src, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(p),
                                   ctypes.byref(st), None), "sssp")
        return out.view(self.dist_dtype), st.as_dict()

    def sssp_nearfar(self, src, delta=0.0):
        """Bulk-synchronous near-far solve (the reference's NF comparison baseline); (dist, stats)."""
        out = np.empty(self.n, dtype=np.uint32)
        st = L.Stats()
        L.check(L._lib().adds_sssp_nearfar(self._h, src, out.ctypes.data_as(ctypes.c_void_p), float(delta),
                                           L.F_HOST_OUT, ctypes.byref(st), None), "sssp_nearfar")
        return out.view(self.dist_dtype), st.as_dict()

    def sssp_nearfar_device(self, src, dist_ptr, delta=0.0):
        """Near-far solve into a device buffer of n words; returns the stats dict."""
        st = L.Stats()
        L.check(L._lib().adds_sssp_nearfar(self._h, src, ctypes.c_void_p(dist_ptr), float(delta), 0,
                                           ctypes.byref(st), None), "sssp_nearfar")
        return st.as_dict()

    def sssp_multi(self, srcs, **params):
        srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
        out = np.empty((len(srcs), self.n), dtype=np.uint32)
        p = L.Params.make(**params)
        p.flags |= L.F_HOST_OUT
        st = L.Stats()
        L.check(L._lib().adds_sssp_multi(self._h, srcs.ctypes.data_as(L.c_u32p), len(srcs),
                                         out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(p),
                                         ctypes.byref(st), None), "sssp_multi")
        return out.view(self.dist_dtype), st.as_dict()

    # ---- solves into device buffers (torch tensors as plumbing) ----
    def sssp_device(self, src, dist_ptr, stream=None, **params):
        p = L.Params.make(**params)
        st = L.Stats()
        L.check(L._lib().adds_sssp(self._h, src, ctypes.c_void_p(dist_ptr), ctypes.byref(p), ctypes.byref(st),
                                   _stream_ptr(stream)), "sssp")
        return st.as_dict()

    def sssp_multi_device(self, srcs, dist_ptr, stream=None, **params):
        srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
        p = L.Params.make(**params)
        st = L.Stats()
        L.check(L._lib().adds_sssp_multi(self._h, srcs.ctypes.data_as(L.c_u32p), len(srcs),
                                         ctypes.c_void_p(dist_ptr), ctypes.byref(p), ctypes.byref(st),
                                         _stream_ptr(stream)), "sssp_multi")
        return st.as_dict()

    def set_weights_device(self, w_ptr, stream=None):
        L.check(L._lib().adds_graph_set_weights_device(self._h, ctypes.c_void_p(w_ptr), _stream_ptr(stream)),
                "set_weights")

    def backward_device(self, srcs, dist_ptr, gdist_ptr, gw_ptr, stream=None):
        """dL/dw into gw_ptr ([m] float32) from dist [k][n] (solved from srcs) and dL/d(dist) [k][n], all
        device pointers; returns the pass's time in ms."""
        srcs = np.ascontiguousarray(np.atleast_1d(srcs), dtype=np.uint32)
        ms = ctypes.c_double()
        L.check(L._lib().adds_sssp_backward_f32(self._h, srcs.ctypes.data_as(L.c_u32p), len(srcs),
                                                ctypes.c_void_p(dist_ptr), ctypes.c_void_p(gdist_ptr),
                                                ctypes.c_void_p(gw_ptr), ctypes.byref(ms), _stream_ptr(stream)),
                "backward")
        return ms.value


def sssp_batched(graphs, srcs, **params):
    """One launch over k (graph, source) instances; returns a list of numpy dist arrays."""
    k = len(graphs)
    srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
    outs = [np.empty(g.n, dtype=np.uint32) for g in graphs]
    hs = (ctypes.c_void_p * k)(*[g.handle.value for g in graphs])
    ds = (ctypes.c_void_p * k)(*[o.ctypes.data for o in outs])
    p = L.Params.make(**params)
    p.flags |= L.F_HOST_OUT
    st = L.Stats()
    L.check(L._lib().adds_sssp_batched(hs, srcs.ctypes.data_as(L.c_u32p), k, ds, ctypes.byref(p),
                                       ctypes.byref(st), None), "sssp_batched")
    return [o.view(g.dist_dtype) for o, g in zip(outs, graphs)], st.as_dict()


class SolveGraph:
    """A solve captured in a HIP graph over fixed device buffers: k sources on one graph
    (``SolveGraph(graph, srcs_ptr, k, dist_ptr)``) or one source on each of several graphs
    (``SolveGraph.batched(graphs, srcs, dist_ptrs)``, one persistent launch over their union)."""

    def __init__(self, graph: Graph, srcs_ptr, k, dist_ptr, **params):
        p = L.Params.make(**params)
        h = ctypes.c_void_p()
        L.check(L._lib().adds_solve_graph_create(graph.handle, ctypes.c_void_p(srcs_ptr), k, ctypes.c_void_p(dist_ptr),
                                                 ctypes.byref(p), ctypes.byref(h)), "solve_graph_create")
        self._h = h
        self.graph = graph

    @classmethod
    def batched(cls, graphs, srcs, dist_ptrs, **params):
        k = len(graphs)
        srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
        hs = (ctypes.c_void_p * k)(*[g.handle.value for g in graphs])
        ds = (ctypes.c_void_p * k)(*[int(d) for d in dist_ptrs])
        p = L.Params.make(**params)
        h = ctypes.c_void_p()
        L.check(L._lib().adds_solve_graph_create_batched(hs, srcs.ctypes.data_as(L.c_u32p), k, ds, ctypes.byref(p),
                                                         ctypes.byref(h)), "solve_graph_create_batched")
        obj = cls.__new__(cls)
        obj._h = h
        obj.graph = graphs[0]
        obj.graphs = list(graphs)  # the members stay alive while the capture exists
        return obj

    def launch(self, stream=None):
        L.check(L._lib().adds_solve_graph_launch(self._h, _stream_ptr(stream)), "solve_graph_launch")

    def status(self):
        st = L.Stats()
        L.check(L._lib().adds_solve_graph_status(self._h, ctypes.byref(st)), "solve_graph")
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            L._lib().adds_solve_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except (TypeError, AttributeError):  # interpreter shutdown: module globals already cleared
            pass
