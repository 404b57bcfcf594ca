"""GPU parity: device distances equal CPU Dijkstra bit for bit (int and float), through the C ABI.

Every case here compares against the oracle's Dijkstra directly; the full-size configurations are in
test_gpu_fullsize.py.
"""
import numpy as np
import pytest

import adds_amd
import oracle
from adds_amd import _lib as L

pytestmark = pytest.mark.gpu

TIMEOUT = 60000


def _check(g, dist, srcs, is_float):
    ref = oracle.dijkstra(g.row_ptr, g.col, g.w, srcs, is_float)
    got = np.ascontiguousarray(dist).view(np.uint32)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: got {got[bad[:5]]} want {ref[bad[:5]]}"


def test_rmat16_int_c1():
    """Config 1: RMAT scale 16 ef 16, uint32 weights in [1,100], source = highest-degree vertex."""
    g = adds_amd.HostGraph.rmat(16, 16, seed=1, wkind="u32", lo=1, hi=100)
    src = g.max_degree_vertex()
    dg = adds_amd.Graph(g)
    dist, st = dg.sssp(src, timeout_ms=TIMEOUT)
    print("c1 stats", st)
    _check(g, dist, src, False)


def test_rmat16_float():
    g = adds_amd.HostGraph.rmat(16, 16, seed=2, wkind="f32", lo=0.0, hi=1.0)
    src = g.max_degree_vertex()
    dist, st = adds_amd.Graph(g).sssp(src, timeout_ms=TIMEOUT)
    _check(g, dist, src, True)


def test_grid_float_small():
    g = adds_amd.HostGraph.grid(256, 256, 64, seed=5, wkind="f32", lo=1.0, hi=2.0)
    dist, st = adds_amd.Graph(g).sssp(0, timeout_ms=TIMEOUT)
    _check(g, dist, 0, True)


def test_grid2048_float_c2():
    """Config 2: 2048x2048 grid + 8385 shortcut pairs, float weights in [1,2), source 0."""
    g = adds_amd.HostGraph.grid(2048, 2048, 8385, seed=2, wkind="f32", lo=1.0, hi=2.0)
    dist, st = adds_amd.Graph(g).sssp(0, timeout_ms=TIMEOUT)
    print("c2 stats", st)
    _check(g, dist, 0, True)


@pytest.mark.parametrize("params", [
    dict(parts=1, nbuckets=4, chunk=1, wg_per_cu=1),
    dict(parts=8, nbuckets=256, chunk=64, wg_per_cu=8, concurrency=3),
    dict(parts=4, nbuckets=8, chunk=48, wg_per_cu=2, concurrency=4),
    dict(flags=0x8 | 0x4, delta=64.0),  # no local continuation, uncached probes
    dict(flags=0x10, delta=64.0),  # flushes do not wait for their atomicMins
    dict(flags=0x10 | 0x8, edge_windows=8, chunk=256, parts=4),
    dict(parts=2, nbuckets=16, chunk=17, delta=1.0, hub_degree=64, task_edges=64),
    dict(parts=4, nbuckets=32, chunk=32, delta=100000.0, concurrency=2),
    dict(parts=4, nbuckets=8, chunk=32, delta=1000.0, concurrency=4),
    dict(parts=2, chunk=8, delta=50000.0, flags=0x8),
    dict(delta=100.0, light_delta=1000.0, parts=4),  # light threshold above the bucket width
    dict(delta=64.0, light_delta=1e30),  # every edge light: no heavy phases
    dict(edge_windows=8, parts=4, concurrency=4),
    dict(tail_items=1 << 30, delta=64.0),  # one bucket once the pending work is a quarter of its peak
    dict(tail_items=4096, delta=16.0, parts=2),
    dict(flags=0x800, parts=4, concurrency=8, delta=64.0),  # full queues first, 8 open buckets
    dict(flags=0x800 | 0x400 | 0x10, parts=2, concurrency=13, chunk=64),  # sparse-only local continuation
])
def test_param_sweep_identical(params):
    """Results never depend on the tuning point (delta, buckets, parts, chunk, hub split)."""
    g = adds_amd.HostGraph.rmat(14, 16, seed=11, wkind="u32", lo=1, hi=1000)
    src = g.max_degree_vertex()
    dist, st = adds_amd.Graph(g).sssp(src, timeout_ms=TIMEOUT, **params)
    _check(g, dist, src, False)


def test_multi_source_replicated():
    g = adds_amd.HostGraph.rmat(13, 8, seed=4, wkind="f32", lo=1.0, hi=10.0)
    srcs = g.sample_component(g.max_degree_vertex(), 16, seed=9)
    dist, st = adds_amd.Graph(g).sssp_multi(srcs, timeout_ms=TIMEOUT)
    for i, s in enumerate(srcs):
        _check(g, dist[i], s, True)


@pytest.mark.parametrize("params", [dict(parts=8), dict(parts=4, edge_windows=8, flags=0x10),
                                    dict(parts=2, flags=0x8), dict(parts=8, light_delta=1e30, delta=2.0)])
@pytest.mark.parametrize("wkind", ["f32", "u32"])
def test_multi_source_replica_parts(params, wkind):
    """ADDS_F_REPLICA_PARTS: each source's items go to sub-queue part (source index mod parts), whatever
    part the pushing wavefront works on; distances are unchanged, on a power-law graph and on a grid
    (where hops leave the wavefront's own part)."""
    lo, hi = (1.0, 10.0) if wkind == "f32" else (1, 1000)
    p = dict(params)
    p["flags"] = p.get("flags", 0) | L.F_REPLICA_PARTS
    for g in (adds_amd.HostGraph.rmat(13, 8, seed=14, wkind=wkind, lo=lo, hi=hi),
              adds_amd.HostGraph.grid(90, 70, 0, seed=15, wkind=wkind, lo=lo, hi=hi)):
        srcs = g.sample_component(0, 19, seed=3)
        dist, st = adds_amd.Graph(g).sssp_multi(srcs, timeout_ms=TIMEOUT, **p)
        for i, s in enumerate(srcs):
            _check(g, dist[i], s, g.is_float)


@pytest.mark.parametrize("wkind", ["u32", "f32"])
def test_split_row_shapes(wkind):
    """The light-first rows at the boundaries of their two build paths: rows of 1023, 1024 (the last
    taken 64 to a wavefront) and 1025 edges (the first taken one to a workgroup), runs of empty rows,
    a row count that is not a multiple of 64, and 64-row groups of more than 16384 edges (built as
    consecutive sub-groups).  Every row mixes light and heavy edges; distances from several sources must
    equal Dijkstra's."""
    rng = np.random.default_rng(77)
    n = 64 * 40 + 37
    deg = rng.integers(0, 40, n)
    deg[rng.random(n) < 0.2] = 0
    deg[64:128] = 300                     # a group of 19200 edges
    deg[200:264] = rng.integers(200, 1100, 64)
    deg[400], deg[401], deg[402], deg[403] = 1023, 1024, 1025, 5000
    deg[500:505] = 0
    deg[n - 3:] = [1025, 0, 7]
    row = np.zeros(n + 1, dtype=np.uint32)
    row[1:] = np.cumsum(deg)
    m = int(row[-1])
    src = np.repeat(np.arange(n), deg)
    col = rng.integers(0, n, m).astype(np.uint32)
    col = np.where(col == src, (col + 1) % n, col).astype(np.uint32)
    if wkind == "f32":
        w = rng.uniform(0.1, 10.0, m).astype(np.float32)
        delta = 2.0
    else:
        w = rng.integers(1, 1000, m).astype(np.uint32)
        delta = 150.0
    g = adds_amd.HostGraph.from_arrays(row, col, w, wkind)
    dg = adds_amd.Graph(g)
    srcs = [0, 65, 401, n - 3]
    for params in (dict(delta=delta), dict(delta=delta, edge_windows=8, hub_degree=512)):
        dist, _ = dg.sssp_multi(srcs, timeout_ms=TIMEOUT, **params)
        for i, s in enumerate(srcs):
            _check(g, dist[i], s, wkind == "f32")


def test_batched_union_graphs():
    gs = [adds_amd.HostGraph.grid(40 + 7 * i, 33, 0, seed=i, wkind="u32", lo=1, hi=1000) if i % 2 == 0 else
          adds_amd.HostGraph.er(3000 + 500 * i, 8, seed=i, wkind="u32", lo=1, hi=1000) for i in range(6)]
    dgs = [adds_amd.Graph(g) for g in gs]
    srcs = [0, 1, 2, 3, 4, 5]
    dists, st = adds_amd.sssp_batched(dgs, srcs, timeout_ms=TIMEOUT)
    for g, d, s in zip(gs, dists, srcs):
        _check(g, d, s, False)


def test_overflow_reports_error_not_hang():
    """A pool too small for the solve returns EOVERFLOW (forced from the host; every wave exits)."""
    g = adds_amd.HostGraph.rmat(12, 16, seed=1, wkind="u32", lo=1, hi=100)
    dg = adds_amd.Graph(g)
    with pytest.raises(adds_amd.AddsError) as ei:
        dg.sssp(g.max_degree_vertex(), pool_entries=4096, timeout_ms=TIMEOUT)
    assert ei.value.status == adds_amd._lib.EOVERFLOW
    # the handle is usable afterwards
    dist, _ = dg.sssp(g.max_degree_vertex(), timeout_ms=TIMEOUT)
    _check(g, dist, g.max_degree_vertex(), False)


@pytest.mark.parametrize("params", [{}, dict(flags=0x10, edge_windows=8), dict(delta=1.0, light_delta=1e30)])
def test_int_sums_saturate_like_the_oracle(params):
    """Integer path sums past 2^32-2 saturate at ADDS_SAT_U32 in adds_sssp, adds_sssp_multi and the
    near-far solver, exactly as the oracle does; a wrapping add would give vertex 2 the distance 0x10
    through the detour 0 -> 5 -> 2."""
    from test_oracle import saturation_case

    row, col, w, want = saturation_case()
    g = adds_amd.HostGraph.from_arrays(row, col, w, "u32")
    ref = oracle.dijkstra(row, col, w, 0, False, saturate=True)
    assert np.array_equal(ref, want)
    dg = adds_amd.Graph(g)
    dist, _ = dg.sssp(0, timeout_ms=TIMEOUT, **params)
    assert np.array_equal(dist.view(np.uint32), want)
    dm, _ = dg.sssp_multi([0, 1, 5], timeout_ms=TIMEOUT, **params)
    for i, s in enumerate([0, 1, 5]):
        assert np.array_equal(dm[i].view(np.uint32), oracle.dijkstra(row, col, w, s, False, saturate=True))
    dn, _ = dg.sssp_nearfar(0)
    assert np.array_equal(dn.view(np.uint32), want)


def test_repeated_solves_same_handle():
    g = adds_amd.HostGraph.rmat(13, 16, seed=3, wkind="u32", lo=1, hi=100)
    dg = adds_amd.Graph(g)
    for s in [0, 5, 17, g.max_degree_vertex()]:
        dist, _ = dg.sssp(s, timeout_ms=TIMEOUT)
        _check(g, dist, s, False)


@pytest.mark.parametrize("is_float", [False, True])
def test_long_rows_light_heavy_partition(is_float):
    """Two hubs with 20000 and 6000 out-edges (longer than one wavefront handles in the light/heavy
    split) plus a ring: distances from hub 0 are the direct edge weights wherever the direct edge is
    shortest, so a misplaced (col, w) pair in the partition shows up as a wrong distance."""
    n = 30000
    rng = np.random.default_rng(11)
    src, dst = [], []
    for v in range(1, 20001):
        src.append(0), dst.append(v)
    for v in range(20001, 26001):
        src.append(1), dst.append(v)
    for v in range(n):
        src.append(v), dst.append((v + 1) % n)
    src, dst = np.array(src, dtype=np.int64), np.array(dst, dtype=np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    row_ptr = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(row_ptr, src + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.uint32)
    if is_float:
        w = rng.uniform(0.0, 1000.0, size=len(dst)).astype(np.float32)
    else:
        w = rng.integers(1, 1000, size=len(dst)).astype(np.uint32)
    g = adds_amd.HostGraph.from_arrays(row_ptr, dst.astype(np.uint32), w, "f32" if is_float else "u32")
    dg = adds_amd.Graph(g)
    for delta in (16.0, 500.0, 4096.0):  # mostly heavy, mixed, all light
        dist, _ = dg.sssp(0, delta=delta, timeout_ms=TIMEOUT)
        _check(g, dist, 0, is_float)
    # 16-edge chunks: the 20000-edge hub needs 1250 chunks, over three consecutive range records
    dist, st = dg.sssp(0, delta=16.0, task_edges=16, hub_degree=64, timeout_ms=TIMEOUT)
    assert st["hub_tasks"] >= 1250
    _check(g, dist, 0, is_float)


def test_weights_update_resplits():
    """New weights through the device setter are seen by later solves (the light-first copy of the rows is
    rebuilt), with the same and with a different bucket width."""
    import torch

    g = adds_amd.HostGraph.rmat(14, 16, seed=12, wkind="f32", lo=1.0, hi=10.0)
    src = g.max_degree_vertex()
    dg = adds_amd.Graph(g)
    dist, _ = dg.sssp(src, delta=2.0, timeout_ms=TIMEOUT)
    _check(g, dist, src, True)
    rng = np.random.default_rng(3)
    # independent per-edge weights (the two directions of an edge may now differ)
    w2 = rng.uniform(0.5, 20.0, size=g.m).astype(np.float32)
    wt = torch.from_numpy(w2).cuda()
    dg.set_weights_device(wt.data_ptr())
    torch.cuda.synchronize()
    g2 = adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, w2, "f32")
    for delta in (2.0, 8.0):
        dist, _ = dg.sssp(src, delta=delta, timeout_ms=TIMEOUT)
        _check(g2, dist, src, True)


@pytest.mark.parametrize("params", [dict(), dict(delta=16.0), dict(delta=16.0, flags=0x8),
                                    dict(delta=2.0, light_delta=1e30, edge_windows=8),
                                    dict(delta=2.0, light_delta=1e30, edge_windows=8, flags=0x10 | 0x8),
                                    dict(delta=2.0, tail_items=20000)])
def test_multi_source_grid(params):
    """Many sources on a high-diameter graph: replicated instances share the buckets, and hops inside a
    bucket go through local continuation (on by default)."""
    g = adds_amd.HostGraph.grid(200, 150, 0, seed=21, wkind="f32", lo=0.5, hi=3.0)
    srcs = g.sample_component(0, 12, seed=4)
    dist, st = adds_amd.Graph(g).sssp_multi(srcs, timeout_ms=TIMEOUT, **params)
    for i, s in enumerate(srcs):
        _check(g, dist[i], s, True)


def test_tail_mode_waits_for_the_peak():
    """Tail mode (the rest of a solve as one bucket) starts only once the pending work has fallen to a
    quarter of its peak.  Sources on one edge of a grid start with a small frontier: entering tail mode
    at the first head move would run the whole solve unordered, with many times the expansions."""
    g = adds_amd.HostGraph.grid(256, 256, 0, seed=22, wkind="f32", lo=1.0, hi=2.0)
    srcs = np.arange(0, g.n, g.n // 8, dtype=np.uint32)  # column 0, every 32nd row
    dg = adds_amd.Graph(g)
    _, st0 = dg.sssp_multi(srcs, timeout_ms=TIMEOUT, delta=4.0)
    dist, st = dg.sssp_multi(srcs, timeout_ms=TIMEOUT, delta=4.0, tail_items=4096)
    for i, s in enumerate(srcs):
        _check(g, dist[i], s, True)
    print("expansions without / with tail mode", st0["vertices_expanded"], st["vertices_expanded"])
    assert st["vertices_expanded"] <= 1.5 * st0["vertices_expanded"]


def _csr(n, edges, wtype):
    """CSR from (u, v, w) triples, rows in the given order."""
    edges = sorted(edges, key=lambda e: e[0])
    row = np.zeros(n + 1, dtype=np.uint32)
    for u, _, _ in edges:
        row[u + 1] += 1
    row = np.cumsum(row).astype(np.uint32)
    col = np.array([v for _, v, _ in edges], dtype=np.uint32)
    w = np.array([x for _, _, x in edges], dtype=np.float32 if wtype == "float" else np.uint32)
    return adds_amd.HostGraph.from_arrays(row, col, w, wtype)


def _zero_heavy_rmat(is_float):
    g = adds_amd.HostGraph.rmat(12, 8, seed=31, wkind="f32" if is_float else "u32", lo=0 if is_float else 1,
                                hi=1.0 if is_float else 50)
    rng = np.random.default_rng(31)
    w = g.w.copy()
    w[rng.random(g.m) < 0.3] = 0  # a third of the edges weigh nothing
    return adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, w, "float" if is_float else "int")


EDGE_CASES = {
    "single_vertex": lambda: (_csr(1, [], "int"), 0),
    "zero_weights_self_loops_parallel_edges": lambda: (_csr(5, [(0, 1, 0), (1, 2, 0), (2, 3, 5), (1, 1, 3),
                                                                (0, 2, 7), (0, 2, 1), (3, 3, 0), (3, 4, 0),
                                                                (4, 0, 0)], "int"), 0),
    "source_without_out_edges": lambda: (_csr(4, [(1, 2, 3), (2, 3, 4), (3, 1, 1)], "int"), 0),
    "extreme_float_weights": lambda: (_csr(6, [(0, 1, 1e-30), (1, 2, 0.0), (2, 3, 1e30), (0, 3, 3e30),
                                               (3, 4, 1e-38), (0, 4, 2e30), (4, 5, 1.5), (5, 0, 0.0)],
                                          "float"), 0),
    "zero_heavy_rmat_int": lambda: (_zero_heavy_rmat(False), 0),
    "zero_heavy_rmat_float": lambda: (_zero_heavy_rmat(True), 0),
}


@pytest.mark.parametrize("case", sorted(EDGE_CASES))
def test_edge_cases(case):
    """Single vertex, zero weights (ties inside a bucket), self loops, parallel edges, a source with no
    out-edges, float weights from 1e-38 to 3e30: ADDS (default, then a narrow bucket width), several sources
    at once and the near-far baseline all equal Dijkstra bit for bit."""
    g, src = EDGE_CASES[case]()
    dg = adds_amd.Graph(g)
    # a narrow bucket width too, except where distances span 1e30 (10^33 buckets of 1e-3 would take the
    # head forever to walk)
    narrow = [] if case == "extreme_float_weights" else [dict(delta=1.0 if not g.is_float else 1e-3)]
    for params in [dict()] + narrow:
        dist, _ = dg.sssp(src, timeout_ms=TIMEOUT, **params)
        _check(g, dist, src, g.is_float)
    srcs = np.array([src, g.n - 1], dtype=np.uint32)
    dist, _ = dg.sssp_multi(srcs, timeout_ms=TIMEOUT)
    for i, s in enumerate(srcs):
        _check(g, dist[i], s, g.is_float)
    dist, _ = dg.sssp_nearfar(src)
    _check(g, dist, src, g.is_float)


@pytest.mark.parametrize("is_float", [False, True])
def test_pair_windows_odd_ranges(is_float):
    """The 8-window form loads 16-byte edge pairs: ranges that start or end on an odd edge (odd task sizes,
    short hub splits, light/heavy boundaries inside rows, the last edge of the array) mask the half that
    lies outside; single and several sources, against Dijkstra bit for bit."""
    g = adds_amd.HostGraph.rmat(14, 12, seed=77, wkind="f32" if is_float else "u32", lo=0 if is_float else 1,
                                hi=1.0 if is_float else 999)
    dg = adds_amd.Graph(g)
    src = g.max_degree_vertex()
    delta = 0.05 if is_float else 64.0
    for params in (dict(edge_windows=8, task_edges=37, hub_degree=64, delta=delta),
                   dict(edge_windows=8, task_edges=1, hub_degree=64, delta=delta, chunk=7),
                   dict(edge_windows=8, light_delta=1e30, delta=delta, flags=0x10)):
        dist, _ = dg.sssp(src, timeout_ms=TIMEOUT, **params)
        _check(g, dist, src, is_float)
        srcs = g.sample_component(src, 5, seed=3)
        dm, _ = dg.sssp_multi(srcs, timeout_ms=TIMEOUT, **params)
        for i, s in enumerate(srcs):
            _check(g, dm[i], s, is_float)
