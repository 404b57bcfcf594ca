"""GPU parity of pull phases: a settled bucket's heavy edges relaxed from their targets (symmetric graphs).

Every case compares against the oracle's Dijkstra bit for bit and checks from the solve's counters that
a pull phase really ran (or, where it must not, that none did).  ADDS_F_PULL_EAGER makes the solver take
a pull at the first heavy phase whatever its cost estimate, so that small graphs exercise the path.
"""
import numpy as np
import pytest
import torch

import adds_amd
import oracle
from adds_amd import _lib as L

pytestmark = pytest.mark.gpu

TIMEOUT = 60000
EAGER = L.F_PULL_EAGER


def _check(g, dist, srcs, is_float, saturate=False):
    ref = oracle.dijkstra(g.row_ptr, g.col, g.w, srcs, is_float, saturate=saturate)
    got = np.ascontiguousarray(dist).view(np.uint32)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: got {got[bad[:5]]} want {ref[bad[:5]]}"


def _symmetrize(row, col, w):
    """Both directions of every edge, one weight per unordered pair (the smaller of the two)."""
    n = len(row) - 1
    u = np.repeat(np.arange(n, dtype=np.int64), np.diff(row.astype(np.int64)))
    v = col.astype(np.int64)
    a, b = np.minimum(u, v), np.maximum(u, v)
    keep = a != b
    a, b, ww = a[keep], b[keep], w[keep]
    order = np.lexsort((ww, b, a))
    a, b, ww = a[order], b[order], ww[order]
    first = np.ones(len(a), dtype=bool)
    first[1:] = (a[1:] != a[:-1]) | (b[1:] != b[:-1])
    a, b, ww = a[first], b[first], ww[first]
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    wts = np.concatenate([ww, ww])
    o = np.lexsort((dst, src))
    src, dst, wts = src[o], dst[o], wts[o]
    rp = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(rp, src + 1, 1)
    rp = np.cumsum(rp).astype(np.uint32)
    return rp, dst.astype(np.uint32), wts


@pytest.mark.parametrize("is_float", [False, True])
@pytest.mark.parametrize("params", [
    {},
    dict(flags=EAGER),
    dict(flags=EAGER | L.F_NO_WAIT | L.F_NO_LOCAL, edge_windows=8, chunk=256, parts=4),
    dict(flags=EAGER, parts=1, nbuckets=4, chunk=16),
    dict(flags=EAGER, tail_items=4096, concurrency=4),
    dict(flags=EAGER, hub_degree=64, task_edges=64),
    dict(flags=EAGER, pull_delta=4.0),  # re-bucketed after the pull: 4 (float) or 16 (int) times wider
    dict(flags=EAGER | L.F_NO_WAIT | L.F_NO_LOCAL, edge_windows=8, pull_delta=4.0, tail_items=4096),
    dict(flags=EAGER, pull_delta=0.5),  # narrower after the pull
    dict(flags=EAGER | L.F_PULL_ALL_LIGHT, pull_delta=4.0),  # ... and every edge light after the pull
    dict(flags=EAGER | L.F_PULL_ALL_LIGHT | L.F_NO_WAIT, edge_windows=8, pull_delta=64.0, hub_degree=64,
         task_edges=32),
])
def test_rmat_pull_identical(is_float, params):
    """RMAT from its highest-degree vertex: bucket 0 holds most of the graph's edges.  With the default
    cost model a graph this size may or may not pull; with EAGER it always does."""
    wk = dict(wkind="f32", lo=0.0, hi=1.0) if is_float else dict(wkind="u32", lo=1, hi=65535)
    g = adds_amd.HostGraph.rmat(15, 32, seed=21, **wk)
    assert g.symmetric
    src = g.max_degree_vertex()
    delta = 0.0625 if is_float else 4096.0
    params = dict(params)
    if "pull_delta" in params:
        params["pull_delta"] *= delta * (1 if is_float else 4)
    dist, st = adds_amd.Graph(g).sssp(src, timeout_ms=TIMEOUT, delta=delta, **params)
    _check(g, dist, src, is_float)
    if params.get("flags", 0) & EAGER:
        assert st["pulls"] == 1 and st["pull_edges"] > 0 and st["pull_improved"] > 0, st


def test_rmat18_default_cost_model_pulls():
    """RMAT-18 ef32 from the hub with the default rule: bucket 0's heavy edges far outnumber the unsettled
    vertices' heavy in-edges, so the solver pulls without being told to, and scans fewer edges than the
    push direction would."""
    g = adds_amd.HostGraph.rmat(18, 32, seed=1, wkind="u32", lo=1, hi=65535)
    src = g.max_degree_vertex()
    dg = adds_amd.Graph(g)
    dist, st = dg.sssp(src, timeout_ms=TIMEOUT, delta=4096.0)
    _check(g, dist, src, False)
    assert st["pulls"] == 1, st
    dpush, sp = dg.sssp(src, timeout_ms=TIMEOUT, delta=4096.0, flags=L.F_NO_PULL)
    _check(g, dpush, src, False)
    assert sp["pulls"] == 0 and sp["pull_edges"] == 0
    assert st["edges_scanned"] < sp["edges_scanned"] / 2, (st["edges_scanned"], sp["edges_scanned"])


def test_low_degree_source_long_unsettled_rows():
    """From a low-degree vertex the hubs are still unsettled when the first heavy phase opens: their
    rows (thousands of heavy edges) go through the long-row tasks, each applied with an atomicMin."""
    g = adds_amd.HostGraph.rmat(15, 32, seed=5, wkind="u32", lo=1, hi=65535)
    deg = g.degrees()
    cand = np.nonzero((deg >= 2) & (deg <= 4))[0]
    src = int(cand[len(cand) // 2])
    assert deg.max() > 4 * 1024  # the hub's row spans several tasks
    for delta in (512.0, 4096.0):
        dist, st = adds_amd.Graph(g).sssp(src, timeout_ms=TIMEOUT, delta=delta, flags=EAGER)
        _check(g, dist, src, False)
        assert st["pulls"] == 1


@pytest.mark.parametrize("flags,pull_delta", [(EAGER, 0.0), (EAGER | L.F_REPLICA_PARTS, 0.0), (EAGER, 16.0),
                                              (EAGER | L.F_PULL_ALL_LIGHT, 16.0)])
def test_multi_source_pull(flags, pull_delta):
    """Many sources in one solve (replicated distance rows): a pull covers every instance's row."""
    g = adds_amd.HostGraph.rmat(13, 16, seed=9, wkind="f32", lo=1.0, hi=10.0)
    srcs = g.sample_component(g.max_degree_vertex(), 12, seed=4)
    dm, st = adds_amd.Graph(g).sssp_multi(srcs, timeout_ms=TIMEOUT, delta=2.0, parts=4, flags=flags,
                                          pull_delta=pull_delta)
    assert st["pulls"] == 1
    for i, s in enumerate(srcs):
        _check(g, dm[i], s, True)


def test_batched_union_pull():
    """Graphs of uneven size in one union solve: the union of symmetric graphs is symmetric."""
    gs = [adds_amd.HostGraph.rmat(11, 16, seed=30 + i, wkind="u32", lo=1, hi=1000) for i in range(3)]
    gs += [adds_amd.HostGraph.grid(40, 50, 10, seed=7, wkind="u32", lo=1, hi=1000),
           adds_amd.HostGraph.er(3000, 8.0, seed=8, wkind="u32", lo=1, hi=1000)]
    dgs = [adds_amd.Graph(h) for h in gs]
    srcs = [h.max_degree_vertex() for h in gs]
    for pd, fl in ((0.0, EAGER), (512.0, EAGER), (512.0, EAGER | L.F_PULL_ALL_LIGHT)):
        dists, st = adds_amd.sssp_batched(dgs, srcs, timeout_ms=TIMEOUT, delta=64.0, flags=fl, pull_delta=pd)
        assert st["pulls"] == 1
        for h, d, s in zip(gs, dists, srcs):
            _check(h, d, s, False)


def test_undeclared_and_reweighted_graphs_do_not_pull():
    """A graph not declared symmetric never pulls; nor does one whose weights were replaced (new weights
    need not be symmetric) until it is declared again."""
    g = adds_amd.HostGraph.rmat(13, 16, seed=3, wkind="f32", lo=1.0, hi=10.0)
    src = g.max_degree_vertex()
    plain = adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, g.w, "f32")
    assert not plain.symmetric
    d0, st0 = adds_amd.Graph(plain).sssp(src, timeout_ms=TIMEOUT, delta=2.0, flags=EAGER)
    assert st0["pulls"] == 0
    _check(g, d0, src, True)
    dg = adds_amd.Graph(g)
    _, st1 = dg.sssp(src, timeout_ms=TIMEOUT, delta=2.0, flags=EAGER)
    assert st1["pulls"] == 1
    # asymmetric new weights: each direction of a pair gets its own value
    rng = np.random.default_rng(1)
    w2 = (rng.random(g.m).astype(np.float32) * 9 + 1).astype(np.float32)
    wt = torch.from_numpy(w2).cuda()
    dg.set_weights_device(wt.data_ptr())
    torch.cuda.synchronize()
    d2, st2 = dg.sssp(src, timeout_ms=TIMEOUT, delta=2.0, flags=EAGER)
    assert st2["pulls"] == 0
    h2 = adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, w2, "f32")
    _check(h2, d2, src, True)


def test_captured_pull_replays_and_weight_update():
    """A captured solve with a pull phase replays bit-exact; after set_weights_device it stops pulling
    (the graph's device word is cleared) and still gives the new weights' distances."""
    g = adds_amd.HostGraph.rmat(13, 16, seed=6, wkind="u32", lo=1, hi=4000)
    srcs = g.sample_component(g.max_degree_vertex(), 4, seed=3)
    dg = adds_amd.Graph(g)
    dist = torch.empty((len(srcs), g.n), dtype=torch.int32, device="cuda")
    sd = torch.from_numpy(srcs.astype(np.int32)).cuda()
    sg = adds_amd.SolveGraph(dg, sd.data_ptr(), len(srcs), dist.data_ptr(), delta=256.0, flags=EAGER,
                             pull_delta=2048.0)
    ref = None
    for i in range(3):
        dist.fill_(5)
        torch.cuda.synchronize()
        sg.launch()
        st = sg.status()
        assert st["pulls"] == 1, st
        if ref is None:
            ref = dist.clone()
            for j, s in enumerate(srcs):
                _check(g, ref[j].cpu().numpy(), s, False)
        assert torch.equal(dist, ref), f"replay {i} differs"
    # symmetric new weights (the same rule per unordered pair): distances change, pulls stop
    g2 = adds_amd.HostGraph.rmat(13, 16, seed=6, wkind="u32", lo=1, hi=4000).reweight(seed=77, wkind="u32", lo=1,
                                                                                     hi=4000)
    wt = torch.from_numpy(g2.w_bits.view(np.int32).copy()).cuda()
    dg.set_weights_device(wt.data_ptr())
    torch.cuda.synchronize()
    sg.launch()
    st = sg.status()
    assert st["pulls"] == 0
    for j, s in enumerate(srcs):
        _check(g2, dist[j].cpu().numpy(), s, False)
    sg.close()


@pytest.mark.parametrize("params", [dict(flags=EAGER), dict(flags=EAGER | L.F_NO_WAIT, delta=1.0)])
def test_saturating_sums_in_pull(params):
    """The saturation case made symmetric: a pull reads unreached neighbours (no path) and saturated
    ones (2^32-2 plus anything stays 2^32-2) exactly as the push direction and the oracle do."""
    from test_oracle import saturation_case

    row, col, w, _ = saturation_case()
    rp, c2, w2 = _symmetrize(np.asarray(row), np.asarray(col), np.asarray(w, dtype=np.uint32))
    g = adds_amd.HostGraph.from_arrays(rp, c2, w2, "u32", symmetric=True)
    assert g.check_symmetric()
    dg = adds_amd.Graph(g)
    for s in range(g.n):
        dist, _ = dg.sssp(s, timeout_ms=TIMEOUT, **params)
        _check(g, dist, s, False, saturate=True)


def test_disconnected_pull_keeps_unreached():
    """Vertices of another component stay unreached: an in-edge from an unreached vertex is no path."""
    a = adds_amd.HostGraph.rmat(12, 16, seed=2, wkind="u32", lo=1, hi=65535)
    n = a.n
    rp = np.concatenate([a.row_ptr, a.row_ptr[1:] + a.m]).astype(np.uint32)
    col = np.concatenate([a.col, a.col + n]).astype(np.uint32)
    w = np.concatenate([a.w_bits, a.w_bits]).astype(np.uint32)
    g = adds_amd.HostGraph.from_arrays(rp, col, w, "u32", symmetric=True)
    src = a.max_degree_vertex()
    dist, st = adds_amd.Graph(g).sssp(src, timeout_ms=TIMEOUT, delta=1024.0, flags=EAGER)
    assert st["pulls"] == 1
    _check(g, dist, src, False)
    assert np.all(dist.view(np.uint32)[n:] == L.INF_U32)


def test_device_symmetry_check():
    """adds_graph_check_symmetric agrees with the exact host check: a generated graph passes (and then
    pulls); one changed weight, one row out of order, or a one-way edge fails (and never pulls)."""
    g = adds_amd.HostGraph.rmat(12, 16, seed=4, wkind="u32", lo=1, hi=65535)
    src = g.max_degree_vertex()
    plain = adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, g.w_bits, "u32")
    dg = adds_amd.Graph(plain)
    assert dg.check_symmetric()
    dist, st = dg.sssp(src, timeout_ms=TIMEOUT, delta=2048.0, flags=EAGER)
    assert st["pulls"] == 1
    _check(g, dist, src, False)
    w = g.w_bits.copy()
    w[7] += 1
    assert not adds_amd.Graph(adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, w, "u32")).check_symmetric()
    # swap the first two entries (destination and weight) of vertex src's row: same edges, unsorted row
    c, w = g.col.copy(), g.w_bits.copy()
    r0 = int(g.row_ptr[src])
    c[r0], c[r0 + 1] = c[r0 + 1], c[r0]
    w[r0], w[r0 + 1] = w[r0 + 1], w[r0]
    h = adds_amd.HostGraph.from_arrays(g.row_ptr, c, w, "u32")
    assert h.check_symmetric()  # as an edge multiset it is symmetric
    dh = adds_amd.Graph(h)
    assert not dh.check_symmetric()
    _, st = dh.sssp(src, timeout_ms=TIMEOUT, delta=2048.0, flags=EAGER)
    assert st["pulls"] == 0
    # one-way edge: drop the last edge of the source's row
    keep = np.ones(g.m, dtype=bool)
    keep[int(g.row_ptr[src + 1]) - 1] = False
    rp = g.row_ptr.astype(np.int64)
    rp[src + 1:] -= 1
    d1 = adds_amd.Graph(adds_amd.HostGraph.from_arrays(rp.astype(np.uint32), g.col[keep], g.w_bits[keep], "u32"))
    assert not d1.check_symmetric()


SPLIT = L.F_PULL_SPLIT


@pytest.mark.parametrize("is_float", [False, True])
@pytest.mark.parametrize("params", [
    dict(flags=SPLIT),
    dict(flags=SPLIT, pull_delta=16.0),
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT, pull_delta=16.0),
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT | L.F_NO_WAIT | L.F_NO_LOCAL, edge_windows=8, pull_delta=128.0),
    dict(flags=SPLIT, tail_items=4096, concurrency=4, pull_delta=16.0),
    dict(flags=SPLIT, parts=1, nbuckets=4, chunk=16),  # the ring narrower than the split bucket
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT, hub_degree=64, task_edges=64, pull_delta=8.0),
    dict(flags=SPLIT, light_scale=3.3),  # a light threshold that is no multiple of the bucket width
    # ordered split: only vertices of at least order_degree light edges wait for their bucket
    dict(flags=SPLIT, order_degree=4),
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT, order_degree=16, pull_delta=16.0, concurrency=1),
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT | L.F_NO_WAIT, order_degree=2, edge_windows=8, pull_delta=64.0,
         tail_items=4096, hub_degree=64, task_edges=32),
    dict(flags=SPLIT, order_degree=1 << 30, parts=2, nbuckets=8),  # no vertex qualifies: none waits
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT | L.F_LOCAL_SPARSE, order_degree=8, pull_delta=16.0, chunk=64),
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT, order_degree=8, pull_delta=16.0, parts=4, concurrency=13),
    dict(flags=SPLIT, order_degree=8, parts=2, concurrency=8, nbuckets=8),  # concurrency capped at nb - 2
    # early pull: queued items carried over into the wide buckets
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT, order_degree=8, pull_delta=16.0, pull_early=4096),
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT, pull_delta=16.0, pull_early=1 << 30),  # at the first chance
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT | L.F_LOCAL_SPARSE | L.F_NO_WAIT, order_degree=4, pull_delta=64.0,
         pull_early=256, hub_degree=64, task_edges=32, parts=4, concurrency=8),
    dict(flags=SPLIT, order_degree=8, pull_early=4096),  # no re-bucketing: pull_early is ignored
    dict(flags=SPLIT | L.F_PULL_ALL_LIGHT | L.F_FULL_FIRST | L.F_LOCAL_SPARSE, order_degree=8, pull_delta=32.0,
         pull_early=2048, parts=4, concurrency=8),
])
def test_rmat_split_first_bucket(is_float, params):
    """ADDS_F_PULL_SPLIT: the buckets below the light threshold run their light phases only, then one
    forced pull relaxes all their heavy edges (counted as one pull, whatever the cost estimate says)."""
    wk = dict(wkind="f32", lo=0.0, hi=1.0) if is_float else dict(wkind="u32", lo=1, hi=65535)
    g = adds_amd.HostGraph.rmat(15, 32, seed=21, **wk)
    src = g.max_degree_vertex()
    delta = 0.0078125 if is_float else 512.0
    params = dict(params)
    light = delta * params.pop("light_scale", 8.0)
    if "pull_delta" in params:
        params["pull_delta"] *= delta
    dist, st = adds_amd.Graph(g).sssp(src, timeout_ms=TIMEOUT, delta=delta, light_delta=light, **params)
    _check(g, dist, src, is_float)
    assert st["pulls"] == 1 and st["pull_edges"] > 0, st


def test_split_first_bucket_runs_dry_before_the_pull():
    """Nothing queued below the split point: the head still stops at it (the solve cannot end while heavy
    edges wait for the pull), the pull reaches the vertex only a heavy edge leads to, and an isolated
    source ends with everything else unreached."""
    # path 0 - 1 - 2 (weights 1) and a heavy edge 2 - 3 (weight 100000); vertex 4 isolated
    rp = np.array([0, 1, 3, 5, 6, 6], dtype=np.uint32)
    col = np.array([1, 0, 2, 1, 3, 2], dtype=np.uint32)
    w = np.array([1, 1, 1, 1, 100000, 100000], dtype=np.uint32)
    g = adds_amd.HostGraph.from_arrays(rp, col, w, "u32", symmetric=True)
    assert g.check_symmetric()
    dg = adds_amd.Graph(g)
    for params in (dict(), dict(pull_delta=4096.0), dict(pull_delta=4096.0, flags=L.F_PULL_ALL_LIGHT),
                   dict(tail_items=4096), dict(order_degree=2),
                   dict(pull_delta=4096.0, flags=L.F_PULL_ALL_LIGHT, pull_early=1 << 30)):
        params = dict(params)
        params["flags"] = params.get("flags", 0) | SPLIT
        for src in range(g.n):
            dist, st = dg.sssp(src, timeout_ms=TIMEOUT, delta=64.0, light_delta=1024.0, **params)
            _check(g, dist, src, False)
            assert st["pulls"] == 1, (src, st)  # an isolated source too: the pull finds nothing
    # the pull's vertex: reached through the heavy edge only
    dist, _ = dg.sssp(0, timeout_ms=TIMEOUT, delta=64.0, light_delta=1024.0, flags=SPLIT)
    assert list(dist.view(np.uint32)[:4]) == [0, 1, 2, 100002]


@pytest.mark.parametrize("flags,order,early", [(SPLIT, 0, 0), (SPLIT | L.F_REPLICA_PARTS | L.F_PULL_ALL_LIGHT, 0, 0),
                                               (SPLIT | L.F_REPLICA_PARTS | L.F_PULL_ALL_LIGHT, 3, 0),
                                               (SPLIT | L.F_PULL_ALL_LIGHT, 3, 2048)])
def test_multi_source_split(flags, order, early):
    g = adds_amd.HostGraph.rmat(13, 16, seed=9, wkind="f32", lo=1.0, hi=10.0)
    srcs = g.sample_component(g.max_degree_vertex(), 12, seed=4)
    dm, st = adds_amd.Graph(g).sssp_multi(srcs, timeout_ms=TIMEOUT, delta=0.5, light_delta=4.0, parts=4,
                                          flags=flags, pull_delta=16.0 if flags & L.F_PULL_ALL_LIGHT else 0.0,
                                          order_degree=order, pull_early=early)
    assert st["pulls"] == 1, st
    for i, s in enumerate(srcs):
        _check(g, dm[i], s, True)


def test_captured_split_replays_and_weight_update():
    """A captured solve with a split first bucket replays bit-exact; once the weights change (and the
    graph is no longer declared symmetric) the same capture runs every heavy phase by pushing."""
    g = adds_amd.HostGraph.rmat(13, 16, seed=6, wkind="u32", lo=1, hi=4000)
    src = np.array([g.max_degree_vertex()], dtype=np.int32)
    dg = adds_amd.Graph(g)
    dist = torch.empty((1, g.n), dtype=torch.int32, device="cuda")
    sd = torch.from_numpy(src).cuda()
    sg = adds_amd.SolveGraph(dg, sd.data_ptr(), 1, dist.data_ptr(), delta=32.0, light_delta=256.0,
                             flags=SPLIT | L.F_PULL_ALL_LIGHT, pull_delta=2048.0, order_degree=4, pull_early=2048)
    ref = None
    for i in range(3):
        dist.fill_(5)
        torch.cuda.synchronize()
        sg.launch()
        st = sg.status()
        assert st["pulls"] == 1, st
        if ref is None:
            ref = dist.clone()
            _check(g, ref[0].cpu().numpy(), int(src[0]), False)
        assert torch.equal(dist, ref), f"replay {i} differs"
    g2 = adds_amd.HostGraph.rmat(13, 16, seed=6, wkind="u32", lo=1, hi=4000).reweight(seed=78, wkind="u32", lo=1,
                                                                                     hi=4000)
    wt = torch.from_numpy(g2.w_bits.view(np.int32).copy()).cuda()
    dg.set_weights_device(wt.data_ptr())
    torch.cuda.synchronize()
    sg.launch()
    assert sg.status()["pulls"] == 0
    _check(g2, dist[0].cpu().numpy(), int(src[0]), False)
    sg.close()


def test_split_flag_without_symmetry_is_ignored():
    g = adds_amd.HostGraph.rmat(12, 16, seed=4, wkind="u32", lo=1, hi=65535)
    plain = adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, g.w_bits, "u32")
    src = g.max_degree_vertex()
    dist, st = adds_amd.Graph(plain).sssp(src, timeout_ms=TIMEOUT, delta=256.0, light_delta=2048.0, flags=SPLIT)
    assert st["pulls"] == 0
    _check(g, dist, src, False)


def _hip_to_host(ptr, nbytes):
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.empty(nbytes, dtype=np.uint8)
    assert hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 2) == 0
    return out


@pytest.mark.parametrize("is_float,light", [(False, 2048.0), (False, 1.0), (True, 0.0625), (True, 1e-4)])
def test_split_rows_light_first_then_bands(is_float, light):
    """The solver's copy of the rows: every row holds the same (col, w) pairs as the input, light ones
    first, then the heavy ones in non-decreasing weight band (what lets a pull phase stop early); rows of
    more than 1024 edges (the workgroup-per-row split) included."""
    import ctypes
    wk = dict(wkind="f32", lo=0.0, hi=1.0) if is_float else dict(wkind="u32", lo=1, hi=65535)
    g = adds_amd.HostGraph.rmat(14, 32, seed=8, **wk)
    deg = g.degrees()
    assert deg.max() > 1024
    dg = adds_amd.Graph(g)
    cw, lc = ctypes.c_void_p(), ctypes.c_void_p()
    L.check(L._lib().adds_graph_split_rows(dg.handle, ctypes.c_float(light), ctypes.byref(cw), ctypes.byref(lc)),
            "split rows")
    pairs = _hip_to_host(cw.value, g.m * 8).view(np.uint32).reshape(-1, 2)
    lcnt = _hip_to_host(lc.value, g.n * 4).view(np.uint32)
    col, wb = pairs[:, 0], pairs[:, 1]
    lf = np.float32(light)
    wf = wb.view(np.float32) if is_float else wb.astype(np.float32)
    is_l = (wb.view(np.float32) <= lf) if is_float else (wb <= np.uint32(int(light)))
    band = np.minimum(np.maximum((wf - lf) * (np.float32(1.0) / lf), np.float32(0)), np.float32(31)).astype(np.int64)
    rows = np.repeat(np.arange(g.n), deg)
    pos = np.arange(g.m) - g.row_ptr[:-1].astype(np.int64)[rows]
    # light exactly in the first lcnt positions of each row
    assert np.array_equal(is_l, pos < lcnt[rows].astype(np.int64))
    assert np.array_equal(lcnt, np.bincount(rows, weights=is_l, minlength=g.n).astype(np.uint32))
    # heavy bands non-decreasing inside each row
    same_row = rows[1:] == rows[:-1]
    both_heavy = ~is_l[1:] & ~is_l[:-1]
    assert np.all(band[1:][same_row & both_heavy] >= band[:-1][same_row & both_heavy])
    # the same multiset of pairs per row
    key_got = np.lexsort((wb, col, rows))
    key_in = np.lexsort((g.w_bits, g.col, rows))
    assert np.array_equal(col[key_got], g.col[key_in]) and np.array_equal(wb[key_got], g.w_bits[key_in])
