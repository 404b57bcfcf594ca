"""Runtime behaviour on the GPU: HIP-graph replays equal eager solves, the cross-XCD litmus test of
the ordering recipe finds no forbidden outcome, the autotuner keeps results identical, and the
asynchronous (no-sync) form reports status later."""
import ctypes

import numpy as np
import pytest
import torch

import adds_amd
import oracle
from adds_amd import _lib as L

pytestmark = pytest.mark.gpu


def _xcd_count():
    """XCDs seen by one wide launch (8 on a whole MI355X in SPX mode)."""
    n = 4096
    xcc = (ctypes.c_uint32 * n)()
    L.check(L._lib().adds_xcd_map(n, 64, xcc), "xcd map")
    return len(set(xcc))


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_cross_xcd_litmus(variant):
    """Every producer/consumer pair really runs on two different XCDs (read from each workgroup's
    XCC_ID register), and no consumer sees a stale payload after the flag."""
    f, lost, cross = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    L.check(L._lib().adds_litmus(256, 2000, variant, ctypes.byref(f), ctypes.byref(lost), ctypes.byref(cross)),
            "litmus")
    assert lost.value == 0
    assert f.value == 0, f"{f.value} forbidden outcomes"
    if _xcd_count() > 1:
        assert cross.value == 256, f"only {cross.value} of 256 pairs crossed XCDs"


def test_litmus_control_without_wait():
    """Variant 3 drops the completion wait; it is a control that reports how often a stale payload is
    seen without it (no assertion: the hardware may or may not reorder in a given run)."""
    f, lost, cross = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    L.check(L._lib().adds_litmus(256, 2000, 3, ctypes.byref(f), ctypes.byref(lost), ctypes.byref(cross)), "litmus")
    assert lost.value == 0
    print(f"litmus without the wait: {f.value} stale payloads in {256 * 2000} rounds, {cross.value} pairs cross-XCD")


@pytest.mark.parametrize("wg_per_cu", [1, 4])
def test_solver_grid_xcd_placement(wg_per_cu):
    """The solver's persistent grid (workgroups_per_cu x CUs workgroups of 256 threads) is dealt
    round-robin over the XCDs: workgroup b runs on the same XCD as workgroup b % nx, and the first nx
    workgroups cover nx different XCDs.  The termination words are grouped by b % 8 on that basis, so
    each group's waves share one XCD's L2."""
    cu = L._lib().adds_device_cu_count()
    grid = cu * wg_per_cu
    xcc = (ctypes.c_uint32 * grid)()
    L.check(L._lib().adds_xcd_map(grid, 256, xcc), "xcd map")
    ids = list(xcc)
    nx = len(set(ids))
    print(f"{grid} workgroups on {nx} XCDs; first 16: {ids[:16]}")
    assert nx in (1, 2, 4, 8)
    assert len(set(ids[:nx])) == nx
    assert all(ids[b] == ids[b % nx] for b in range(grid)), "placement is not round-robin"
    if nx > 1:
        assert 8 % nx == 0  # b % 8 groups then never straddle two XCDs


def test_graph_replays_match_eager():
    host = adds_amd.HostGraph.rmat(14, 16, seed=6, wkind="f32", lo=1.0, hi=10.0)
    srcs = host.sample_component(host.max_degree_vertex(), 8, seed=3)
    dg = adds_amd.Graph(host)
    eager = torch.empty((len(srcs), host.n), dtype=torch.int32, device="cuda")
    dg.sssp_multi_device(srcs, eager.data_ptr())
    ref = eager.clone()
    dist = torch.empty_like(eager)
    sd = torch.from_numpy(srcs.astype(np.int32)).cuda()
    sg = adds_amd.SolveGraph(dg, sd.data_ptr(), len(srcs), dist.data_ptr())
    for i in range(4):
        dist.fill_(7)  # eager work between replays
        torch.cuda.synchronize()
        sg.launch()
        st = sg.status()
        assert torch.equal(dist, ref), f"replay {i} differs"
        assert st["edges_scanned"] > 0
        dg.sssp_multi_device(srcs, eager.data_ptr())  # an eager solve on the same handle in between
        assert torch.equal(eager, ref)
    for i in (0, len(srcs) - 1):
        assert np.array_equal(ref[i].cpu().numpy().view(np.uint32),
                              oracle.dijkstra(host.row_ptr, host.col, host.w, srcs[i], True))


def test_graph_replay_survives_workspace_changes():
    """A capture owns its workspace: eager solves with another bucket ring, sub-queue count, light
    threshold or a pool that must grow (which reallocate the handle's own workspace and re-split its rows)
    leave later replays correct."""
    host = adds_amd.HostGraph.rmat(13, 16, seed=8, wkind="u32", lo=1, hi=1000)
    srcs = host.sample_component(host.max_degree_vertex(), 4, seed=5)
    dg = adds_amd.Graph(host)
    dist = torch.empty((len(srcs), host.n), dtype=torch.int32, device="cuda")
    sd = torch.from_numpy(srcs.astype(np.int32)).cuda()
    sg = adds_amd.SolveGraph(dg, sd.data_ptr(), len(srcs), dist.data_ptr(), nbuckets=64, parts=8, delta=64.0)
    sg.launch()
    sg.status()
    ref = dist.clone()
    for i in range(len(srcs)):
        assert np.array_equal(ref[i].cpu().numpy().view(np.uint32),
                              oracle.dijkstra(host.row_ptr, host.col, host.w, srcs[i], False))
    eager = torch.empty(host.n, dtype=torch.int32, device="cuda")
    for params in (dict(nbuckets=256, parts=2), dict(light_delta=1e30, delta=8.0), dict(pool_entries=1 << 24),
                   dict(nbuckets=16, parts=4, delta=512.0)):
        dg.sssp_device(int(srcs[0]), eager.data_ptr(), **params)
        dist.fill_(3)
        sg.launch()
        sg.status()
        assert torch.equal(dist, ref), f"replay after an eager solve with {params} differs"
    from adds_amd.autotune import tune
    tune(dg, host, [int(srcs[1])], trials=1, cache_path=None, refresh=True)
    dist.fill_(3)
    sg.launch()
    sg.status()
    assert torch.equal(dist, ref)
    sg.close()


def test_batched_graph_capture_replays():
    """Several graphs of uneven size in one captured persistent launch: every replay fills every
    caller buffer with that graph's distances, bit-exact against Dijkstra, with eager work in between."""
    hosts = [adds_amd.HostGraph.grid(30 + 11 * i, 25, 0, seed=i, wkind="u32", lo=1, hi=1000) if i % 2 == 0 else
             adds_amd.HostGraph.er(2000 + 700 * i, 8, seed=i, wkind="u32", lo=1, hi=1000) for i in range(7)]
    dgs = [adds_amd.Graph(h) for h in hosts]
    srcs = [i % 3 for i in range(len(hosts))]
    outs = [torch.full((h.n,), 5, dtype=torch.int32, device="cuda") for h in hosts]
    sg = adds_amd.SolveGraph.batched(dgs, srcs, [o.data_ptr() for o in outs], delta=256.0)
    refs = [oracle.dijkstra(h.row_ptr, h.col, h.w, s, False) for h, s in zip(hosts, srcs)]
    for rep in range(3):
        for o in outs:
            o.fill_(rep)
        torch.cuda.synchronize()
        sg.launch()
        st = sg.status()
        assert st["vertices_expanded"] > 0
        for h, o, r in zip(hosts, outs, refs):
            assert np.array_equal(o.cpu().numpy().view(np.uint32), r), f"replay {rep} differs (n={h.n})"
        adds_amd.sssp_batched(dgs[:3], srcs[:3])  # eager batched work on some of the members in between
    sg.close()


def test_union_cache_sees_member_weight_updates():
    """A batched solve caches the union of its graphs on the first handle; changing the weights of
    another member, or replacing a member with a new handle, must rebuild it."""
    a = adds_amd.HostGraph.er(3000, 8, seed=1, wkind="f32", lo=1.0, hi=10.0)
    b = adds_amd.HostGraph.grid(40, 50, 0, seed=2, wkind="f32", lo=1.0, hi=10.0)
    da, db = adds_amd.Graph(a), adds_amd.Graph(b)
    dists, _ = adds_amd.sssp_batched([da, db], [0, 0])
    assert np.array_equal(dists[1].view(np.uint32), oracle.dijkstra(b.row_ptr, b.col, b.w, 0, True))
    w2 = np.random.default_rng(4).uniform(0.5, 3.0, size=b.m).astype(np.float32)
    wt = torch.from_numpy(w2).cuda()
    db.set_weights_device(wt.data_ptr())
    torch.cuda.synchronize()
    dists, _ = adds_amd.sssp_batched([da, db], [0, 0])
    assert np.array_equal(dists[1].view(np.uint32), oracle.dijkstra(b.row_ptr, b.col, w2, 0, True))
    assert np.array_equal(dists[0].view(np.uint32), oracle.dijkstra(a.row_ptr, a.col, a.w, 0, True))
    # a new member handle (possibly at the old one's address) of a larger graph
    db.close()
    c = adds_amd.HostGraph.grid(60, 70, 0, seed=3, wkind="f32", lo=1.0, hi=10.0)
    dc = adds_amd.Graph(c)
    dists, _ = adds_amd.sssp_batched([da, dc], [0, 5])
    assert np.array_equal(dists[1].view(np.uint32), oracle.dijkstra(c.row_ptr, c.col, c.w, 5, True))


def test_autotune_keeps_results():
    from adds_amd.autotune import tune

    host = adds_amd.HostGraph.rmat(13, 16, seed=8, wkind="u32", lo=1, hi=1000)
    src = host.max_degree_vertex()
    dg = adds_amd.Graph(host)
    params = tune(dg, host, [src], trials=1, cache_path="/tmp/adds_tune_test.json", refresh=True)
    dist, _ = dg.sssp(src, **params)
    assert np.array_equal(dist.view(np.uint32), oracle.dijkstra(host.row_ptr, host.col, host.w, src, False))


def test_no_sync_then_status():
    host = adds_amd.HostGraph.rmat(12, 16, seed=2, wkind="u32", lo=1, hi=100)
    src = host.max_degree_vertex()
    dg = adds_amd.Graph(host)
    dist = torch.empty(host.n, dtype=torch.int32, device="cuda")
    dg.sssp_device(src, dist.data_ptr(), flags=L.F_NO_SYNC)
    st = L.Stats()
    L.check(L._lib().adds_last_status(dg.handle, ctypes.byref(st)), "status")
    assert st.vertices_expanded > 0
    assert np.array_equal(dist.cpu().numpy().view(np.uint32),
                          oracle.dijkstra(host.row_ptr, host.col, host.w, src, False))


def test_source_out_of_range_is_einval():
    host = adds_amd.HostGraph.rmat(8, 4, seed=2)
    dg = adds_amd.Graph(host)
    with pytest.raises(adds_amd.AddsError) as ei:
        dg.sssp(host.n + 5)
    assert ei.value.status == L.EINVAL


def test_autotune_batched_keeps_results():
    from adds_amd.autotune import tune_batched

    hosts = [adds_amd.HostGraph.grid(60 + 10 * i, 40, 0, seed=30 + i, wkind="u32", lo=1, hi=500) if i % 2 == 0 else
             adds_amd.HostGraph.er(4000 + 700 * i, 6, seed=30 + i, wkind="u32", lo=1, hi=500) for i in range(5)]
    dgs = [adds_amd.Graph(h) for h in hosts]
    srcs = [0, 3, 1, 7, 2]
    params = tune_batched(dgs, hosts, srcs, trials=1, cache_path="/tmp/adds_tune_batched_test.json", refresh=True)
    dists, _ = adds_amd.sssp_batched(dgs, srcs, **params)
    for h, d, s in zip(hosts, dists, srcs):
        assert np.array_equal(d.view(np.uint32), oracle.dijkstra(h.row_ptr, h.col, h.w, s, False))


def test_gather_ceiling_reports_a_rate():
    """The random-gather diagnostic returns a plausible rate, higher for a distance array that fits the
    caches than for one that does not, and refuses empty shapes."""
    small, large = ctypes.c_double(), ctypes.c_double()
    L.check(L._lib().adds_gather_ceiling(1 << 16, 1 << 24, ctypes.byref(small)), "gather ceiling")
    L.check(L._lib().adds_gather_ceiling(1 << 24, 1 << 24, ctypes.byref(large)), "gather ceiling")
    print(f"random gathers/s: {small.value:.3e} (256 KiB array), {large.value:.3e} (64 MiB array)")
    assert 1e9 < large.value < small.value < 1e13
    assert L._lib().adds_gather_ceiling(0, 1 << 20, ctypes.byref(small)) == L.EINVAL
