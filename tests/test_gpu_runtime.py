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


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_cross_xcd_litmus(variant):
    f, lost = ctypes.c_uint64(), ctypes.c_uint64()
    L.check(L._lib().adds_litmus(256, 2000, variant, ctypes.byref(f), ctypes.byref(lost)), "litmus")
    assert lost.value == 0
    assert f.value == 0, f"{f.value} forbidden outcomes"


def test_litmus_control_without_wait():
    """Variant 3 drops the completion wait; it is a control that reports how often a stale payload is
    seen without it (no assertion: the hardware may or may not reorder in a given run)."""
    f, lost = ctypes.c_uint64(), ctypes.c_uint64()
    L.check(L._lib().adds_litmus(256, 2000, 3, ctypes.byref(f), ctypes.byref(lost)), "litmus")
    assert lost.value == 0
    print(f"litmus without the wait: {f.value} stale payloads in {256 * 2000} rounds")


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
