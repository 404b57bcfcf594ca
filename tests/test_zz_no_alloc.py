"""The allocation-free entry point (adds_sssp_noalloc) against adds_sssp, through the C ABI.

Two calls in a row -- the second on a smaller graph, with another bucket layout -- must each give
exactly the distances adds_sssp gives (and the oracle's Dijkstra); the scratch keeps the last call's
contents, so the second call also shows that every call clears what must start empty.  Sizes beyond the
scratch, host output and asynchronous calls are refused with a clear error.
"""
import ctypes

import numpy as np
import pytest
import torch

import adds_amd
import oracle
from adds_amd import _lib as L

pytestmark = pytest.mark.gpu


def noalloc(dg, src, dist, **params):
    p = L.Params.make(**params)
    st = L.Stats()
    rc = L._lib().adds_sssp_noalloc(dg.handle, src, ctypes.c_void_p(dist.data_ptr()), ctypes.byref(p),
                                     ctypes.byref(st), None)
    return rc, st.as_dict()


def eager(dg, src, dist, **params):
    dg.sssp_device(src, dist.data_ptr(), **params)
    torch.cuda.synchronize()


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large_then_small"])
def test_noalloc_matches_sssp(order):
    a = adds_amd.HostGraph.rmat(15, 16, seed=21, wkind="u32", lo=1, hi=1000)
    b = adds_amd.HostGraph.grid(200, 150, 40, seed=22, wkind="f32", lo=1.0, hi=2.0)
    cases = [(a, a.max_degree_vertex(), dict()), (b, 7, dict(nbuckets=32, parts=2, delta=4.0))]
    if order != "large_then_small":
        cases = [cases[1], cases[0], (b, 123, dict(edge_windows=8, flags=0x10))]
    for host, src, params in cases:
        print(f"noalloc case n={host.n} m={host.m} src={src} params={params}", flush=True)
        dg = adds_amd.Graph(host)
        ref = torch.empty(host.n, dtype=torch.int32, device="cuda")
        got = torch.full((host.n,), 12345, dtype=torch.int32, device="cuda")
        eager(dg, src, ref, **params)
        rc, st = noalloc(dg, src, got, **params)
        assert rc == L.OK, L._lib().adds_last_error_detail().decode()
        assert st["vertices_expanded"] > 0
        assert torch.equal(got, ref)
        want = oracle.dijkstra(host.row_ptr, host.col, host.w, src, host.is_float)
        assert np.array_equal(got.cpu().numpy().view(np.uint32), want)
        # the same call again on the same handle
        got.fill_(0)
        rc, _ = noalloc(dg, src, got, **params)
        assert rc == L.OK
        assert torch.equal(got, ref)
        dg.close()


def test_noalloc_refuses_what_it_cannot_hold():
    big = adds_amd.HostGraph.grid(1500, 1500, 0, seed=3, wkind="u32", lo=1, hi=10)  # 4n + m/16 > 8M
    dg = adds_amd.Graph(big)
    dist = torch.empty(big.n, dtype=torch.int32, device="cuda")
    rc, _ = noalloc(dg, 0, dist)
    assert rc == L.EINVAL
    assert "allocation-free" in L._lib().adds_last_error_detail().decode()
    dg.close()
    small = adds_amd.HostGraph.rmat(10, 8, seed=1, wkind="u32", lo=1, hi=100)
    ds = adds_amd.Graph(small)
    d2 = torch.empty(small.n, dtype=torch.int32, device="cuda")
    assert noalloc(ds, 0, d2, flags=L.F_HOST_OUT)[0] == L.EINVAL
    assert noalloc(ds, 0, d2, flags=L.F_NO_SYNC)[0] == L.EINVAL
    assert noalloc(ds, 0, d2, nbuckets=256, parts=4)[0] == L.EINVAL  # 1024 queues per phase
    assert noalloc(ds, small.n, d2)[0] == L.EINVAL
    # still usable afterwards
    rc, _ = noalloc(ds, 0, d2)
    assert rc == L.OK
    assert np.array_equal(d2.cpu().numpy().view(np.uint32),
                          oracle.dijkstra(small.row_ptr, small.col, small.w, 0, False))
