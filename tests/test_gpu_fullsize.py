"""Full-size parity on the benchmark workloads (adds_amd.workloads), through the C ABI.

* C3: RMAT-24 ef32 (16.8M vertices, 1.02G directed edges), uint32 and float32 weights, single source:
  bit-exact against the oracle's Dijkstra at the default point, the pinned point (pull phase and
  re-bucketing) and the pinned point without pulls.
* C4a: 256 sources on RMAT-20 in one replicated launch: every source checked with the oracle's
  certificate (feasible + tight-reachable + INF exactly on the unreachable set, which pins the solution
  uniquely) and 16 of them against Dijkstra bit for bit.
* C4b: 64 graphs of uneven size in one union launch: every graph bit-exact against Dijkstra.
* C5: the backward pass against the committed finite differences (tests/golden/c5_fd.npz, made by
  tests/golden/make_c5_fd.py with scipy in float64): 1000 tree edges, rel tol 1e-3.

The oracle checks run on a thread pool (each is one C call that releases the GIL).
"""
import os
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

import adds_amd
import oracle
from adds_amd import workloads

pytestmark = pytest.mark.gpu

TIMEOUT = 60000
POOL = min(16, os.cpu_count() or 1)
HERE = os.path.dirname(os.path.abspath(__file__))


def _dist_words(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.fixture(scope="module")
def c3_graph():
    g, src = workloads.c3("u32")
    print(f"c3 graph n={g.n} m={g.m} src={src}", flush=True)
    return g, src


@pytest.mark.parametrize("kind", ["u32", "f32"])
def test_c3_full_size_bit_exact(c3_graph, kind):
    """The default parameters, the pinned benchmark point (a pull phase for bucket 0, then the rest of the
    solve re-bucketed at pull_delta) and the pinned point with every heavy phase pushed: all three equal
    Dijkstra bit for bit."""
    from adds_amd import _lib as L, autotune

    g, src = c3_graph
    if kind == "f32":
        workloads.c3_reweight_float(g)
    assert g.is_float == (kind == "f32")
    pinned, _ = autotune.lookup(autotune.workload_key(g, 1))
    assert pinned is not None and pinned.get("pull_delta")
    dg = adds_amd.Graph(g)
    dist = torch.empty(g.n, dtype=torch.int32, device="cuda")
    got = {}
    for name, params in (("default", {}), ("pinned", pinned),
                         ("pinned, push only", dict(pinned, flags=pinned.get("flags", 0) | L.F_NO_PULL))):
        st = dg.sssp_device(src, dist.data_ptr(), timeout_ms=TIMEOUT, **params)
        print(f"c3 {kind} {name}: kernel {st['kernel_ms']:.2f} ms, edges scanned {st['edges_scanned']}, "
              f"pulls {st['pulls']}", flush=True)
        if name == "pinned":
            assert st["pulls"] == 1
        got[name] = _dist_words(dist).copy()
    dg.close()
    del dist
    ref = oracle.dijkstra(g.row_ptr, g.col, g.w, src, g.is_float)
    for name, d in got.items():
        bad = np.nonzero(d != ref)[0]
        assert bad.size == 0, f"{name}: {bad.size} mismatches, first at {bad[:5]}"
    n_r = int(np.count_nonzero(ref != (0x7F800000 if g.is_float else 0xFFFFFFFF)))
    assert n_r > g.n // 2


@pytest.mark.parametrize("point", ["default", "pinned"])
def test_c4a_all_256_sources(point):
    """All 256 sources at the default parameters and at the pinned benchmark point (which places each
    source's items in one sub-queue part, ADDS_F_REPLICA_PARTS)."""
    from adds_amd import autotune

    g, srcs = workloads.c4a()
    assert len(srcs) == 256
    params = {}
    if point == "pinned":
        params, _ = autotune.lookup(autotune.workload_key(g, len(srcs)))
        assert params is not None
    dg = adds_amd.Graph(g)
    dist = torch.empty((len(srcs), g.n), dtype=torch.int32, device="cuda")
    st = dg.sssp_multi_device(srcs, dist.data_ptr(), timeout_ms=TIMEOUT, **params)
    print(f"c4a ({point} {params}): kernel {st['kernel_ms']:.2f} ms", flush=True)
    d = dist.cpu().numpy().view(np.uint32)
    del dist
    dg.close()
    with ThreadPoolExecutor(POOL) as ex:
        certs = list(ex.map(lambda i: oracle.certify(g.row_ptr, g.col, g.w, srcs[i], d[i], True), range(len(srcs))))
        bad = [i for i, c in enumerate(certs) if c != 0]
        assert not bad, f"certificate fails for sources {bad[:8]} (codes {[certs[i] for i in bad[:8]]})"
        pick = list(range(0, len(srcs), len(srcs) // 16))
        refs = list(ex.map(lambda i: oracle.dijkstra(g.row_ptr, g.col, g.w, srcs[i], True), pick))
    for i, r in zip(pick, refs):
        assert np.array_equal(d[i], r), f"source {i} differs from Dijkstra"


def test_c4b_all_64_graphs():
    gs, srcs = workloads.c4b()
    assert len(gs) == 64
    dgs = [adds_amd.Graph(h) for h in gs]
    dists, st = adds_amd.sssp_batched(dgs, srcs, timeout_ms=TIMEOUT)
    print(f"c4b: total n={sum(h.n for h in gs)}, kernel {st['kernel_ms']:.2f} ms", flush=True)
    for d in dgs:
        d.close()
    with ThreadPoolExecutor(POOL) as ex:
        refs = list(ex.map(lambda i: oracle.dijkstra(gs[i].row_ptr, gs[i].col, gs[i].w, srcs[i], False),
                           range(len(gs))))
    bad = [i for i in range(len(gs)) if not np.array_equal(dists[i].view(np.uint32), refs[i])]
    assert not bad, f"graphs {bad} differ from Dijkstra"


def test_c5_backward_against_finite_differences():
    fx = np.load(os.path.join(HERE, "golden", "c5_fd.npz"))
    g, srcs, gd = workloads.c5()
    assert zlib.crc32(g.w_bits.tobytes()) == int(fx["w_crc32"][0]), "C5 generator changed: rerun make_c5_fd.py"
    k = len(srcs)
    dg = adds_amd.Graph(g)
    dist = torch.empty((k, g.n), dtype=torch.float32, device="cuda")
    dg.sssp_multi_device(srcs, dist.data_ptr(), timeout_ms=TIMEOUT)
    gdt = torch.from_numpy(gd).cuda()
    gw = torch.empty(g.m, dtype=torch.float32, device="cuda")
    src_idx, edges, fd = fx["source"], fx["edge"], fx["fd"]
    assert len(fd) == 1000
    worst = 0.0
    for s in range(k):
        sel = np.nonzero(src_idx == s)[0]
        if sel.size == 0:
            continue
        # the gradient of L_s alone: the backward pass of source s
        dg.backward_device([srcs[s]], dist[s].data_ptr(), gdt[s].data_ptr(), gw.data_ptr())
        got = gw.cpu().numpy()[edges[sel]].astype(np.float64)
        want = fd[sel]
        err = np.abs(got - want) / np.maximum(np.abs(want), 1e-3)
        worst = max(worst, float(err.max()))
        bad = np.nonzero(err > 1e-3)[0]
        assert bad.size == 0, (f"source {s}: {bad.size} edges off, e.g. edge {edges[sel][bad[:3]]}: "
                               f"{got[bad[:3]]} vs fd {want[bad[:3]]}")
    print(f"c5 fd: 1000 edges, worst relative error {worst:.2e}", flush=True)
    dg.close()
