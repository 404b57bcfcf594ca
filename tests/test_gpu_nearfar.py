"""The near-far GPU baseline (adds_sssp_nearfar, the reference's NF comparison runs): distances equal CPU
Dijkstra bit for bit over weight types, graph shapes and threshold widths, through the C ABI."""
import numpy as np
import pytest

import adds_amd
import oracle

pytestmark = pytest.mark.gpu


def _check(g, dist, src, is_float):
    ref = oracle.dijkstra(g.row_ptr, g.col, g.w, src, is_float)
    got = np.ascontiguousarray(dist).view(np.uint32)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: got {got[bad[:5]]} want {ref[bad[:5]]}"


@pytest.mark.parametrize("delta", [0.0, 1.0, 37.0, 1e9])  # default, many phases, some, one (label-correcting)
def test_rmat16_int(delta):
    g = adds_amd.HostGraph.rmat(16, 16, seed=1, wkind="u32", lo=1, hi=100)
    src = g.max_degree_vertex()
    dist, st = adds_amd.Graph(g).sssp_nearfar(src, delta=delta)
    print("nf rmat16 int", delta, st["probes"], st["rebuckets"], st["vertices_expanded"])
    _check(g, dist, src, False)
    nr, _ = g.reached_counts(dist)
    assert st["vertices_expanded"] >= nr and st["edges_scanned"] > 0


@pytest.mark.parametrize("delta", [0.0, 0.01, 1e30])
def test_rmat_float(delta):
    g = adds_amd.HostGraph.rmat(15, 16, seed=2, wkind="f32", lo=0.0, hi=1.0)
    src = g.max_degree_vertex()
    dist, _ = adds_amd.Graph(g).sssp_nearfar(src, delta=delta)
    _check(g, dist, src, True)


def test_grid_float_small_and_hub_rows():
    """A grid (many phases at the default delta) and an RMAT graph whose hub rows exceed the 1024-edge task
    size, so both relax paths run."""
    g = adds_amd.HostGraph.grid(256, 256, 64, seed=5, wkind="f32", lo=1.0, hi=2.0)
    dist, st = adds_amd.Graph(g).sssp_nearfar(0)
    assert st["rebuckets"] > 10
    _check(g, dist, 0, True)
    h = adds_amd.HostGraph.rmat(18, 16, seed=3, wkind="u32", lo=1, hi=1000)
    assert h.degrees().max() > 4 * 1024
    src = h.max_degree_vertex()
    dist, _ = adds_amd.Graph(h).sssp_nearfar(src)
    _check(h, dist, src, False)


def test_disconnected_and_er():
    g = adds_amd.HostGraph.er(20000, 3.0, seed=7, wkind="u32", lo=1, hi=1000)  # several components
    dist, _ = adds_amd.Graph(g).sssp_nearfar(5)
    _check(g, dist, 5, False)
    assert (dist == 0xFFFFFFFF).any()


def test_grid2048_float_c2():
    """The C2 graph (2048x2048 grid + 8385 shortcut pairs, float weights in [1,2), source 0)."""
    g = adds_amd.HostGraph.grid(2048, 2048, 8385, seed=2, wkind="f32", lo=1.0, hi=2.0)
    dist, st = adds_amd.Graph(g).sssp_nearfar(0)
    print("nf c2", st["ms"], st["probes"], st["rebuckets"])
    _check(g, dist, 0, True)


def test_same_as_adds_device_output():
    """Into a device buffer, and equal to the ADDS solve of the same graph."""
    import torch
    g = adds_amd.HostGraph.rmat(16, 8, seed=9, wkind="u32", lo=1, hi=65535)
    dg = adds_amd.Graph(g)
    buf = torch.empty(g.n, dtype=torch.int32, device="cuda")
    dg.sssp_nearfar_device(3, buf.data_ptr())
    torch.cuda.synchronize()
    ref, _ = dg.sssp(3)
    assert np.array_equal(buf.cpu().numpy().view(np.uint32), np.ascontiguousarray(ref).view(np.uint32))
