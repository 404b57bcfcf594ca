"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

The reference ships no golden vectors for this path (its only correctness mechanism is text equality
against Galois Dijkstra, verify.py), so the oracle is pinned by: scipy's Dijkstra (float64, exact for
integer sums < 2^53), brute-force Bellman-Ford on tiny graphs (bit-equal, int and float32: with
non-negative weights and a monotone add every relaxation order reaches the same fixed point), closed
forms on grids/paths/stars, and planted mistakes that the certificate must reject.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

import oracle

INF_U32 = 0xFFFFFFFF
INF_F32 = 0x7F800000


def rand_graph(rng, n, m, wmax=100, is_float=False, zero_ok=False):
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    pairs = np.unique(src[keep].astype(np.int64) * n + dst[keep])  # sorted, no duplicate edges
    src, dst = pairs // n, pairs % n
    row = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row).astype(np.uint32)
    if is_float:
        lo = 0.0 if zero_ok else 0.05
        w = rng.uniform(lo, 3.0, len(src)).astype(np.float32)
    else:
        w = rng.integers(0 if zero_ok else 1, wmax + 1, len(src)).astype(np.uint32)
    return row, dst.astype(np.uint32), w


def scipy_dist(row, col, w, src):
    n = len(row) - 1
    M = sp.csr_matrix((w.astype(np.float64), col, row), shape=(n, n))
    d = cg.dijkstra(M, indices=src)
    return d


@pytest.mark.parametrize("seed", range(6))
def test_dijkstra_int_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    row, col, w = rand_graph(rng, 300, 1500)
    src = int(rng.integers(0, 300))
    d = oracle.dijkstra(row, col, w, src, False)
    ref = scipy_dist(row, col, w, src)
    exp = np.where(np.isinf(ref), INF_U32, ref).astype(np.uint64)
    assert np.array_equal(d.astype(np.uint64), exp)


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("is_float", [False, True])
def test_dijkstra_equals_bellman_ford(seed, is_float):
    rng = np.random.default_rng(100 + seed)
    row, col, w = rand_graph(rng, 60, 240, is_float=is_float, zero_ok=True)
    srcs = rng.integers(0, 60, 1 + seed % 3)
    d = oracle.dijkstra(row, col, w, srcs, is_float)
    b = oracle.bellman_ford(row, col, w, srcs, is_float)
    assert np.array_equal(d, b)
    assert oracle.certify(row, col, w, srcs, d, is_float) == 0


def test_float_dijkstra_close_to_float64():
    rng = np.random.default_rng(7)
    row, col, w = rand_graph(rng, 500, 3000, is_float=True)
    d = oracle.dijkstra(row, col, w, 0, True).view(np.float32)
    ref = scipy_dist(row, col, w, 0)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(d), fin)
    assert np.allclose(d[fin], ref[fin], rtol=1e-5, atol=1e-5)


def grid(rows, cols, wfun):
    n = rows * cols
    adj = [[] for _ in range(n)]
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            for dr, dc in ((0, 1), (1, 0), (0, -1), (-1, 0)):
                rr, cc = r + dr, c + dc
                if 0 <= rr < rows and 0 <= cc < cols:
                    adj[v].append(rr * cols + cc)
    row = np.zeros(n + 1, dtype=np.uint32)
    for v in range(n):
        row[v + 1] = row[v] + len(adj[v])
    col = np.array([u for a in adj for u in sorted(a)], dtype=np.uint32)
    w = np.array([wfun() for _ in col])
    return row, col, w


def test_unit_grid_is_manhattan():
    row, col, w = grid(17, 23, lambda: 1)
    d = oracle.dijkstra(row, col, w.astype(np.uint32), 0, False)
    rr, cc = np.divmod(np.arange(17 * 23), 23)
    assert np.array_equal(d, (rr + cc).astype(np.uint32))


def test_dyadic_float_grid_is_exact_multiple():
    row, col, w = grid(20, 20, lambda: 1.5)
    d = oracle.dijkstra(row, col, w.astype(np.float32), 0, True).view(np.float32)
    rr, cc = np.divmod(np.arange(400), 20)
    assert np.array_equal(d, (1.5 * (rr + cc)).astype(np.float32))


def test_path_star_disconnected():
    # path 0-1-2-3 (directed forward, weights 2,3,4), star centre 4 -> 5,6,7 (weight 7), vertex 8 isolated
    row = np.array([0, 1, 2, 3, 3, 6, 6, 6, 6, 6], dtype=np.uint32)
    col = np.array([1, 2, 3, 5, 6, 7], dtype=np.uint32)
    w = np.array([2, 3, 4, 7, 7, 7], dtype=np.uint32)
    d = oracle.dijkstra(row, col, w, 0, False)
    assert list(d) == [0, 2, 5, 9, INF_U32, INF_U32, INF_U32, INF_U32, INF_U32]
    d = oracle.dijkstra(row, col, w, 4, False)
    assert list(d) == [INF_U32] * 4 + [0, 7, 7, 7, INF_U32]


@pytest.mark.parametrize("is_float", [False, True])
def test_certificate_rejects_planted_mistakes(is_float):
    rng = np.random.default_rng(3)
    row, col, w = rand_graph(rng, 400, 2400, is_float=is_float)
    d = oracle.dijkstra(row, col, w, 0, is_float)
    assert oracle.certify(row, col, w, 0, d, is_float) == 0
    reach = np.nonzero((d != (INF_F32 if is_float else INF_U32)) & (np.arange(400) != 0))[0]
    unreach = np.nonzero(d == (INF_F32 if is_float else INF_U32))[0]
    assert len(reach) > 10
    v = reach[len(reach) // 2]
    bad = d.copy(); bad[v] -= 1          # one unit (int) / one ulp (float) too small
    assert oracle.certify(row, col, w, 0, bad, is_float) != 0
    bad = d.copy(); bad[v] += 1          # too large
    assert oracle.certify(row, col, w, 0, bad, is_float) != 0
    bad = d.copy(); bad[v] = INF_F32 if is_float else INF_U32
    assert oracle.certify(row, col, w, 0, bad, is_float) != 0
    bad = d.copy(); bad[0] = 1
    assert oracle.certify(row, col, w, 0, bad, is_float) == 1
    if len(unreach):
        bad = d.copy(); bad[unreach[0]] = 5
        assert oracle.certify(row, col, w, 0, bad, is_float) == 3


def test_int_overflow_is_reported():
    row = np.array([0, 1, 2, 2], dtype=np.uint32)
    col = np.array([1, 2], dtype=np.uint32)
    w = np.array([0xF0000000, 0x20000000], dtype=np.uint32)
    with pytest.raises(OverflowError):
        oracle.dijkstra(row, col, w, 0, False)


def saturation_case():
    """A path 0 -> 1 -> 2 -> 3 whose sums pass 2^32 (0xF0000000 + 0x20000000), a branch 1 -> 4 that ends
    exactly at 2^32 - 3 (the largest exact sum), and a detour 0 -> 5 -> 2 that is shorter than the
    saturated direct route only by a wrapped sum: the saturated solution is fixed by hand."""
    # edges in CSR order: 0->1, 0->5, 1->2, 1->4, 2->3, 5->2
    row = np.array([0, 2, 4, 5, 5, 5, 6], dtype=np.uint32)
    col = np.array([1, 5, 2, 4, 3, 2], dtype=np.uint32)
    w = np.array([0xF0000000, 0xFFFFFFF0, 0x20000000, 0x0FFFFFFD, 7, 0x20], dtype=np.uint32)
    # d1 = 0xF0000000; d4 = 0xF0000000 + 0x0FFFFFFD = 0xFFFFFFFD (exact); d5 = 0xFFFFFFF0;
    # d2 = min(sat(0xF0000000 + 0x20000000), sat(0xFFFFFFF0 + 0x20)) = 0xFFFFFFFE; d3 = sat(d2 + 7)
    want = np.array([0, 0xF0000000, 0xFFFFFFFE, 0xFFFFFFFE, 0xFFFFFFFD, 0xFFFFFFF0], dtype=np.uint32)
    return row, col, w, want


def test_int_saturation_closed_form():
    """With saturate=True the oracle returns the saturating solution (sums >= 2^32-1 become 2^32-2),
    the contract the device solvers follow; Dijkstra and Bellman-Ford agree on it."""
    row, col, w, want = saturation_case()
    assert np.array_equal(oracle.dijkstra(row, col, w, 0, False, saturate=True), want)
    assert np.array_equal(oracle.bellman_ford(row, col, w, 0, False, saturate=True), want)
    with pytest.raises(OverflowError):
        oracle.bellman_ford(row, col, w, 0, False)
    # a sum of exactly 2^32 - 3 does not count as overflow
    d = oracle.dijkstra(np.array([0, 1, 2, 2], dtype=np.uint32), np.array([1, 2], dtype=np.uint32),
                        np.array([0xF0000000, 0x0FFFFFFD], dtype=np.uint32), 0, False)
    assert np.array_equal(d, [0, 0xF0000000, 0xFFFFFFFD])


# ---- backward reference ----
def exact_tree_grad(row, col, w, srcs, dist, gdist):
    """Independent evaluation for dyadic gradients (multiples of 2^-23): the tree by brute force over the
    edge list (round 0: smallest tight in-edge from a strictly closer vertex; then rounds over
    equal-distance tight in-edges from vertices resolved earlier), the gradient as exact Python-integer
    sums of c * 2^23 over each edge's tree descendants, rounded once to float32."""
    k, n = dist.shape
    m = len(col)
    tail = np.repeat(np.arange(n), np.diff(row.astype(np.int64)))
    total = [0] * m
    for s in range(k):
        d = dist[s]
        pe = [None] * n
        rnd = [None] * n
        for v in range(n):
            if not np.isfinite(d[v]):
                continue
            if v == srcs[s]:
                rnd[v] = 0
                continue
            cands = [e for e in range(m) if col[e] == v and d[tail[e]] < d[v]
                     and np.float32(d[tail[e]] + w[e]) == d[v]]
            if cands:
                pe[v], rnd[v] = min(cands), 0
        r = 1
        while True:
            new = {}
            for v in range(n):
                if not np.isfinite(d[v]) or rnd[v] is not None:
                    continue
                cands = [e for e in range(m) if col[e] == v and d[tail[e]] == d[v]
                         and np.float32(d[tail[e]] + w[e]) == d[v] and rnd[tail[e]] is not None]
                if cands:
                    new[v] = min(cands)
            if not new:
                break
            for v, e in new.items():
                pe[v], rnd[v] = e, r
            r += 1
        for t in range(n):
            if gdist[s, t] == 0 or not np.isfinite(d[t]):
                continue
            c = int(gdist[s, t] * (1 << 23))
            assert c == gdist[s, t] * (1 << 23)
            x = t
            while pe[x] is not None:
                total[pe[x]] += c
                x = tail[pe[x]]
    return np.array([np.float32(t / float(1 << 23)) for t in total], dtype=np.float32)


def dyadic(rng, shape, density=1.0):
    g = (rng.integers(-(1 << 23), 1 << 23, shape) / float(1 << 23)).astype(np.float32)
    return np.where(rng.random(shape) < density, g, 0).astype(np.float32)


@pytest.mark.parametrize("seed,zero_ok", [(0, False), (1, False), (2, True), (3, True)])
def test_tree_grad_matches_exact_path_sums(seed, zero_ok):
    """Bit-exact against the independent evaluation, with zero weights allowed in half the cases (so
    equal-distance tree edges occur); a wrong parent, a dropped edge or a wrong sign changes bits."""
    rng = np.random.default_rng(50 + seed)
    row, col, w = rand_graph(rng, 40, 200, is_float=True, zero_ok=zero_ok)
    if zero_ok:
        w[rng.random(len(w)) < 0.3] = 0.0
    srcs = [0, 7, 13]
    dist = np.stack([oracle.dijkstra(row, col, w, s, True).view(np.float32) for s in srcs])
    gd = dyadic(rng, dist.shape)
    gw, pe, scale = oracle.tree_grad(row, col, w, srcs, dist, gd)
    assert scale >= 23
    ref = exact_tree_grad(row, col, w, srcs, dist, gd)
    assert np.array_equal(gw.view(np.uint32), ref.view(np.uint32))
    assert np.count_nonzero(gw) > 0
    # planted mistakes must be visible: another sign or another coefficient
    gw2, _, _ = oracle.tree_grad(row, col, w, srcs, dist, -gd)
    assert np.array_equal(gw2, -gw)
    gd3 = gd.copy()
    t = int(np.nonzero(np.isfinite(dist[0]) & (np.arange(40) != srcs[0]))[0][0])
    gd3[0, t] += np.float32(0.5)
    assert not np.array_equal(oracle.tree_grad(row, col, w, srcs, dist, gd3)[0], gw)


def test_tree_grad_zero_weight_chain_and_cycle():
    """Zero-weight edges are tree edges: source 0 -0-> 1 -0-> 2 -1.5-> 3, plus a zero-weight back edge
    2 -> 1 (a cycle at equal distance).  With dL/dd[3] = 1 and dL/dd[2] = 0.25, d[3] = w01 + w12 + w23
    gives dL/dw = 1.25 on 0->1 and 1->2, 1 on 2->3, and 0 on the back edge."""
    # edges (CSR order): 0->1, 1->2, 2->1, 2->3
    row = np.array([0, 1, 2, 4, 4], dtype=np.uint32)
    col = np.array([1, 2, 1, 3], dtype=np.uint32)
    w = np.array([0.0, 0.0, 0.0, 1.5], dtype=np.float32)
    d = oracle.dijkstra(row, col, w, 0, True).view(np.float32)
    assert list(d) == [0.0, 0.0, 0.0, 1.5]
    gd = np.array([[0.0, 0.0, 0.25, 1.0]], dtype=np.float32)
    gw, pe, _ = oracle.tree_grad(row, col, w, [0], d[None, :], gd)
    assert list(gw) == [1.25, 1.25, 0.0, 1.0]
    assert list(pe[0]) == [0xFFFFFFFF, 0, 1, 3]
    # an absorbed add: 2^25 + 1 rounds to 2^25 in float32, so the unit edge is a tree edge too
    row = np.array([0, 1, 2, 2], dtype=np.uint32)
    col = np.array([1, 2], dtype=np.uint32)
    w = np.array([2.0 ** 25, 1.0], dtype=np.float32)
    d = oracle.dijkstra(row, col, w, 0, True).view(np.float32)
    assert d[2] == d[1]
    gw, _, _ = oracle.tree_grad(row, col, w, [0], d[None, :], np.array([[0, 0, 1.0]], dtype=np.float32))
    assert list(gw) == [1.0, 1.0]


def test_tree_grad_scale_rule():
    """S = 62 - eM - eN (M = max |g| = m 2^eM with 0.5 <= m < 1; 2^eN >= nonzero count); unreached
    vertices do not count; inputs off the grid 2^-S are rounded to it."""
    dist = np.array([[0.0, 1.0, np.inf, 2.0]], dtype=np.float32)
    g = np.array([[0.5, -3.0, 1e30, 0.0]], dtype=np.float32)  # M = 3 = 0.75 * 2^2, N = 2 = 2^1
    assert oracle.grad_scale(dist, g) == (62 - 2 - 1, 2)
    assert oracle.grad_scale(dist, np.zeros_like(g))[1] == 0


def fd_loss(row, col, w64, src, gd):
    n = len(row) - 1
    M = sp.csr_matrix((w64, col, row), shape=(n, n))
    d = cg.dijkstra(M, indices=src)
    fin = np.isfinite(d)
    return float(np.sum(gd[fin] * d[fin]))


def test_tree_grad_central_differences():
    """dL/dw against central differences of the float64 shortest-path function (h = 1e-3) on tree
    edges whose tight-predecessor gap exceeds 1e-2."""
    rng = np.random.default_rng(9)
    row, col, w = rand_graph(rng, 80, 480, is_float=True)
    src = 3
    d = oracle.dijkstra(row, col, w, src, True).view(np.float32)
    gd = dyadic(rng, 80)
    gw, pe, _ = oracle.tree_grad(row, col, w, [src], d[None, :], gd[None, :])
    src_of = np.repeat(np.arange(80), np.diff(row.astype(np.int64)))
    h = 1e-3
    checked = 0
    for v in range(80):
        e = pe[0, v]
        if e == 0xFFFFFFFF:
            continue
        alts = [float(d[src_of[x]]) + float(w[x]) - float(d[v]) for x in range(len(col)) if col[x] == v and x != e]
        if alts and min(alts) <= 1e-2:
            continue
        w64 = w.astype(np.float64)
        wp, wm = w64.copy(), w64.copy()
        wp[e] += h
        wm[e] -= h
        fd = (fd_loss(row, col, wp, src, gd) - fd_loss(row, col, wm, src, gd)) / (2 * h)
        assert abs(fd - gw[e]) <= 1e-3 * max(1.0, abs(fd)), (v, e, fd, gw[e])
        checked += 1
    assert checked >= 10


def test_tree_grad_c5_fixture_sources():
    """The committed C5 finite differences (scipy, float64, full 1024^2 grid) against the oracle's
    gradient of single sources (four of the 32; the GPU test checks all 1000 edges)."""
    import os
    import zlib

    from adds_amd import workloads

    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c5_fd.npz"))
    g, srcs, gd = workloads.c5()
    assert zlib.crc32(g.w_bits.tobytes()) == int(fx["w_crc32"][0])
    for s in (0, 9, 22, 31):
        sel = np.nonzero(fx["source"] == s)[0]
        d = oracle.dijkstra(g.row_ptr, g.col, g.w, srcs[s], True).view(np.float32)
        gw, _, _ = oracle.tree_grad(g.row_ptr, g.col, g.w, [srcs[s]], d[None, :], gd[s][None, :])
        got = gw[fx["edge"][sel]].astype(np.float64)
        want = fx["fd"][sel]
        err = np.abs(got - want) / np.maximum(np.abs(want), 1e-3)
        assert err.max() <= 1e-3, (s, err.max())
