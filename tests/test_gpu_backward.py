"""Differentiable SSSP on the GPU: the hand-written backward kernel equals the CPU tree-gradient
reference bit for bit, agrees with central finite differences of the float64 shortest-path function on
tree edges whose tight-predecessor gap exceeds 1e-2, and works through torch.autograd."""
import numpy as np
import pytest
import torch

import adds_amd
import oracle
from adds_amd.autograd import sssp as sssp_ad

pytestmark = pytest.mark.gpu


def make_case(side, k, seed):
    host = adds_amd.HostGraph.grid(side, side, 0, seed=seed, wkind="lognormal", lo=0.5, hi=0.0)
    srcs = host.sample_component(0, k, seed=seed + 1)
    rng = np.random.default_rng(seed)
    gd = np.zeros((k, host.n), dtype=np.float32)
    ntg = min(4096, host.n)
    for s in range(k):
        tg = rng.choice(host.n, ntg, replace=False)
        gd[s, tg] = rng.integers(-(1 << 23), 1 << 23, ntg).astype(np.float32) / np.float32(1 << 23)
    return host, srcs, gd


@pytest.mark.parametrize("side,k", [(64, 4), (1024, 32)])
def test_backward_bit_exact_vs_cpu_tree(side, k):
    host, srcs, gd = make_case(side, k, seed=side)
    dg = adds_amd.Graph(host)
    dist = torch.empty((k, host.n), dtype=torch.float32, device="cuda")
    dg.sssp_multi_device(srcs, dist.data_ptr())
    d = dist.cpu().numpy()
    for i in (0, k - 1):
        assert np.array_equal(d[i].view(np.uint32), oracle.dijkstra(host.row_ptr, host.col, host.w, srcs[i], True))
    gdt = torch.from_numpy(gd).cuda()
    gw = torch.empty(host.m, dtype=torch.float32, device="cuda")
    ms = dg.backward_device(k, dist.data_ptr(), gdt.data_ptr(), gw.data_ptr())
    ref, _ = oracle.tree_grad(host.row_ptr, host.col, host.w, d, gd)
    got = gw.cpu().numpy()
    bad = np.nonzero(got.view(np.uint32) != ref.view(np.uint32))[0]
    assert bad.size == 0, f"{bad.size} differing gradients, e.g. {bad[:4]} {got[bad[:4]]} vs {ref[bad[:4]]}"
    print(f"backward side={side} k={k}: {ms:.3f} ms")
    # repeated backward is deterministic
    gw2 = torch.empty_like(gw)
    dg.backward_device(k, dist.data_ptr(), gdt.data_ptr(), gw2.data_ptr())
    assert torch.equal(gw.view(torch.int32), gw2.view(torch.int32))


def test_backward_finite_differences():
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    host, srcs, gd = make_case(48, 2, seed=7)
    dg = adds_amd.Graph(host)
    dist = torch.empty((2, host.n), dtype=torch.float32, device="cuda")
    dg.sssp_multi_device(srcs, dist.data_ptr())
    d = dist.cpu().numpy()
    gw = torch.empty(host.m, dtype=torch.float32, device="cuda")
    # per-source gradients: backward on one source at a time
    s = 0
    g1 = torch.from_numpy(gd[s:s + 1].copy()).cuda()
    d1 = dist[s:s + 1].contiguous()
    dg.backward_device(1, d1.data_ptr(), g1.data_ptr(), gw.data_ptr())
    grad = gw.cpu().numpy()
    _, pe = oracle.tree_grad(host.row_ptr, host.col, host.w, d[s:s + 1], gd[s:s + 1])
    src_of = np.repeat(np.arange(host.n), np.diff(host.row_ptr.astype(np.int64)))
    # tight-predecessor gap of every tree edge
    w64 = host.w.astype(np.float64)
    slack = d[s][src_of].astype(np.float64) + w64 - d[s][host.col].astype(np.float64)
    n = host.n
    h = 1e-3
    rng = np.random.default_rng(3)
    verts = [v for v in rng.permutation(n) if pe[0, v] != 0xFFFFFFFF]
    parent = np.full(n, -1)
    has = pe[0] != 0xFFFFFFFF
    parent[has] = src_of[pe[0][has]]
    children = [[] for _ in range(n)]
    for x in np.nonzero(has)[0]:
        children[parent[x]].append(x)
    tree_edge = np.zeros(host.m, dtype=bool)
    tree_edge[pe[0][has]] = True
    checked = 0
    for v in verts:
        e = int(pe[0, v])
        ins = np.nonzero(host.col == v)[0]
        alts = slack[ins[ins != e]]
        if alts.size and alts.min() <= 1e-2:  # tight-predecessor gap of the edge itself
            continue
        # The whole subtree below e moves by +-h: the tree is stable only if every non-tree edge that
        # crosses the subtree boundary has slack above the step as well (otherwise the function has a
        # kink inside [w-h, w+h] and central differences do not measure the derivative).
        sub = np.zeros(n, dtype=bool)
        stack = [v]
        while stack:
            x = stack.pop()
            sub[x] = True
            stack.extend(children[x])
        cross = (sub[src_of] != sub[host.col]) & ~tree_edge
        if cross.any() and slack[cross].min() <= 1e-2:
            continue

        def loss(wv):
            M = sp.csr_matrix((wv, host.col, host.row_ptr), shape=(n, n))
            dd = cg.dijkstra(M, indices=int(srcs[s]))
            fin = np.isfinite(dd)
            return float(np.sum(gd[s][fin].astype(np.float64) * dd[fin]))

        wp, wm = w64.copy(), w64.copy()
        wp[e] += h
        wm[e] -= h
        fd = (loss(wp) - loss(wm)) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-3 * max(1.0, abs(fd)), (v, e, fd, grad[e])
        checked += 1
        if checked == 40:
            break
    assert checked >= 20


def test_autograd_function():
    host, srcs, gd = make_case(96, 3, seed=11)
    dg = adds_amd.Graph(host)
    w = torch.from_numpy(host.w.copy()).cuda().requires_grad_(True)
    c = torch.from_numpy(gd).cuda()
    dist = sssp_ad(w, dg, srcs)
    fin = torch.isfinite(dist)
    loss = (torch.where(fin, dist, torch.zeros_like(dist)) * c).sum()
    loss.backward()
    ref, _ = oracle.tree_grad(host.row_ptr, host.col, host.w, dist.detach().cpu().numpy(), gd)
    assert np.array_equal(w.grad.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    # a second forward with other weights, then backward of the first graph again, must not mix them
    w2 = (w.detach() * 1.25).requires_grad_(True)
    dist2 = sssp_ad(w2, dg, srcs)
    assert not torch.equal(dist2, dist.detach())
