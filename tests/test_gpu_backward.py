"""Differentiable SSSP on the GPU: the hand-written backward pass equals the CPU reference gradient
(oracle.tree_grad: same tree rule, exact fixed-point sums rounded once) bit for bit, in its sparse form
(one walk per nonzero dL/dd) and its dense form (one climb per leaf); zero-weight edges are tree edges;
bad inputs are refused; and the torch.autograd binding gives the same gradient.  The full-size C5
finite-difference check is in test_gpu_fullsize.py."""
import numpy as np
import pytest
import torch

import adds_amd
import oracle
from adds_amd import _lib as L
from adds_amd.autograd import sssp as sssp_ad

pytestmark = pytest.mark.gpu


def make_case(side, k, seed, targets=4096):
    host = adds_amd.HostGraph.grid(side, side, 0, seed=seed, wkind="lognormal", lo=0.5, hi=0.0)
    srcs = host.sample_component(0, k, seed=seed + 1)
    rng = np.random.default_rng(seed)
    gd = np.zeros((k, host.n), dtype=np.float32)
    ntg = min(targets, host.n)
    for s in range(k):
        tg = rng.choice(host.n, ntg, replace=False)
        gd[s, tg] = rng.integers(-(1 << 23), 1 << 23, ntg).astype(np.float32) / np.float32(1 << 23)
    return host, srcs, gd


def device_grad(dg, srcs, dist, gd):
    gdt = torch.from_numpy(np.ascontiguousarray(gd)).cuda()
    gw = torch.full((dg.m,), 7.0, dtype=torch.float32, device="cuda")
    ms = dg.backward_device(srcs, dist.data_ptr(), gdt.data_ptr(), gw.data_ptr())
    return gw.cpu().numpy(), ms


def assert_bits_equal(got, ref):
    bad = np.nonzero(got.view(np.uint32) != ref.view(np.uint32))[0]
    assert bad.size == 0, f"{bad.size} differing gradients, e.g. {bad[:4]} {got[bad[:4]]} vs {ref[bad[:4]]}"


@pytest.mark.parametrize("side,k,targets", [(64, 4, 4096), (64, 4, 64), (256, 8, 1 << 30), (1024, 32, 4096)])
def test_backward_bit_exact_vs_cpu_reference(side, k, targets):
    """Sparse targets take the walk form, all-vertex targets (targets >= n) the climb form."""
    host, srcs, gd = make_case(side, k, seed=side + targets % 97, targets=targets)
    dg = adds_amd.Graph(host)
    dist = torch.empty((k, host.n), dtype=torch.float32, device="cuda")
    dg.sssp_multi_device(srcs, dist.data_ptr())
    d = dist.cpu().numpy()
    for i in (0, k - 1):
        assert np.array_equal(d[i].view(np.uint32), oracle.dijkstra(host.row_ptr, host.col, host.w, srcs[i], True))
    got, ms = device_grad(dg, srcs, dist, gd)
    ref, _, scale = oracle.tree_grad(host.row_ptr, host.col, host.w, srcs, d, gd)
    assert_bits_equal(got, ref)
    print(f"backward side={side} k={k} targets={min(targets, host.n)}: {ms:.3f} ms, scale {scale}")
    # repeated backward is bit-identical (no order dependence)
    got2, _ = device_grad(dg, srcs, dist, gd)
    assert_bits_equal(got2, got)


def test_backward_zero_weights_and_plateaus():
    """A grid where a third of the weights are zero: many vertices are reached only through zero-weight
    (equal-distance) edges and get their parents in later rounds; sparse and dense forms."""
    host = adds_amd.HostGraph.grid(48, 40, 30, seed=4, wkind="f32", lo=0.0, hi=3.0)
    rng = np.random.default_rng(8)
    w = host.w.copy()
    w[rng.random(host.m) < 0.33] = 0.0
    g = adds_amd.HostGraph.from_arrays(host.row_ptr, host.col, w, "f32")
    srcs = np.array([0, 777, 1500], dtype=np.uint32)
    dg = adds_amd.Graph(g)
    dist = torch.empty((3, g.n), dtype=torch.float32, device="cuda")
    dg.sssp_multi_device(srcs, dist.data_ptr())
    d = dist.cpu().numpy()
    for density in (0.01, 1.0):
        gd = (rng.integers(-(1 << 23), 1 << 23, d.shape) / float(1 << 23)).astype(np.float32)
        gd[rng.random(d.shape) >= density] = 0
        got, _ = device_grad(dg, srcs, dist, gd)
        ref, pe, _ = oracle.tree_grad(g.row_ptr, g.col, g.w, srcs, d, gd)
        assert_bits_equal(got, ref)
    # some tree edges really are zero-weight edges between equal distances
    tails = np.repeat(np.arange(g.n), np.diff(g.row_ptr.astype(np.int64)))
    e = pe[0][pe[0] != 0xFFFFFFFF]
    assert np.any((w[e] == 0) & (d[0][tails[e]] == d[0][g.col[e]]))


def test_backward_zero_weight_chain_closed_form():
    """0 -0-> 1 -0-> 2 -1.5-> 3 with a zero-weight back edge 2 -> 1: dL/dd[3] = 1, dL/dd[2] = 0.25 give
    dL/dw = (1.25, 1.25, 0, 1)."""
    g = adds_amd.HostGraph.from_arrays(np.array([0, 1, 2, 4, 4]), np.array([1, 2, 1, 3]),
                                       np.array([0.0, 0.0, 0.0, 1.5]), "f32")
    dg = adds_amd.Graph(g)
    dist = torch.empty((1, 4), dtype=torch.float32, device="cuda")
    dg.sssp_multi_device([0], dist.data_ptr())
    got, _ = device_grad(dg, [0], dist, np.array([[0, 0, 0.25, 1.0]], dtype=np.float32))
    assert list(got) == [1.25, 1.25, 0.0, 1.0]


def test_backward_refuses_bad_inputs():
    host, srcs, gd = make_case(32, 2, seed=3, targets=100)
    dg = adds_amd.Graph(host)
    dist = torch.empty((2, host.n), dtype=torch.float32, device="cuda")
    dg.sssp_multi_device(srcs, dist.data_ptr())
    bad = gd.copy()
    bad[1, int(srcs[1]) ^ 1] = np.inf
    with pytest.raises(adds_amd.AddsError) as ei:
        device_grad(dg, srcs, dist, bad)
    assert ei.value.status == L.EINVAL
    wrong = dist.clone()
    wrong[0, (int(srcs[0]) + 5) % host.n] -= 0.5  # no tight path reaches this vertex any more
    with pytest.raises(adds_amd.AddsError) as ei:
        device_grad(dg, srcs, wrong, gd)
    assert ei.value.status == L.EINVAL
    got, _ = device_grad(dg, srcs, dist, gd)  # still fine afterwards
    assert_bits_equal(got, oracle.tree_grad(host.row_ptr, host.col, host.w, srcs, dist.cpu().numpy(), gd)[0])


def test_autograd_function():
    host, srcs, gd = make_case(96, 3, seed=11)
    dg = adds_amd.Graph(host)
    w = torch.from_numpy(host.w.copy()).cuda().requires_grad_(True)
    c = torch.from_numpy(gd).cuda()
    dist = sssp_ad(w, dg, srcs)
    fin = torch.isfinite(dist)
    loss = (torch.where(fin, dist, torch.zeros_like(dist)) * c).sum()
    loss.backward()
    ref, _, _ = oracle.tree_grad(host.row_ptr, host.col, host.w, srcs, dist.detach().cpu().numpy(), gd)
    assert np.array_equal(w.grad.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    # a second forward with other weights, then backward of the first graph again, must not mix them
    w2 = (w.detach() * 1.25).requires_grad_(True)
    dist2 = sssp_ad(w2, dg, srcs)
    assert not torch.equal(dist2, dist.detach())
