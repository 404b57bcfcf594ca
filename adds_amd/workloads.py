"""The benchmark workloads of this path, generated from seeds (no dataset): one function per configuration,
shared by bench.py and the full-size parity tests so that both run on identical inputs.

    c1()        RMAT scale 16, edge factor 16, uint32 weights in [1, 100]; source = highest-degree vertex
    c2()        2048 x 2048 4-neighbour grid + 8385 shortcut pairs (0.1% of the undirected grid edges),
                float32 weights in [1, 2); source = vertex 0 (a corner)
    c3(kind)    RMAT scale 24, edge factor 32, uint32 weights in [1, 65535] ("u32") or float32 weights in
                [0, 1) ("f32"); source = highest-degree vertex
    c4a()       RMAT scale 20, edge factor 16, float32 weights in [1, 10); 256 sources sampled from the
                component of the highest-degree vertex
    c4b()       64 graphs, sizes log-uniform in [1e4, 1e6], alternating 2-D grids and Erdos-Renyi graphs of
                mean degree 8, uint32 weights in [1, 1000]; source 0 of each
    c5()        1024 x 1024 grid, float32 weights exp(N(0, 0.5^2)); 32 sources; per source 4096 target
                vertices with loss coefficients c = k / 2^23, k uniform in [-2^23, 2^23)

RMAT uses a/b/c/d = .57/.19/.19/.05; every graph is symmetric with one weight per undirected pair, no
self loops and no duplicate edges (adds_gen_*).
"""
import numpy as np

from .graph import HostGraph


def c1():
    g = HostGraph.rmat(16, 16, seed=1, wkind="u32", lo=1, hi=100)
    return g, g.max_degree_vertex()


def c2():
    return HostGraph.grid(2048, 2048, 8385, seed=2, wkind="f32", lo=1.0, hi=2.0), 0


def c3(kind="u32", scale=24):
    g = HostGraph.rmat(scale, 32, seed=1, wkind="u32", lo=1, hi=65535)
    if kind == "f32":
        g.reweight(seed=1, wkind="f32", lo=0.0, hi=1.0)
    return g, g.max_degree_vertex()


def c3_reweight_float(g):
    """The float variant of c3's graph in place (same edges, float32 weights in [0, 1))."""
    return g.reweight(seed=1, wkind="f32", lo=0.0, hi=1.0)


def c4a(k=256):
    g = HostGraph.rmat(20, 16, seed=4, wkind="f32", lo=1.0, hi=10.0)
    return g, g.sample_component(g.max_degree_vertex(), k, seed=40)


def c4b(count=64):
    rng = np.random.default_rng(4242)  # sizes only; graph content comes from the library's seeded PRNG
    gs = []
    for i in range(count):
        n = int(round(np.exp(np.log(1e4) + rng.random() * (np.log(1e6) - np.log(1e4)))))
        if i % 2 == 0:
            side = max(2, int(round(np.sqrt(n))))
            gs.append(HostGraph.grid(side, side, 0, seed=100 + i, wkind="u32", lo=1, hi=1000))
        else:
            gs.append(HostGraph.er(n, 8.0, seed=100 + i, wkind="u32", lo=1, hi=1000))
    return gs, [0] * len(gs)


def c5(side=1024, k=32, targets=4096):
    """(graph, sources, grad_dist [k][n] float32 with `targets` nonzero coefficients per source)."""
    g = HostGraph.grid(side, side, 0, seed=5, wkind="lognormal", lo=0.5, hi=0.0)
    srcs = g.sample_component(0, k, seed=55)
    rng = np.random.default_rng(5)
    gd = np.zeros((len(srcs), g.n), dtype=np.float32)
    for s in range(len(srcs)):
        tg = rng.choice(g.n, targets, replace=False)
        gd[s, tg] = rng.integers(-(1 << 23), 1 << 23, size=targets).astype(np.float32) / np.float32(1 << 23)
    return g, srcs, gd
