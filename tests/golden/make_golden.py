"""Writes the committed golden fixtures of tests/golden/.

Each fixture is a small .gr graph plus its expected per-vertex distance text (the reference's output
format, ads_int/support.cu:89-95 and cpu_sssp_float.patch:241-248).  Expected distances come from the
CPU oracle's Dijkstra and are cross-checked here against scipy's float64 Dijkstra (exact for the int
fixtures, within 1e-5 relative for the float ones) before anything is written.  gen_hashes.json pins
the seeded generators.  Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import struct
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import adds_amd  # noqa: E402
import oracle  # noqa: E402

INF_U32, INF_F32 = 0xFFFFFFFF, 0x7F800000


def dist_text(d, is_float):
    out = []
    for i, x in enumerate(d):
        if x == (INF_F32 if is_float else INF_U32):
            out.append(f"{i} INF\n")
        elif is_float:
            out.append(f"{i} {float(np.uint32(x).view(np.float32)):f}\n")
        else:
            out.append(f"{i} {int(x)}\n")
    return "".join(out)


def disconnected():
    # component A: 0-1-2 triangle (weights 4, 9, 3); component B: 3-4 (weight 6); 5 isolated
    edges = {(0, 1): 4, (1, 2): 9, (0, 2): 3, (3, 4): 6}
    adj = {v: [] for v in range(6)}
    for (a, b), w in edges.items():
        adj[a].append((b, w))
        adj[b].append((a, w))
    row, col, w = [0], [], []
    for v in range(6):
        for u, x in sorted(adj[v]):
            col.append(u)
            w.append(x)
        row.append(len(col))
    return adds_amd.HostGraph.from_arrays(row, col, w, "int")


def main():
    fixtures = {
        "rmat8_int": (adds_amd.HostGraph.rmat(8, 8, seed=21, wkind="u32", lo=1, hi=100), None),
        "grid12_float": (adds_amd.HostGraph.grid(12, 12, 4, seed=22, wkind="f32", lo=1.0, hi=2.0), 0),
        "er500_float": (adds_amd.HostGraph.er(500, 4, seed=23, wkind="f32", lo=0.0, hi=1.0), 0),
        "disconnected_int": (disconnected(), 1),
    }
    meta = {}
    for name, (g, src) in fixtures.items():
        if src is None:
            src = g.max_degree_vertex()
        d = oracle.dijkstra(g.row_ptr, g.col, g.w, src, g.is_float)
        M = sp.csr_matrix((g.w.astype(np.float64), g.col, g.row_ptr), shape=(g.n, g.n))
        ref = cg.dijkstra(M, indices=src)
        fin = np.isfinite(ref)
        dv = d.view(np.float32).astype(np.float64) if g.is_float else d.astype(np.float64)
        assert np.array_equal(fin, d != (INF_F32 if g.is_float else INF_U32)), name
        if g.is_float:
            assert np.allclose(dv[fin], ref[fin], rtol=1e-5), name
        else:
            assert np.array_equal(dv[fin], ref[fin]), name
        g.write_gr(os.path.join(HERE, name + ".gr"))
        with open(os.path.join(HERE, name + ".dist"), "w") as f:
            f.write(dist_text(d, g.is_float))
        meta[name] = {"wtype": "float" if g.is_float else "int", "src": int(src), "n": g.n, "m": g.m}
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)

    def h(g):
        x = hashlib.sha256()
        for a in (g.row_ptr, g.col, g.w_bits):
            x.update(np.ascontiguousarray(a).tobytes())
        return x.hexdigest()[:16]

    hashes = {
        "rmat10_ef8_u32": h(adds_amd.HostGraph.rmat(10, 8, seed=3, wkind="u32", lo=1, hi=100)),
        "grid30x40_sc25_f32": h(adds_amd.HostGraph.grid(30, 40, 25, seed=2, wkind="f32", lo=1.0, hi=2.0)),
        "er2000_u32": h(adds_amd.HostGraph.er(2000, 8, seed=5, wkind="u32", lo=1, hi=1000)),
    }
    with open(os.path.join(HERE, "gen_hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1, sort_keys=True)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
