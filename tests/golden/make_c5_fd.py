#!/usr/bin/env python3
"""Writes tests/golden/c5_fd.npz: central finite differences of the C5 loss on 1000 tree edges.

C5 (adds_amd.workloads.c5): 1024 x 1024 grid with float32 weights exp(N(0, 0.5^2)), 32 sources, per
source 4096 targets with coefficients c = k / 2^23.  The loss of source s is L_s(w) = sum_t c_t d_s[t](w).

For each source the shortest-path distances are computed in float64 by scipy (independently of this
project's code).  The tree edge of a vertex v is its tight in-edge (the one attaining d[v]); its gap is
the amount by which the second-best in-edge loses.  Sampled are tree edges with gap > 1e-2 whose head has
a target in its subtree (so that the gradient is not trivially zero) and whose subtree no near-tie (a
vertex with two in-edges within 1e-2 of each other) can change: the device builds its tree from float32
distances, whose rounding may resolve such a tie the other way and move a subtree (the float64 function
would then not be the one differentiated).  32 edges per source are sampled (seeded); the first 1000 are
kept.  For each sampled (source s, edge e):

    fd = (L_s(w + h e_e) - L_s(w - h e_e)) / (2 h),   h = 1e-3,

each L_s from a full scipy Dijkstra on the perturbed weights.  A gap above 1e-2 > h keeps the tree fixed
inside the stencil, so fd is the derivative dL_s/dw_e up to float64 rounding.  The fixture stores
(source index, edge id, fd, gap) and the CRC-32 of the weight words, so a change of the generator is
detected instead of silently comparing different graphs.

Run from the repository root (CPU only, about 4 minutes on 8 cores):  python tests/golden/make_c5_fd.py
"""
import os
import sys
import time
import zlib
from multiprocessing import Pool

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from adds_amd import workloads  # noqa: E402

H = 1e-3
GAP = 1e-2
PER_SOURCE = 32
TOTAL = 1000
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c5_fd.npz")

_G = {}


def _solve(w64, src):
    g = _G["g"]
    a = sp.csr_matrix((w64, g.col.astype(np.int32), g.row_ptr.astype(np.int64)), shape=(g.n, g.n))
    return cg.dijkstra(a, indices=int(src))


def _fd(job):
    s, src, e = job
    w = _G["w64"].copy()
    tg, c = _G["targets"][s]
    w[e] += H
    lp = float(np.dot(c, _solve(w, src)[tg]))
    w[e] -= 2 * H
    lm = float(np.dot(c, _solve(w, src)[tg]))
    return (lp - lm) / (2 * H)


def tree_and_gaps(g, d, eps=None):
    """Tight in-edge of every vertex (edge id, or -1 for the source) and its gap, from float64 distances.
    With eps: also the near-ties, as pairs (tail of the best in-edge, tail of another in-edge whose
    candidate distance is within eps of the best)."""
    tail = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.row_ptr.astype(np.int64)))
    val = d[tail] + g.w.astype(np.float64)
    head = g.col.astype(np.int64)
    order = np.lexsort((val, head))  # by head, then candidate distance
    hs = head[order]
    first = np.r_[0, np.nonzero(np.diff(hs))[0] + 1]
    cnt = np.diff(np.r_[first, len(hs)])
    best = order[first]
    second_val = np.where(cnt > 1, val[order[np.minimum(first + 1, len(hs) - 1)]], np.inf)
    te = np.full(g.n, -1, dtype=np.int64)
    gap = np.full(g.n, np.inf)
    te[hs[first]] = best
    gap[hs[first]] = second_val - val[best]
    if eps is None:
        return te, gap, tail
    pos_in_group = np.arange(len(hs)) - np.repeat(first, cnt)
    best_val = np.repeat(val[best], cnt)
    near = (pos_in_group > 0) & (val[order] - best_val <= eps)
    ties = np.stack([np.repeat(tail[best], cnt)[near], tail[order][near]], axis=1)
    return te, gap, tail, ties


def depths(parent):
    """Tree depth of every vertex (roots: 0) by pointer jumping."""
    dep = (parent >= 0).astype(np.int64)
    jump = parent.copy()
    while np.any(jump >= 0):
        ok = jump >= 0
        dep[ok] += dep[jump[ok]]
        nxt = np.full_like(jump, -1)
        nxt[ok] = jump[jump[ok]]
        jump = nxt
    return dep


def tainted_by_ties(parent, dep, ties):
    """Vertices v whose subtree gains or loses a part when a near-tie resolves the other way: a tie between
    candidate parents p1 and p2 moves a subtree from under one to under the other, which changes the
    subtree of v exactly when v lies on the path from p1 or p2 up to (not including) their lowest common
    ancestor.  Two-pointer walk over all ties at once."""
    taint = np.zeros(len(parent), dtype=bool)
    a, b = ties[:, 0].copy(), ties[:, 1].copy()
    live = a != b
    a, b = a[live], b[live]
    while a.size:
        da, db = dep[a], dep[b]
        up_a, up_b = da >= db, db >= da
        taint[a[up_a]] = True
        taint[b[up_b]] = True
        a = np.where(up_a, parent[a], a)
        b = np.where(up_b, parent[b], b)
        live = (a != b) & (a >= 0) & (b >= 0)
        a, b = a[live], b[live]
    return taint


def main():
    t0 = time.time()
    g, srcs, gd = workloads.c5()
    _G["g"] = g
    _G["w64"] = g.w.astype(np.float64)
    _G["targets"] = []
    for s in range(len(srcs)):
        tg = np.nonzero(gd[s])[0]
        _G["targets"].append((tg, gd[s, tg].astype(np.float64)))
    rng = np.random.default_rng(2024)
    jobs = []
    for s, src in enumerate(srcs):
        d = _solve(_G["w64"], src)
        te, gap, tail, ties = tree_and_gaps(g, d, eps=GAP)
        te[src] = -1
        parent = np.where(te >= 0, tail[np.maximum(te, 0)], -1)
        # the device computes the tree from float32 distances: a near-tie that its rounding resolves the
        # other way moves a subtree, so edges whose subtree such a move changes are not sampled
        taint = tainted_by_ties(parent, depths(parent), ties)
        # vertices with a target in their subtree: walk up from the targets, one tree level at a time
        marked = np.zeros(g.n, dtype=bool)
        front = np.unique(_G["targets"][s][0])
        while front.size:
            marked[front] = True
            front = parent[front]
            front = np.unique(front[front >= 0])
            front = front[~marked[front]]
        cand = np.nonzero(marked & (te >= 0) & (gap > GAP) & ~taint)[0]
        pick = rng.choice(cand, size=min(PER_SOURCE, cand.size), replace=False)
        jobs += [(s, int(src), int(te[v])) for v in np.sort(pick)]
        _G.setdefault("gaps", {}).update({(s, int(te[v])): float(gap[v]) for v in pick})
        print(f"source {s} ({src}): {cand.size} candidate edges ({int(taint.sum())} vertices near a tie)",
              flush=True)
    jobs = jobs[:TOTAL]
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        fd = pool.map(_fd, jobs, chunksize=4)
    np.savez_compressed(
        OUT,
        source=np.array([j[0] for j in jobs], dtype=np.int32),
        edge=np.array([j[2] for j in jobs], dtype=np.int64),
        fd=np.array(fd, dtype=np.float64),
        gap=np.array([_G["gaps"][(j[0], j[2])] for j in jobs], dtype=np.float64),
        w_crc32=np.array([zlib.crc32(g.w_bits.tobytes())], dtype=np.uint64),
        h=np.array([H]),
    )
    print(f"wrote {OUT}: {len(jobs)} edges, |fd| median {np.median(np.abs(fd)):.4g}, "
          f"{time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
