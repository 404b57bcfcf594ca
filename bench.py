#!/usr/bin/env python3
"""bench.py -- ADDS SSSP on one MI355X.

Headline workload (the metric's single-GPU configuration): single-source SSSP on RMAT scale 24, edge
factor 32 (a/b/c/d = .57/.19/.19/.05, symmetrised, deduplicated: 16.8M vertices, ~0.5B directed edges),
uint32 weights uniform in [1, 65535], source = highest-degree vertex.  One step = one solve (dist fill +
persistent kernel + worklist cleanup) with the CSR resident in HBM.  value = GTEPS = m_r / t, where m_r
is the sum of out-degrees of the reached vertices and t the wall time per step.

Also reported (extras): the float variant of the same graph; RMAT-16 (C1) and the 2048^2 grid (C2);
256 sources on RMAT-20 in one persistent launch replayed from a HIP graph vs a loop of single-source
launches (C4a); 64 uneven grid/ER graphs in one launch vs a loop (C4b); forward and hand-written
backward on a 1024^2 grid with 32 sources (C5).

roofline: HBM-bound; algorithmic bytes per solve B = 4n + 8 n_r + 12 m_r; achieved = B / (mean
persistent-kernel time from HIP events on the solve stream); peak 8 TB/s.
cpu_baseline: the oracle's single-thread Dijkstra on the same graph and source (kind "port").
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--only", default="all", help="all | c3int | c3float | c1 | c2 | c4a | c4b | c5 (comma list)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--scale", type=int, default=24, help="RMAT scale of the headline graph")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="per-launch HBM bytes measured by rocprofv3 --pmc (optional)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read MIB, write MIB), 3 times: the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def bytes_alg(n, n_r, m_r):
    return 4 * n + 8 * n_r + 12 * m_r


def time_single(dg, src, dist_ptr, steps, warmup, torch, **params):
    for _ in range(warmup):
        dg.sssp_device(src, dist_ptr, **params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [dg.sssp_device(src, dist_ptr, **params) for _ in range(steps)]
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, stats


def near_far(dg, src, adds_dist, adds_ms, torch, reps=3):
    """The reference's GPU comparison baseline (bulk-synchronous near-far at its default delta) on the same
    graph and source: median solve ms, its work, and whether its distances equal the ADDS solve's."""
    buf = torch.empty_like(adds_dist)
    sts = [dg.sssp_nearfar_device(src, buf.data_ptr()) for _ in range(reps)]
    torch.cuda.synchronize()
    ms = float(np.median([s["ms"] for s in sts]))
    st = sts[-1]
    return {"ms": ms, "rounds": st["probes"], "threshold_moves": st["rebuckets"],
            "items": st["vertices_expanded"], "edges_scanned": st["edges_scanned"],
            "same_as_adds": bool(torch.equal(buf, adds_dist)), "adds_speedup": ms / adds_ms}


def run_c3(host, name, steps, warmup, torch, adds_amd, oracle, do_cpu, certify=True, tune=True):
    src = host.max_degree_vertex()
    dg = adds_amd.Graph(host)
    params = {}
    if tune:
        from adds_amd.autotune import tune as autotune
        t0 = time.perf_counter()
        params = autotune(dg, host, [src])
        log(f"{name}: tuned {params} ({time.perf_counter() - t0:.1f}s)")
    dist = torch.empty(host.n, dtype=torch.int32, device="cuda")
    wall, stats = time_single(dg, src, dist.data_ptr(), steps, warmup, torch, **params)
    d = dist.cpu().numpy().view(np.uint32)
    n_r, m_r = host.reached_counts(d)
    kms = float(np.mean([s["kernel_ms"] for s in stats]))
    sms = float(np.mean([s["ms"] for s in stats]))
    res = {
        "n": host.n, "m": host.m, "src": src, "n_r": n_r, "m_r": m_r, "params": params,
        "ms_per_step": wall * 1e3, "solve_ms_event": sms, "kernel_ms_event": kms,
        "gteps": m_r / wall / 1e9, "gteps_kernel": m_r / (kms / 1e3) / 1e9,
        "bytes_alg": bytes_alg(host.n, n_r, m_r),
        "work": {k: stats[-1][k] for k in ("vertices_expanded", "edges_scanned", "relaxations", "pushes",
                                           "stale_skips", "hub_tasks", "rebuckets")},
    }
    if certify and not do_cpu:
        t0 = time.perf_counter()
        c = oracle.certify(host.row_ptr, host.col, host.w, src, d, host.is_float)
        res["certificate"] = "ok" if c == 0 else f"FAILED({c})"
        log(f"{name}: certificate {res['certificate']} ({time.perf_counter() - t0:.1f}s)")
    # the same solve with the distances copied back to host memory (the CLI's path): PCIe-inclusive
    hw = []
    for _ in range(3):
        t0 = time.perf_counter()
        dg.sssp(src, **params)
        hw.append(time.perf_counter() - t0)
    res["host_out_ms"] = float(np.median(hw)) * 1e3
    res["host_out_gteps"] = m_r / np.median(hw) / 1e9
    res["near_far"] = near_far(dg, src, dist, sms, torch)
    log(f"{name}: near-far baseline {res['near_far']}")
    if do_cpu:
        t0 = time.perf_counter()
        ref = oracle.dijkstra(host.row_ptr, host.col, host.w, src, host.is_float)
        el = time.perf_counter() - t0
        res["cpu_dijkstra_s"] = el
        res["bit_exact_vs_dijkstra"] = bool(np.array_equal(ref, d))
        log(f"{name}: cpu dijkstra {el:.1f}s, bit-exact {res['bit_exact_vs_dijkstra']}")
        # the reference's other CPU baseline: parallel delta-stepping on the host's cores
        threads = min(os.cpu_count() or 1, 16)
        delta = params.get("delta") or 4096.0
        t0 = time.perf_counter()
        ds = oracle.delta_step(host.row_ptr, host.col, host.w, src, host.is_float, delta, threads)
        el = time.perf_counter() - t0
        res["cpu_delta_step"] = {"s": el, "threads": threads, "delta": delta, "gteps": m_r / el / 1e9,
                                 "bit_exact": bool(np.array_equal(ds, ref))}
        log(f"{name}: cpu delta-step {el:.1f}s on {threads} threads")
    dg.close()
    del dist
    log(f"{name}: {res['ms_per_step']:.3f} ms/step, {res['gteps']:.2f} GTEPS, kernel {kms:.3f} ms")
    return res


def tuned(dg, host, srcs, name):
    from adds_amd.autotune import tune as autotune
    t0 = time.perf_counter()
    params = autotune(dg, host, srcs)
    log(f"{name}: tuned {params} ({time.perf_counter() - t0:.1f}s)")
    return params


def run_small(host, name, src, steps, warmup, torch, adds_amd, oracle):
    dg = adds_amd.Graph(host)
    params = tuned(dg, host, [src], name)
    dist = torch.empty(host.n, dtype=torch.int32, device="cuda")
    wall, stats = time_single(dg, src, dist.data_ptr(), steps, warmup, torch, **params)
    d = dist.cpu().numpy().view(np.uint32)
    n_r, m_r = host.reached_counts(d)
    ref = oracle.dijkstra(host.row_ptr, host.col, host.w, src, host.is_float)
    res = {"n": host.n, "m": host.m, "n_r": n_r, "m_r": m_r, "params": params, "ms_per_step": wall * 1e3,
           "kernel_ms_event": float(np.mean([s["kernel_ms"] for s in stats])),
           "gteps": m_r / wall / 1e9, "bit_exact_vs_dijkstra": bool(np.array_equal(ref, d))}
    res["near_far"] = near_far(dg, src, dist, float(np.mean([s["ms"] for s in stats])), torch)
    dg.close()
    log(f"{name}: {res}")
    return res


def run_c4a(steps, torch, adds_amd, oracle):
    host = adds_amd.HostGraph.rmat(20, 16, seed=4, wkind="f32", lo=1.0, hi=10.0)
    srcs = host.sample_component(host.max_degree_vertex(), 256, seed=40)
    k = len(srcs)
    dg = adds_amd.Graph(host)
    params = tuned(dg, host, srcs, "c4a")
    dist = torch.empty((k, host.n), dtype=torch.int32, device="cuda")
    srcs_dev = torch.from_numpy(srcs.astype(np.int32)).cuda()
    sg = adds_amd.SolveGraph(dg, srcs_dev.data_ptr(), k, dist.data_ptr(), **params)
    sg.launch()
    sg.status()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sg.launch()
        st = sg.status()
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t0) / steps
    dall = dist.cpu().numpy().view(np.uint32)
    # loop of single-source launches into the same buffer rows
    one = torch.empty(host.n, dtype=torch.int32, device="cuda")
    p1 = tuned(dg, host, [int(srcs[0])], "c4a single-source")  # the loop gets its own best parameters
    dg.sssp_device(int(srcs[0]), one.data_ptr(), **p1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in srcs:
        dg.sssp_device(int(s), one.data_ptr(), **p1)
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    check = [0, k // 2, k - 1]
    exact = all(np.array_equal(oracle.dijkstra(host.row_ptr, host.col, host.w, srcs[i], True), dall[i])
                for i in check)
    m_r_total = sum(host.reached_counts(dall[i])[1] for i in range(k))
    res = {"n": host.n, "m": host.m, "sources": k, "batched_ms": batched_s * 1e3, "loop_ms": loop_s * 1e3,
           "batched_sources_per_s": k / batched_s, "loop_sources_per_s": k / loop_s,
           "speedup": loop_s / batched_s, "batched_gteps": m_r_total / batched_s / 1e9,
           "bit_exact_sampled": exact, "work": {"edges_scanned": st["edges_scanned"]}}
    sg.close()
    dg.close()
    log(f"c4a: {res}")
    return res


def c4b_graphs(adds_amd):
    rng = np.random.default_rng(4242)  # sizes only; graph content comes from the library's seeded PRNG
    gs = []
    for i in range(64):
        n = int(round(np.exp(np.log(1e4) + rng.random() * (np.log(1e6) - np.log(1e4)))))
        if i % 2 == 0:
            side = max(2, int(round(np.sqrt(n))))
            gs.append(adds_amd.HostGraph.grid(side, side, 0, seed=100 + i, wkind="u32", lo=1, hi=1000))
        else:
            gs.append(adds_amd.HostGraph.er(n, 8.0, seed=100 + i, wkind="u32", lo=1, hi=1000))
    return gs


def run_c4b(steps, torch, adds_amd, oracle):
    hosts = c4b_graphs(adds_amd)
    dgs = [adds_amd.Graph(h) for h in hosts]
    srcs = [0] * len(hosts)
    from adds_amd.autotune import tune_batched
    t0 = time.perf_counter()
    params = tune_batched(dgs, hosts, srcs)  # builds and tunes the union graph
    log(f"c4b: tuned {params} ({time.perf_counter() - t0:.1f}s)")
    dists, _ = adds_amd.sssp_batched(dgs, srcs, **params)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        dists, st = adds_amd.sssp_batched(dgs, srcs, **params)
    batched_s = (time.perf_counter() - t0) / steps
    outs = [torch.empty(h.n, dtype=torch.int32, device="cuda") for h in hosts]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for dg, o in zip(dgs, outs):
        dg.sssp_device(0, o.data_ptr())
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    exact = all(np.array_equal(oracle.dijkstra(h.row_ptr, h.col, h.w, 0, False), d.view(np.uint32))
                for h, d in zip(hosts, dists))
    res = {"graphs": len(hosts), "total_n": int(sum(h.n for h in hosts)), "total_m": int(sum(h.m for h in hosts)),
           "params": params, "loop_params": "defaults of each graph",
           "batched_ms": batched_s * 1e3, "loop_ms": loop_s * 1e3, "speedup": loop_s / batched_s,
           "instances_per_s": len(hosts) / batched_s, "bit_exact_all": exact}
    for d in dgs:
        d.close()
    log(f"c4b: {res}")
    return res


def run_c5(steps, torch, adds_amd, oracle):
    from adds_amd.autograd import SSSPFunction  # noqa: F401  (binding exercised by tests)
    host = adds_amd.HostGraph.grid(1024, 1024, 0, seed=5, wkind="lognormal", lo=0.5, hi=0.0)
    dg = adds_amd.Graph(host)
    k = 32
    srcs = host.sample_component(0, k, seed=55)
    dist = torch.empty((k, host.n), dtype=torch.float32, device="cuda")
    gdist = torch.zeros((k, host.n), dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(5)
    for s in range(k):
        tg = rng.choice(host.n, 4096, replace=False)
        c = rng.integers(-(1 << 23), 1 << 23, size=4096).astype(np.float32) / np.float32(1 << 23)
        gdist[s, torch.from_numpy(tg).cuda()] = torch.from_numpy(c).cuda()
    gw = torch.empty(host.m, dtype=torch.float32, device="cuda")
    params = tuned(dg, host, srcs, "c5")
    fwd, bwd = [], []
    for i in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dg.sssp_multi_device(srcs, dist.data_ptr(), **params)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        bms = dg.backward_device(srcs, dist.data_ptr(), gdist.data_ptr(), gw.data_ptr())
        torch.cuda.synchronize()
        if i:
            fwd.append((t1 - t0) * 1e3)
            bwd.append(bms)
    ref_gw, _, _ = oracle.tree_grad(host.row_ptr, host.col, host.w, srcs, dist.cpu().numpy(), gdist.cpu().numpy())
    res = {"n": host.n, "m": host.m, "sources": k, "params": params, "forward_ms": float(np.median(fwd)),
           "backward_ms": float(np.median(bwd)),
           "grad_bit_exact_vs_cpu_tree": bool(np.array_equal(ref_gw.view(np.uint32),
                                                             gw.cpu().numpy().view(np.uint32)))}
    dg.close()
    log(f"c5: {res}")
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of this "
              "path is a design in DESIGN.md, not a runnable configuration.", file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_calibration:
        n = args.copy_calibration << 18  # int32 words
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n}))
        return

    import adds_amd
    import oracle

    only = set(args.only.split(",")) if args.only != "all" else {"c3int", "c3float", "c1", "c2", "c4a", "c4b", "c5"}
    torch.cuda.init()
    ncu = adds_amd._lib._lib().adds_device_cu_count()
    log(f"device CUs: {ncu}")
    extras = {}
    head = None
    host = None
    if "c3int" in only or "c3float" in only:
        t0 = time.perf_counter()
        host = adds_amd.HostGraph.rmat(args.scale, 32, seed=1, wkind="u32", lo=1, hi=65535)
        log(f"rmat{args.scale} generated n={host.n} m={host.m} in {time.perf_counter() - t0:.1f}s")
    if "c3int" in only:
        head = run_c3(host, "c3int", args.steps, args.warmup, torch, adds_amd, oracle, not args.no_cpu)
    if "c3float" in only:
        host.reweight(seed=1, wkind="f32", lo=0.0, hi=1.0)
        extras["c3_float"] = run_c3(host, "c3float", args.steps, args.warmup, torch, adds_amd, oracle,
                                    do_cpu=False)
    host = None
    if "c1" in only:
        h = adds_amd.HostGraph.rmat(16, 16, seed=1, wkind="u32", lo=1, hi=100)
        extras["c1_rmat16_int"] = run_small(h, "c1", h.max_degree_vertex(), 20, 3, torch, adds_amd, oracle)
    if "c2" in only:
        h = adds_amd.HostGraph.grid(2048, 2048, 8385, seed=2, wkind="f32", lo=1.0, hi=2.0)
        extras["c2_grid2048_float"] = run_small(h, "c2", 0, 5, 2, torch, adds_amd, oracle)
    if "c4a" in only:
        extras["c4a_rmat20_256src"] = run_c4a(3, torch, adds_amd, oracle)
    if "c4b" in only:
        extras["c4b_64graphs"] = run_c4b(3, torch, adds_amd, oracle)
    if "c5" in only:
        extras["c5_grid1024_backward"] = run_c5(3, torch, adds_amd, oracle)

    if head is None:
        print(json.dumps({"extras": extras}))
        return
    traffic = None
    if os.path.exists(args.traffic_json):
        try:
            with open(args.traffic_json) as f:
                traffic = json.load(f).get("c3int_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    achieved = head["bytes_alg"] / (head["kernel_ms_event"] / 1e3) / 1e9
    cpu = None
    if "cpu_dijkstra_s" in head:
        cpu = {"value": head["m_r"] / head["cpu_dijkstra_s"] / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"one full single-source Dijkstra (oracle, binary heap) on the same RMAT-{args.scale} "
                         f"graph and source: {head['cpu_dijkstra_s']:.1f}s"}
    line = {
        "metric": "SSSP GTEPS, RMAT-24 ef32 uint32 weights, single source, 1 GPU",
        "value": head["gteps"], "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (seeded RMAT generator, no dataset)",
        "config": {"workload": f"rmat{args.scale}-ef32-u32[1,65535]-single-source", "n": head["n"], "m": head["m"],
                   "source": head["src"], "n_reached": head["n_r"], "m_reached": head["m_r"],
                   "certificate": head.get("certificate"),
                   "bit_exact_vs_dijkstra": head.get("bit_exact_vs_dijkstra")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"k_sssp<false, false, {head['params'].get('edge_windows') or 4}>",
                     "kernel_ms": head["kernel_ms_event"],
                     "bytes_alg_per_launch": head["bytes_alg"]},
        "cpu_baseline": cpu,
        "extras": dict(extras, c3_int={k: head[k] for k in ("params", "kernel_ms_event", "solve_ms_event",
                                                             "gteps_kernel", "host_out_ms", "host_out_gteps",
                                                             "work", "cpu_delta_step", "near_far") if k in head}),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
