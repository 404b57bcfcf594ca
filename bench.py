#!/usr/bin/env python3
"""bench.py -- ADDS SSSP on one MI355X.

Headline workload (the metric's single-GPU configuration): single-source SSSP on RMAT scale 24, edge
factor 32 (a/b/c/d = .57/.19/.19/.05, symmetrised, deduplicated: 16.8M vertices, 1.02G directed edges),
uint32 weights uniform in [1, 65535], source = highest-degree vertex.  One step = one solve (dist fill +
persistent kernel + worklist cleanup) with the CSR resident in HBM.  value = GTEPS = m_r / t, where m_r
is the sum of out-degrees of the reached vertices and t the wall time per step.

Also reported (extras), each with its own roofline entry and CPU leg: the float variant of the same graph
(C3 float); RMAT-16 (C1) and the 2048^2 grid (C2); 256 sources on RMAT-20 in one persistent launch replayed
from a HIP graph vs a loop of single-source launches (C4a); 64 uneven grid/ER graphs in one captured
launch vs a loop of per-graph launches (C4b); forward and backward on a 1024^2 grid with 32 sources (C5).
Workloads: adds_amd/workloads.py.  Solver parameters: the pinned entries of adds_amd/tuned_params.json
(--retune searches again and --write-params stores the result).

roofline: HBM-bound; algorithmic bytes per solve B = 4n + 8 n_r + 12 m_r (backward: 4m + 20 k n_r);
achieved = B / (mean persistent-kernel time from HIP events on the solve stream); peak 8 TB/s.
cpu_baseline: the oracle's single-thread Dijkstra on the same graph and source (kind "port"); the
parallel delta-stepping leg (baselines/, Galois -algo=deltaStep restated) runs on the host's CPU share
(OMP_NUM_THREADS, else the affinity mask).
"""
import argparse
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
KERNEL_SOURCES = ("adds_amd/csrc/sssp_kernels.hip", "adds_amd/csrc/device.h")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--only", default="all", help="all | c3int | c3float | c1 | c2 | c4a | c4b | c5 (comma list)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU legs")
    ap.add_argument("--no-cli", action="store_true", help="skip the drop-in binary on the headline graph")
    ap.add_argument("--profile", action="store_true",
                    help="run only the warmup and timed solves of the selected configurations (pinned parameters, "
                         "no checks, comparisons or CPU legs): the command the rocprofv3 summaries come from")
    ap.add_argument("--retune", action="store_true", help="search the solver parameters again (ignore pinned)")
    ap.add_argument("--write-params", default=None, help="store the parameters used per workload key here")
    ap.add_argument("--scale", type=int, default=24, help="RMAT scale of the headline graph")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="per-launch HBM bytes measured by rocprofv3 --pmc (used when the kernel source matches)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read MIB, write MIB), 3 times: the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def bytes_alg(n, n_r, m_r):
    return 4 * n + 8 * n_r + 12 * m_r


def roofline_entry(bytes_, ms, note=None):
    gbs = bytes_ / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    r = {"bytes_alg": int(bytes_), "ms": ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    if note:
        r["note"] = note
    return r


def run_cli(host, src, wtype):
    """The drop-in binary on the same graph and source, as the reference's run_all.sh calls it
    (`sssp -s <src> <graph>` with fd 3 redirected): the graph goes to a .gr file in /dev/shm (memory), the
    binary reads it, copies it to the device, looks its tuned parameters up and solves RUN_LOOP times
    with the distances copied back to host memory.  Reports the per-run times it prints (the first
    includes building the light-first rows) and its fd-3 line (mean ms, mean work)."""
    import re
    import subprocess
    import tempfile
    binp = os.path.join(ROOT, "adds_amd", "bin", f"sssp_{wtype}")
    if not os.path.exists(binp):
        return {"skipped": f"{binp} not built"}
    d = "/dev/shm" if os.path.isdir("/dev/shm") else None
    with tempfile.TemporaryDirectory(dir=d) as tmp:
        gr = os.path.join(tmp, "c3.gr")
        res = os.path.join(tmp, "result")
        t0 = time.perf_counter()
        host.write_gr(gr)
        write_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = subprocess.run(["bash", "-c", '"$@" 3>"$RES"', "cli", binp, "-s", str(src), gr], capture_output=True,
                           text=True, env=dict(os.environ, RES=res), timeout=600)
        wall_s = time.perf_counter() - t0
        if r.returncode != 0:
            return {"failed": r.returncode, "stderr": r.stderr[-500:]}
        runs = [float(x) * 1e3 for x in re.findall(r"Measured time for sample = ([0-9.]+)s", r.stdout)]
        tuned = re.findall(r"tuned parameters from (\S+)", r.stdout)
        with open(res) as f:
            fd3 = f.read().strip()
    tok = fd3.split()
    return {"binary": os.path.relpath(binp, ROOT), "run_loop": len(runs), "first_ms": runs[0] if runs else None,
            "runs_ms": runs, "fd3_line": fd3, "fd3_mean_ms": float(tok[1]) if len(tok) == 3 else None,
            "fd3_mean_gteps": host.m / (float(tok[1]) * 1e6) if len(tok) == 3 and float(tok[1]) > 0 else None,
            "tuned_parameters_from": tuned[0] if tuned else None, "gr_write_s": write_s, "process_wall_s": wall_s,
            "note": "per-run times as the binary prints them (\"Measured time for sample = %.3fs\", 1 ms "
                    "resolution) and the fd-3 mean: the library's solve time (fill, persistent kernel, "
                    "cleanup), the light-first rows built before the first run; GTEPS on m"}


def gather_ceiling(adds_amd, n, edges_scanned, kernel_ms):
    """The device's random-gather rate for an n-word distance array (adds_gather_ceiling: 2^28 uniformly
    random targets, no locality) beside the solve's edge rate (edges scanned, one d[v] gather each, per
    second of kernel time)."""
    import ctypes
    g = ctypes.c_double()
    adds_amd._lib.check(adds_amd._lib._lib().adds_gather_ceiling(n, 1 << 28, ctypes.byref(g)), "gather ceiling")
    rate = edges_scanned / (kernel_ms / 1e3)
    return {"random_gathers_per_s": g.value, "edges_scanned": int(edges_scanned), "edge_gathers_per_s": rate,
            "edge_rate_over_ceiling": rate / g.value if g.value else None}


def kernel_source_sha():
    h = hashlib.sha256()
    for p in KERNEL_SOURCES:
        with open(os.path.join(ROOT, p), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS") or affinity)
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "threads": max(1, min(share, affinity))}


class Params:
    """Parameter lookup per workload: pinned / cached entries, or a new search (--retune)."""

    def __init__(self, retune):
        from adds_amd import autotune
        self.at = autotune
        self.retune = retune
        self.used = {}

    @staticmethod
    def _origin(src):
        return os.path.basename(src) if src and os.sep in src else src

    def single(self, dg, host, srcs, name):
        key = self.at.workload_key(host, len(np.atleast_1d(srcs)))
        p, src = (None, None) if self.retune else self.at.lookup(key)
        if p is None:
            t0 = time.perf_counter()
            p = self.at.tune(dg, host, srcs, refresh=True)
            src = f"tuned in {time.perf_counter() - t0:.1f}s"
        self.used[key] = p
        log(f"{name}: params {p} ({self._origin(src)})")
        return p

    def batched(self, dgs, hosts, srcs, name):
        key = self.at.batched_key(hosts)
        p, src = (None, None) if self.retune else self.at.lookup(key)
        if p is None:
            t0 = time.perf_counter()
            p = self.at.tune_batched(dgs, hosts, srcs, refresh=True)
            src = f"tuned in {time.perf_counter() - t0:.1f}s"
        self.used[key] = p
        log(f"{name}: params {p} ({self._origin(src)})")
        return p


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


def cpu_delta(params, default):
    """Bucket width for the CPU delta-stepping leg: the GPU point's delta, or with a split first bucket
    (narrow buckets below the light threshold, no counterpart in plain delta-stepping) its light threshold."""
    import adds_amd._lib as L
    if params.get("flags", 0) & L.F_PULL_SPLIT and params.get("light_delta"):
        return params["light_delta"]
    return params.get("delta") or default


def cpu_single(host, src, m_r, oracle, hinfo, delta):
    """CPU legs of a single-source workload: the oracle's Dijkstra on 1 thread and the parallel
    delta-stepping on the CPU share; both checked against each other."""
    t0 = time.perf_counter()
    dj = oracle.dijkstra(host.row_ptr, host.col, host.w, src, host.is_float)
    tdj = time.perf_counter() - t0
    t0 = time.perf_counter()
    import baselines
    ds, _ = baselines.delta_step(host.row_ptr, host.col, host.w, src, host.is_float, delta, hinfo["threads"])
    tds = time.perf_counter() - t0
    out = {"dijkstra_1thread": {"s": tdj, "gteps": m_r / tdj / 1e9},
           "delta_step": {"s": tds, "gteps": m_r / tds / 1e9, "threads": hinfo["threads"], "delta": delta,
                          "same_as_dijkstra": bool(np.array_equal(ds, dj))}}
    return out, dj


def run_c3(host, src, name, steps, warmup, torch, adds_amd, oracle, prm, do_cpu, hinfo):
    t0 = time.perf_counter()
    dg = adds_amd.Graph(host)
    upload_s = time.perf_counter() - t0
    params = prm.single(dg, host, [src], name)
    dist = torch.empty(host.n, dtype=torch.int32, device="cuda")
    # cold: the first solve on a fresh handle builds the light-first rows for its light threshold and
    # allocates the worklist (the one-time cost of a drop-in run); then the warm steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cold = dg.sssp_device(src, dist.data_ptr(), **params)
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    wall, stats = time_single(dg, src, dist.data_ptr(), steps, warmup, torch, **params)
    d = dist.cpu().numpy().view(np.uint32).copy()
    n_r, m_r = host.reached_counts(d)
    kms = float(np.mean([s["kernel_ms"] for s in stats]))
    sms = float(np.mean([s["ms"] for s in stats]))
    res = {
        "n": host.n, "m": host.m, "src": src, "n_r": n_r, "m_r": m_r, "params": params,
        "ms_per_step": wall * 1e3, "solve_ms_event": sms, "kernel_ms_event": kms,
        "kernel_ms_min": float(min(s["kernel_ms"] for s in stats)),
        "gteps": m_r / wall / 1e9, "gteps_kernel": m_r / (kms / 1e3) / 1e9,
        "gteps_undirected": m_r / 2 / wall / 1e9,
        "bytes_alg": bytes_alg(host.n, n_r, m_r),
        "cold_first_solve_ms": cold_ms, "cold_kernel_ms": cold["kernel_ms"], "graph_upload_s": upload_s,
        "work": {k: stats[-1][k] for k in ("vertices_expanded", "edges_scanned", "relaxations", "pushes",
                                           "stale_skips", "hub_tasks", "rebuckets", "heavy_edges", "pulls",
                                           "pull_edges", "pull_improved")},
    }
    # the same solve with every heavy phase in the push direction (ADDS_F_NO_PULL), for comparison
    from adds_amd import _lib as L
    _, pst = time_single(dg, src, dist.data_ptr(), 3, 1, torch,
                         **dict(params, flags=params.get("flags", 0) | L.F_NO_PULL))
    pk = float(np.mean([s["kernel_ms"] for s in pst]))
    res["push_direction_only"] = {
        "kernel_ms_event": pk, "gteps_kernel": m_r / (pk / 1e3) / 1e9, "edges_scanned": pst[-1]["edges_scanned"],
        "heavy_edges": pst[-1]["heavy_edges"],
        "same_distances": bool(np.array_equal(dist.cpu().numpy().view(np.uint32), d))}
    res["roofline"] = roofline_entry(res["bytes_alg"], kms)
    res["work_efficiency"] = {"relaxations_per_m_r": res["work"]["relaxations"] / m_r,
                              "edges_scanned_per_m_r": res["work"]["edges_scanned"] / m_r}
    res["gather_ceiling"] = gather_ceiling(adds_amd, host.n, res["work"]["edges_scanned"], kms)
    log(f"{name}: random-gather ceiling {res['gather_ceiling']}")
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
    dg.close()
    del dist
    if do_cpu:
        cpu, ref = cpu_single(host, src, m_r, oracle, hinfo, cpu_delta(params, 4096.0))
        res["cpu"] = cpu
        res["bit_exact_vs_dijkstra"] = bool(np.array_equal(ref, d))
        log(f"{name}: cpu {cpu}, bit-exact {res['bit_exact_vs_dijkstra']}")
    else:
        c = oracle.certify(host.row_ptr, host.col, host.w, src, d, host.is_float)
        res["certificate"] = "ok" if c == 0 else f"FAILED({c})"
    log(f"{name}: {res['ms_per_step']:.3f} ms/step, {res['gteps']:.2f} GTEPS, kernel {kms:.3f} ms, "
        f"cold {cold_ms:.1f} ms")
    return res


def run_small(host, name, src, steps, warmup, torch, adds_amd, oracle, prm, do_cpu, hinfo, note=None):
    dg = adds_amd.Graph(host)
    params = prm.single(dg, host, [src], name)
    dist = torch.empty(host.n, dtype=torch.int32, device="cuda")
    wall, stats = time_single(dg, src, dist.data_ptr(), steps, warmup, torch, **params)
    d = dist.cpu().numpy().view(np.uint32)
    n_r, m_r = host.reached_counts(d)
    kms = float(np.mean([s["kernel_ms"] for s in stats]))
    res = {"n": host.n, "m": host.m, "n_r": n_r, "m_r": m_r, "params": params, "ms_per_step": wall * 1e3,
           "kernel_ms_event": kms, "gteps": m_r / wall / 1e9,
           "roofline": roofline_entry(bytes_alg(host.n, n_r, m_r), kms, note)}
    res["near_far"] = near_far(dg, src, dist, float(np.mean([s["ms"] for s in stats])), torch)
    dg.close()
    if do_cpu:
        cpu, ref = cpu_single(host, src, m_r, oracle, hinfo, cpu_delta(params, 1.0))
        res["cpu"] = cpu
    else:
        ref = oracle.dijkstra(host.row_ptr, host.col, host.w, src, host.is_float)
    res["bit_exact_vs_dijkstra"] = bool(np.array_equal(ref, d))
    log(f"{name}: {res}")
    return res


def run_c4a(steps, torch, adds_amd, oracle, workloads, prm, do_cpu, hinfo):
    host, srcs = workloads.c4a()
    k = len(srcs)
    dg = adds_amd.Graph(host)
    params = prm.single(dg, host, srcs, "c4a")
    dist = torch.empty((k, host.n), dtype=torch.int32, device="cuda")
    srcs_dev = torch.from_numpy(srcs.astype(np.int32)).cuda()
    sg = adds_amd.SolveGraph(dg, srcs_dev.data_ptr(), k, dist.data_ptr(), **params)
    sg.launch()
    sg.status()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    replay_ms = []
    for _ in range(steps):
        sg.launch()
        st = sg.status()
        replay_ms.append(st["ms"])
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t0) / steps
    dall = dist.cpu().numpy().view(np.uint32)
    # loop of single-source launches into one buffer, with the best parameters for one source
    one = torch.empty(host.n, dtype=torch.int32, device="cuda")
    p1 = prm.single(dg, host, [int(srcs[0])], "c4a single-source")
    dg.sssp_device(int(srcs[0]), one.data_ptr(), **p1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in srcs:
        dg.sssp_device(int(s), one.data_ptr(), **p1)
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    sg.close()
    dg.close()
    # every source checked: the certificate (pins the solution) on all, Dijkstra bit for bit on 16
    with ThreadPoolExecutor(hinfo["threads"]) as ex:
        certs = list(ex.map(lambda i: oracle.certify(host.row_ptr, host.col, host.w, srcs[i], dall[i], True),
                            range(k)))
        pick = list(range(0, k, max(1, k // 16)))
        t0 = time.perf_counter()
        refs = list(ex.map(lambda i: oracle.dijkstra(host.row_ptr, host.col, host.w, srcs[i], True), pick))
        par_s = time.perf_counter() - t0
    exact16 = all(np.array_equal(dall[i], r) for i, r in zip(pick, refs))
    reach = [host.reached_counts(dall[i]) for i in range(k)]
    m_r_total = sum(r[1] for r in reach)
    b_total = sum(bytes_alg(host.n, r[0], r[1]) for r in reach)
    mean_replay = float(np.mean(replay_ms))
    res = {"n": host.n, "m": host.m, "sources": k, "params": params, "loop_params": p1,
           "batched_ms": batched_s * 1e3, "replay_ms_event": mean_replay, "loop_ms": loop_s * 1e3,
           "batched_sources_per_s": k / batched_s, "loop_sources_per_s": k / loop_s,
           "speedup": loop_s / batched_s, "batched_gteps": m_r_total / batched_s / 1e9,
           "certificate_all_256": all(c == 0 for c in certs), "bit_exact_16_vs_dijkstra": exact16,
           "roofline": roofline_entry(b_total, mean_replay, "replay time (fill + kernel + cleanup)"),
           "work": {k: st[k] for k in ("edges_scanned", "heavy_edges", "pulls", "pull_edges")}}
    if do_cpu:
        t0 = time.perf_counter()
        for i in pick[:4]:
            oracle.dijkstra(host.row_ptr, host.col, host.w, srcs[i], True)
        one_s = (time.perf_counter() - t0) / 4
        res["cpu"] = {"dijkstra_1thread_sources_per_s": 1.0 / one_s,
                      "dijkstra_parallel_sources_per_s": len(pick) / par_s, "threads": hinfo["threads"],
                      "sample": f"4 sources on 1 thread, {len(pick)} on {hinfo['threads']} threads"}
    log(f"c4a: {res}")
    return res


def run_c4b(steps, torch, adds_amd, oracle, workloads, prm, do_cpu, hinfo):
    hosts, srcs = workloads.c4b()
    dgs = [adds_amd.Graph(h) for h in hosts]
    params = prm.batched(dgs, hosts, srcs, "c4b")
    outs = [torch.empty(h.n, dtype=torch.int32, device="cuda") for h in hosts]
    sg = adds_amd.SolveGraph.batched(dgs, srcs, [o.data_ptr() for o in outs], **params)
    sg.launch()
    sg.status()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    replay_ms = []
    for _ in range(steps):
        sg.launch()
        st = sg.status()
        replay_ms.append(st["ms"])
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t0) / steps
    dists = [o.cpu().numpy().view(np.uint32) for o in outs]
    sg.close()
    # loop: each graph with its own tuned parameters
    lp = [prm.single(dg, h, [0], f"c4b graph {i}") for i, (dg, h) in enumerate(zip(dgs, hosts))]
    for dg, o, p in zip(dgs, outs, lp):
        dg.sssp_device(0, o.data_ptr(), **p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for dg, o, p in zip(dgs, outs, lp):
        dg.sssp_device(0, o.data_ptr(), **p)
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    for d in dgs:
        d.close()
    with ThreadPoolExecutor(hinfo["threads"]) as ex:
        t0 = time.perf_counter()
        refs = list(ex.map(lambda i: oracle.dijkstra(hosts[i].row_ptr, hosts[i].col, hosts[i].w, srcs[i], False),
                           range(len(hosts))))
        par_s = time.perf_counter() - t0
    exact = all(np.array_equal(r, d) for r, d in zip(refs, dists))
    reach = [h.reached_counts(d) for h, d in zip(hosts, dists)]
    b_total = sum(bytes_alg(h.n, r[0], r[1]) for h, r in zip(hosts, reach))
    mean_replay = float(np.mean(replay_ms))
    res = {"graphs": len(hosts), "total_n": int(sum(h.n for h in hosts)), "total_m": int(sum(h.m for h in hosts)),
           "params": params, "loop_params": "each graph's own tuned parameters",
           "batched_ms": batched_s * 1e3, "replay_ms_event": mean_replay, "loop_ms": loop_s * 1e3,
           "speedup": loop_s / batched_s, "instances_per_s": len(hosts) / batched_s, "bit_exact_all": exact,
           "work": {k: st[k] for k in ("edges_scanned", "heavy_edges", "pulls", "pull_edges")},
           "roofline": roofline_entry(b_total, mean_replay, "replay time (fill + kernel + cleanup + copies)")}
    if do_cpu:
        t0 = time.perf_counter()
        for i in range(len(hosts)):
            oracle.dijkstra(hosts[i].row_ptr, hosts[i].col, hosts[i].w, srcs[i], False)
        res["cpu"] = {"dijkstra_1thread_instances_per_s": len(hosts) / (time.perf_counter() - t0),
                      "dijkstra_parallel_instances_per_s": len(hosts) / par_s, "threads": hinfo["threads"]}
    log(f"c4b: {res}")
    return res


def run_c5(steps, torch, adds_amd, oracle, workloads, prm, do_cpu, hinfo):
    host, srcs, gd = workloads.c5()
    dg = adds_amd.Graph(host)
    k = len(srcs)
    dist = torch.empty((k, host.n), dtype=torch.float32, device="cuda")
    gdist = torch.from_numpy(gd).cuda()
    gw = torch.empty(host.m, dtype=torch.float32, device="cuda")
    params = prm.single(dg, host, srcs, "c5")
    fwd, fwd_k, bwd = [], [], []
    for i in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = dg.sssp_multi_device(srcs, dist.data_ptr(), **params)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        bms = dg.backward_device(srcs, dist.data_ptr(), gdist.data_ptr(), gw.data_ptr())
        torch.cuda.synchronize()
        if i:
            fwd.append((t1 - t0) * 1e3)
            fwd_k.append(st["kernel_ms"])
            bwd.append(bms)
    d = dist.cpu().numpy()
    got = gw.cpu().numpy()
    dg.close()
    t0 = time.perf_counter()
    ref_gw, _, scale = oracle.tree_grad(host.row_ptr, host.col, host.w, srcs, d, gd)
    cpu_grad_s = time.perf_counter() - t0
    reach = [host.reached_counts(d[i]) for i in range(k)]
    n_r_total = sum(r[0] for r in reach)
    res = {"n": host.n, "m": host.m, "sources": k, "params": params, "forward_ms": float(np.median(fwd)),
           "forward_kernel_ms": float(np.median(fwd_k)), "backward_ms": float(np.median(bwd)),
           "grad_scale": scale,
           "grad_bit_exact_vs_cpu_reference": bool(np.array_equal(ref_gw.view(np.uint32), got.view(np.uint32))),
           "roofline_forward": roofline_entry(sum(bytes_alg(host.n, r[0], r[1]) for r in reach),
                                              float(np.median(fwd_k))),
           "roofline_backward": roofline_entry(4 * host.m + 20 * n_r_total, float(np.median(bwd)),
                                               "B = 4m + 20 k n_r (a latency-bound tree pass)")}
    if do_cpu:
        t0 = time.perf_counter()
        for i in range(4):
            oracle.dijkstra(host.row_ptr, host.col, host.w, srcs[i], True)
        res["cpu"] = {"dijkstra_1thread_forward_ms_per_batch": (time.perf_counter() - t0) / 4 * k * 1e3,
                      "tree_gradient_1thread_ms_per_batch": cpu_grad_s * 1e3,
                      "sample": "forward: 4 sources scaled to 32; gradient: the whole batch (oracle.tree_grad)"}
    log(f"c5: {res}")
    return res


def profile_runs(only, args, torch, adds_amd, workloads, prm):
    """The solves of each selected configuration and nothing else, so that every dispatch a profiler sees
    belongs to a timed solve: `warmup` + `steps` solves (single source), captured replays (c4a, c4b) or
    forward + backward passes (c5).  Prints one JSON line of mean kernel times."""
    out = {}

    def single(name, host, src):
        dg = adds_amd.Graph(host)
        p = prm.single(dg, host, [src], name)
        dist = torch.empty(host.n, dtype=torch.int32, device="cuda")
        _, st = time_single(dg, src, dist.data_ptr(), args.steps, args.warmup, torch, **p)
        out[name] = {"kernel_ms": float(np.mean([s["kernel_ms"] for s in st])), "solves": args.warmup + args.steps,
                     "pulls_per_solve": st[-1]["pulls"], "params": p}
        dg.close()

    if "c3int" in only or "c3float" in only:
        host, src = workloads.c3("u32", scale=args.scale)
        if "c3int" in only:
            single("c3int", host, src)
        if "c3float" in only:
            workloads.c3_reweight_float(host)
            single("c3float", host, src)
    if "c1" in only:
        single("c1", *workloads.c1())
    if "c2" in only:
        single("c2", *workloads.c2())
    if "c4a" in only or "c4b" in only:
        for name in ("c4a", "c4b"):
            if name not in only:
                continue
            if name == "c4a":
                host, srcs = workloads.c4a()
                dg = adds_amd.Graph(host)
                p = prm.single(dg, host, srcs, name)
                dist = torch.empty((len(srcs), host.n), dtype=torch.int32, device="cuda")
                sd = torch.from_numpy(srcs.astype(np.int32)).cuda()
                sg = adds_amd.SolveGraph(dg, sd.data_ptr(), len(srcs), dist.data_ptr(), **p)
                keep = [dg]
            else:
                hosts, srcs = workloads.c4b()
                keep = [adds_amd.Graph(h) for h in hosts]
                p = prm.batched(keep, hosts, srcs, name)
                outs = [torch.empty(h.n, dtype=torch.int32, device="cuda") for h in hosts]
                sg = adds_amd.SolveGraph.batched(keep, srcs, [o.data_ptr() for o in outs], **p)
            ms = []
            for i in range(args.warmup + args.steps):
                sg.launch()
                ms.append(sg.status()["ms"])
            out[name] = {"replay_ms": float(np.mean(ms[args.warmup:])), "replays": len(ms), "params": p}
            sg.close()
            for g in keep:
                g.close()
    if "c5" in only:
        host, srcs, gd = workloads.c5()
        dg = adds_amd.Graph(host)
        p = prm.single(dg, host, srcs, "c5")
        dist = torch.empty((len(srcs), host.n), dtype=torch.float32, device="cuda")
        gdist = torch.from_numpy(gd).cuda()
        gw = torch.empty(host.m, dtype=torch.float32, device="cuda")
        fk, bk = [], []
        for i in range(args.warmup + args.steps):
            st = dg.sssp_multi_device(srcs, dist.data_ptr(), **p)
            bk.append(dg.backward_device(srcs, dist.data_ptr(), gdist.data_ptr(), gw.data_ptr()))
            fk.append(st["kernel_ms"])
        out["c5"] = {"forward_kernel_ms": float(np.mean(fk[args.warmup:])), "backward_ms": float(np.mean(bk[args.warmup:])),
                     "passes": len(fk), "params": p}
        dg.close()
    torch.cuda.synchronize()
    print(json.dumps({"profile": out}))


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
    from adds_amd import workloads

    only = set(args.only.split(",")) if args.only != "all" else {"c3int", "c3float", "c1", "c2", "c4a", "c4b", "c5"}
    torch.cuda.init()
    ncu = adds_amd._lib._lib().adds_device_cu_count()
    hinfo = host_info()
    log(f"device CUs: {ncu}; host {hinfo}")
    prm = Params(args.retune)
    if args.profile:
        profile_runs(only, args, torch, adds_amd, workloads, prm)
        return
    do_cpu = not args.no_cpu
    extras = {}
    head = None
    if "c3int" in only or "c3float" in only:
        t0 = time.perf_counter()
        host, src = workloads.c3("u32", scale=args.scale)
        log(f"rmat{args.scale} generated n={host.n} m={host.m} in {time.perf_counter() - t0:.1f}s")
        if "c3int" in only:
            head = run_c3(host, src, "c3int", args.steps, args.warmup, torch, adds_amd, oracle, prm, do_cpu, hinfo)
            if not args.no_cli:
                head["cli"] = run_cli(host, src, "int")
                log(f"c3int: drop-in binary {head['cli']}")
        if "c3float" in only:
            workloads.c3_reweight_float(host)
            extras["c3_float"] = run_c3(host, src, "c3float", args.steps, args.warmup, torch, adds_amd, oracle, prm,
                                        do_cpu, hinfo)
        host = None
    if "c1" in only:
        h, s = workloads.c1()
        extras["c1_rmat16_int"] = run_small(h, "c1", s, 20, 3, torch, adds_amd, oracle, prm, do_cpu, hinfo,
                                            note="16.5 MB working set: L2/MALL-resident, not an HBM case")
    if "c2" in only:
        h, s = workloads.c2()
        extras["c2_grid2048_float"] = run_small(h, "c2", s, 5, 2, torch, adds_amd, oracle, prm, do_cpu, hinfo,
                                                note="high-diameter grid: latency-bound")
    if "c4a" in only:
        extras["c4a_rmat20_256src"] = run_c4a(3, torch, adds_amd, oracle, workloads, prm, do_cpu, hinfo)
    if "c4b" in only:
        extras["c4b_64graphs"] = run_c4b(3, torch, adds_amd, oracle, workloads, prm, do_cpu, hinfo)
    if "c5" in only:
        extras["c5_grid1024_backward"] = run_c5(3, torch, adds_amd, oracle, workloads, prm, do_cpu, hinfo)

    if args.write_params:
        cache = {}
        if os.path.exists(args.write_params):
            with open(args.write_params) as f:
                cache = json.load(f)
        for key, p in prm.used.items():
            cache[key] = {"params": p, "kernel_ms": None}
        with open(args.write_params, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        log(f"wrote {len(prm.used)} parameter sets to {args.write_params}")

    if head is None:
        print(json.dumps({"extras": extras, "host": hinfo}))
        return
    traffic, traffic_note = None, "no rocprofv3 traffic measurement for this kernel source"
    sha = kernel_source_sha()
    if os.path.exists(args.traffic_json):
        try:
            with open(args.traffic_json) as f:
                tj = json.load(f)
            if tj.get("kernel_source_sha") == sha and tj.get("params") == head["params"]:
                traffic = tj.get("c3int_bytes_per_launch")
                traffic_note = tj.get("source")
            else:
                traffic_note = "profiles/traffic.json was measured on another kernel source or parameter set"
        except (OSError, ValueError):
            pass
    achieved = head["bytes_alg"] / (head["kernel_ms_event"] / 1e3) / 1e9
    cpu = None
    if "cpu" in head:
        dj = head["cpu"]["dijkstra_1thread"]
        cpu = {"value": dj["gteps"], "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"one full single-source Dijkstra (oracle, binary heap, 1 thread) on the same "
                         f"RMAT-{args.scale} graph and source: {dj['s']:.1f}s",
               "host_cpu_model": hinfo["cpu_model"], "host_nproc": hinfo["nproc"],
               "host_cpu_share": hinfo["threads"],
               "parallel_delta_step": head["cpu"]["delta_step"]}
    line = {
        "metric": "SSSP GTEPS, RMAT-24 ef32 uint32 weights, single source, 1 GPU",
        "value": head["gteps"], "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (seeded RMAT generator, no dataset)",
        "config": {"workload": f"rmat{args.scale}-ef32-u32[1,65535]-single-source", "n": head["n"], "m": head["m"],
                   "source": head["src"], "n_reached": head["n_r"], "m_reached": head["m_r"],
                   "gteps_undirected": head["gteps_undirected"],
                   "certificate": head.get("certificate"),
                   "bit_exact_vs_dijkstra": head.get("bit_exact_vs_dijkstra")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "kernel": f"one solve: k_sssp<false, false, {head['params'].get('edge_windows') or 4}> "
                               "(persistent launch), k_pull_count, k_pull_decide, k_pull_a, k_pull_b, "
                               "k_pull_publish, k_sssp (continuation); HIP events around the sequence",
                     "kernel_ms": head["kernel_ms_event"], "kernel_source_sha": sha,
                     "bytes_alg_per_launch": head["bytes_alg"],
                     "random_gather_ceiling": head["gather_ceiling"]},
        "cpu_baseline": cpu,
        "extras": dict(extras, c3_int={k: head[k] for k in ("params", "kernel_ms_event", "kernel_ms_min",
                                                             "solve_ms_event", "gteps_kernel", "host_out_ms",
                                                             "host_out_gteps", "cold_first_solve_ms",
                                                             "cold_kernel_ms", "work", "work_efficiency", "cpu",
                                                             "near_far", "cli", "push_direction_only")
                                       if k in head}),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
