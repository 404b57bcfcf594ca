"""The evaluation driver of the reference (run_all.sh + verify_against_*), restated over this build.

    python baselines/run_all.py --inputs-int <dir> --inputs-float <dir> [--lists <dir>] [--threads N]
                            [--legs ads,nf,dj,ds] [--workdir <dir>]

For each weight type it runs, like run_all.sh:
  * ads_{int,float}: adds_amd/bin/sssp_{int,float} -s <src> -o ads_X_final_dist/<g> <inputs>/<g> over
    graph_list_X (run_sssp, run_all.sh:4-18, 20-37);
  * nf_{int,float}: the near-far GPU baseline, adds_amd/bin/sssp_nf_{int,float}, with the same command line
    (the nf blocks of run_all.sh:34-38, 59-63);
  * dj_{int,float}: Dijkstra on one thread over cpu_graph_list_X (run_dj, run_all.sh:94-108, 113-135);
  * ds_{int,float}: parallel delta-step on --threads threads with the list's third column as Galois'
    -delta (run_ds, run_all.sh:78-92);
each program appending "<graph> <ms> <work>" to fd 3, redirected to <impl>_result.  The CPU legs run
baselines/sssp_cpu (Galois sssp-cpu's command line over this directory's CPU baselines).  Then it checks
ads against dj, ds and nf with adds_amd/bin/adds_verify (verify.py
semantics; verify_against_dj, verify_against_ds, verify_against_nf) and aggregates the fd-3 files into one
table per type (mean ms and work per graph and program, ads speed-up over each other leg), also written
to run_all_summary.json.  Exit status 1 when any comparison mismatches.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SSSP_BIN = {t: os.path.join(ROOT, "adds_amd", "bin", f"sssp_{t}") for t in ("int", "float")}
NF_BIN = {t: os.path.join(ROOT, "adds_amd", "bin", f"sssp_nf_{t}") for t in ("int", "float")}
VERIFY_BIN = os.path.join(ROOT, "adds_amd", "bin", "adds_verify")
CPU_BIN = os.path.join(ROOT, "baselines", "sssp_cpu")


def read_list(path):
    """Graph-list lines: graph, source[, delta] separated by white space (the lists use tabs)."""
    rows = []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if tok:
                rows.append(tok)
    return rows


def run_fd3(cmd, result, cwd):
    """Runs cmd with fd 3 appended to `result` (the `3>${impl_name}_result` of run_all.sh)."""
    return subprocess.run(["bash", "-c", '"$@" 3>>"$ADDS_RESULT"', "run_fd3"] + cmd, cwd=cwd,
                          env=dict(os.environ, ADDS_RESULT=result), stdout=subprocess.DEVNULL,
                          stderr=subprocess.PIPE, text=True)


def fresh_dir(path):
    shutil.rmtree(path, ignore_errors=True)
    os.makedirs(path)


def run_leg(impl, wtype, rows, inputs, workdir, threads, log):
    """One run_sssp / run_dj / run_ds block: fresh <impl>_final_dist/ and <impl>_result."""
    kind = impl.split("_")[0]
    out_dir = os.path.join(workdir, f"{impl}_final_dist")
    result = os.path.join(workdir, f"{impl}_result")
    fresh_dir(out_dir)
    if os.path.exists(result):
        os.remove(result)
    failed = []
    for tok in rows:
        g, src = tok[0], tok[1]
        gpath = os.path.join(os.path.abspath(inputs), g)
        if kind in ("ads", "nf"):
            cmd = [(SSSP_BIN if kind == "ads" else NF_BIN)[wtype], "-s", src, "-o", os.path.join(out_dir, g), gpath]
        else:
            delta = tok[2] if len(tok) > 2 else ("1" if wtype == "float" else "13")
            cmd = [CPU_BIN, wtype, gpath, "-noverify", "-algo=" + ("deltaStep" if kind == "ds" else "dijkstra"),
                   "-startNode=" + src, "-delta=" + delta, "-t", str(threads if kind == "ds" else 1)]
        r = run_fd3(cmd, result, workdir)
        if r.returncode != 0:
            failed.append(g)
            log(f"{impl}: {g} failed ({r.returncode}): {r.stderr.strip()[-300:]}")
    return failed


def read_results(path):
    """fd-3 lines -> {graph: (ms, work)}; several lines for one graph are averaged."""
    acc = {}
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        for line in f:
            tok = line.split()
            if len(tok) != 3:
                continue
            a = acc.setdefault(tok[0], [0.0, 0, 0])
            a[0] += float(tok[1])
            a[1] += int(tok[2])
            a[2] += 1
    return {g: (a[0] / a[2], a[1] // a[2]) for g, a in acc.items()}


def verify(a_dir, b_dir, list_path, workdir):
    """adds_verify A/ B/ list (verify.py semantics); True when every graph matches."""
    r = subprocess.run([VERIFY_BIN, a_dir + "/", b_dir + "/", list_path], cwd=workdir, capture_output=True,
                       text=True)
    return r.returncode == 0 and "all results match!!!" in r.stdout, r.stdout


def aggregate(wtype, impls, graphs, workdir):
    res = {impl: read_results(os.path.join(workdir, f"{impl}_result")) for impl in impls}
    table = []
    for g in graphs:
        row = {"graph": g}
        for impl in impls:
            if g in res[impl]:
                row[impl.split("_")[0] + "_ms"], row[impl.split("_")[0] + "_work"] = res[impl][g]
        ads = row.get("ads_ms")
        for cpu in ("dj", "ds", "nf"):
            if ads and row.get(cpu + "_ms"):
                row["speedup_vs_" + cpu] = row[cpu + "_ms"] / ads
        table.append(row)
    return table


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--inputs-int", help="directory of the int .gr files (run_all.sh: ./inputs/graph-int)")
    ap.add_argument("--inputs-float", help="directory of the float .gr files (./inputs/graph-float)")
    ap.add_argument("--lists", default=".", help="directory holding graph_list_{int,float} and "
                                                 "cpu_graph_list_{int,float}")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1, help="delta-step threads (NUM_THREADS)")
    ap.add_argument("--legs", default="ads,nf,dj,ds")
    ap.add_argument("--workdir", default=".")
    args = ap.parse_args(argv)
    legs = [x for x in args.legs.split(",") if x]
    workdir = os.path.abspath(args.workdir)
    os.makedirs(workdir, exist_ok=True)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ok = True
    summary = {}
    for wtype, inputs in (("int", args.inputs_int), ("float", args.inputs_float)):
        if not inputs:
            continue
        glist = os.path.join(args.lists, f"graph_list_{wtype}")
        clist = os.path.join(args.lists, f"cpu_graph_list_{wtype}")
        rows_gpu = read_list(glist) if os.path.exists(glist) else []
        rows_cpu = read_list(clist) if os.path.exists(clist) else rows_gpu
        impls = []
        for leg in legs:
            rows = rows_gpu if leg in ("ads", "nf") else rows_cpu
            if not rows:
                continue
            impl = f"{leg}_{wtype}"
            impls.append(impl)
            failed = run_leg(impl, wtype, rows, inputs, workdir, args.threads, log)
            ok &= not failed
            log(f"{impl}: {len(rows) - len(failed)}/{len(rows)} graphs")
        checks = {}
        if "ads" in legs and rows_gpu:
            for cpu in ("dj", "ds", "nf"):
                if cpu in legs:
                    match, out = verify(os.path.join(workdir, f"ads_{wtype}_final_dist"),
                                        os.path.join(workdir, f"{cpu}_{wtype}_final_dist"), glist, workdir)
                    checks[f"ads_vs_{cpu}"] = match
                    ok &= match
                    log(f"verify ads_{wtype} vs {cpu}_{wtype}: {'all results match' if match else 'MISMATCH'}")
                    if not match:
                        log(out[-2000:])
        graphs = [t[0] for t in (rows_gpu or rows_cpu)]
        table = aggregate(wtype, impls, graphs, workdir)
        summary[wtype] = {"checks": checks, "graphs": table}
        cols = [c for c in ("ads_ms", "nf_ms", "dj_ms", "ds_ms", "speedup_vs_nf", "speedup_vs_dj", "speedup_vs_ds",
                            "ads_work", "nf_work", "ds_work")
                if any(c in r for r in table)]
        print(f"# {wtype}: " + "  ".join(["graph"] + cols))
        for r in table:
            print("  ".join([r["graph"]] + [f"{r[c]:.3f}" if isinstance(r.get(c), float) else str(r.get(c, "-"))
                                            for c in cols]))
    with open(os.path.join(workdir, "run_all_summary.json"), "w") as f:
        json.dump(summary, f, indent=1)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
