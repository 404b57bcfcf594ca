"""baselines/run_all.py, the reference's evaluation driver (run_all.sh + verify_against_*) over this build: the
CPU legs (Dijkstra on one thread, parallel delta-step with Galois' -delta) on the golden fixtures, their
output files and fd-3 lines; on the GPU, the whole driver with the drop-in binaries and the verifier."""
import json
import os
import shutil
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, os.path.dirname(HERE))
from baselines import run_all  # noqa: E402

DELTA = {"int": "3", "float": "0.25"}  # Galois -delta: shift 3 (bucket width 8) / divisor 0.25


def _setup(tmp):
    with open(os.path.join(GOLD, "fixtures.json")) as f:
        fx = json.load(f)
    for wtype in ("int", "float"):
        d = os.path.join(tmp, f"graph-{wtype}")
        os.makedirs(d)
        rows = []
        for name, meta in sorted(fx.items()):
            if meta["wtype"] == wtype:
                shutil.copy(os.path.join(GOLD, f"{name}.gr"), os.path.join(d, f"{name}.gr"))
                rows.append((f"{name}.gr", meta["src"]))
        with open(os.path.join(tmp, f"graph_list_{wtype}"), "w") as f:
            f.writelines(f"{g}\t{s}\n" for g, s in rows)
        with open(os.path.join(tmp, f"cpu_graph_list_{wtype}"), "w") as f:
            f.writelines(f"{g}\t{s}\t{DELTA[wtype]}\n" for g, s in rows)
    return fx


def _args(tmp, legs):
    return ["--inputs-int", os.path.join(tmp, "graph-int"), "--inputs-float", os.path.join(tmp, "graph-float"),
            "--lists", tmp, "--legs", legs, "--threads", "4", "--workdir", os.path.join(tmp, "w")]


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def test_cpu_legs(tmp_path):
    tmp = str(tmp_path)
    fx = _setup(tmp)
    assert run_all.main(_args(tmp, "dj,ds")) == 0
    w = os.path.join(tmp, "w")
    for name, meta in fx.items():
        gold = _read(os.path.join(GOLD, f"{name}.dist"))
        for leg in ("dj", "ds"):
            assert _read(os.path.join(w, f"{leg}_{meta['wtype']}_final_dist", f"{name}.gr")) == gold, (leg, name)
    for wtype in ("int", "float"):
        names = [n for n, m in fx.items() if m["wtype"] == wtype]
        for leg in ("dj", "ds"):
            res = run_all.read_results(os.path.join(w, f"{leg}_{wtype}_result"))
            assert sorted(res) == sorted(f"{n}.gr" for n in names)
            for n in names:
                ms, work = res[f"{n}.gr"]
                reached = sum(not ln.endswith(b"INF") for ln in _read(os.path.join(GOLD, f"{n}.dist")).splitlines())
                assert ms >= 0
                # Dijkstra expands each reached vertex once; delta-step at least once
                assert (work == reached) if leg == "dj" else (work >= reached), (leg, n, work, reached)
    with open(os.path.join(w, "run_all_summary.json")) as f:
        summary = json.load(f)
    assert {r["graph"] for r in summary["int"]["graphs"]} == {f"{n}.gr" for n, m in fx.items() if m["wtype"] == "int"}


def test_cpu_leg_failure_is_reported(tmp_path):
    """A graph that does not exist fails its run and the driver's exit status, like a failed verify."""
    tmp = str(tmp_path)
    _setup(tmp)
    with open(os.path.join(tmp, "cpu_graph_list_int"), "a") as f:
        f.write("missing.gr\t0\t3\n")
    assert run_all.main(_args(tmp, "dj")) == 1


@pytest.mark.gpu
def test_full_driver(tmp_path):
    """The drop-in binaries, the near-far GPU baseline, both CPU legs and verify_against_{dj,ds,nf}, as
    run_all.sh runs them."""
    tmp = str(tmp_path)
    fx = _setup(tmp)
    assert run_all.main(_args(tmp, "ads,nf,dj,ds")) == 0
    w = os.path.join(tmp, "w")
    with open(os.path.join(w, "run_all_summary.json")) as f:
        summary = json.load(f)
    for wtype in ("int", "float"):
        assert summary[wtype]["checks"] == {"ads_vs_dj": True, "ads_vs_ds": True, "ads_vs_nf": True}
        for r in summary[wtype]["graphs"]:
            assert r["ads_ms"] >= 0 and r["nf_ms"] >= 0 and "speedup_vs_dj" in r
    for name, meta in fx.items():
        for leg in ("ads", "nf"):
            assert _read(os.path.join(w, f"{leg}_{meta['wtype']}_final_dist", f"{name}.gr")) == \
                _read(os.path.join(GOLD, f"{name}.dist")), (leg, name)


def test_cpu_command_contract(tmp_path):
    """sssp_cpu: usage errors exit 2; a missing output directory is reported (exit 1); the int -delta is
    a shift and the float -delta a divisor, and neither changes the distances."""
    import subprocess
    g = os.path.join(GOLD, "rmat8_int.gr")
    r = subprocess.run([run_all.CPU_BIN, "int", g, "-algo=bellman"], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([run_all.CPU_BIN, "int", g, "-algo=dijkstra", "-startNode=0"], cwd=str(tmp_path),
                       capture_output=True, text=True)
    assert r.returncode == 1 and "cannot write" in r.stderr
    outs = []
    for d in ("0", "3", "12"):
        os.makedirs(tmp_path / "ds_int_final_dist", exist_ok=True)
        r = subprocess.run([run_all.CPU_BIN, "int", g, "-algo=deltaStep", "-startNode=0", "-delta=" + d, "-t", "3"],
                           cwd=str(tmp_path), capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        outs.append(_read(os.path.join(tmp_path, "ds_int_final_dist", "rmat8_int.gr")))
    assert outs[0] == outs[1] == outs[2] == _read(os.path.join(GOLD, "rmat8_int.dist"))
