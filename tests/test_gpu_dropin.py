"""The drop-in contract on the GPU: `sssp_int` / `sssp_float` read a .gr file, write the per-vertex
distance text and the fd-3 line exactly like the reference's binaries, and adds_verify (verify.py
semantics) finds them equal to the oracle's Dijkstra output -- the reference's verify_against_dj flow
(run_all.sh:13-17, 93-107; verify_against_dj:1)."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "adds_amd", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_cli_on_golden_fixtures(tmp_path):
    with open(os.path.join(GOLDEN, "fixtures.json")) as f:
        fx = json.load(f)
    out = tmp_path / "ads_final_dist"
    out.mkdir()
    lst = tmp_path / "graph_list"
    lst.write_text("".join(f"{name}\t{meta['src']}\n" for name, meta in fx.items()))
    fd3 = tmp_path / "ads_result"
    for name, meta in fx.items():
        exe = os.path.join(BIN, "sssp_float" if meta["wtype"] == "float" else "sssp_int")
        # the same invocation as run_all.sh:16 (`$bin -s <src> -o <dir>/<g> <path>/<g>` with fd 3 redirected)
        cmd = f"{exe} -s {meta['src']} -o {out / name} {os.path.join(GOLDEN, name + '.gr')} 3>>{fd3}"
        r = subprocess.run(cmd, shell=True, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
    # golden distance files renamed as a second result directory, compared with verify semantics
    gold = tmp_path / "dj_final_dist"
    gold.mkdir()
    for name in fx:
        (gold / name).write_text(open(os.path.join(GOLDEN, name + ".dist")).read())
    r = subprocess.run([os.path.join(BIN, "adds_verify"), f"{out}/", f"{gold}/", str(lst)], capture_output=True,
                       text=True)
    assert r.returncode == 0 and "all results match!!!" in r.stdout, r.stdout
    lines = fd3.read_text().splitlines()
    assert len(lines) == len(fx)
    for line, name in zip(lines, fx):
        parts = line.split(" ")
        assert parts[0] == name + ".gr" and len(parts) == 3
        float(parts[1]), int(parts[2])


@pytest.mark.parametrize("kind", ["int", "float"])
def test_cli_uses_tuned_parameters(tmp_path, kind):
    """The binaries find the tuned parameters of their graph under the key autotune.workload_key()
    writes (here in a file named by ADDS_TUNE_FILE, with distinctive values), and the output is unchanged."""
    import numpy as np

    import adds_amd
    import oracle
    from adds_amd import autotune

    wk = ("u32", 1, 500) if kind == "int" else ("f32", 0.0, 1.0)
    g = adds_amd.HostGraph.rmat(11, 8, seed=5, wkind=wk[0], lo=wk[1], hi=wk[2])
    path = tmp_path / "g.gr"
    g.write_gr(path)
    tune = tmp_path / "tune.json"
    params = {"delta": 37.0 if kind == "int" else 0.013, "light_delta": 0.0, "wg_per_cu": 1, "chunk": 32,
              "nbuckets": 16, "parts": 2, "concurrency": 2, "task_edges": 0, "flags": 16, "edge_windows": 8,
              "tail_items": 0, "hub_degree": 0}
    tune.write_text(json.dumps({"other": {"params": {}}, autotune.workload_key(g, 1): {"params": params}}, indent=1))
    out = tmp_path / "dist"
    env = dict(os.environ, ADDS_TUNE_FILE=str(tune))
    r = subprocess.run(f"{os.path.join(BIN, 'sssp_' + kind)} -s 3 -o {out} {path} 3>/dev/null", shell=True,
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    assert f"tuned parameters from {tune}" in r.stdout
    ref = oracle.dijkstra(g.row_ptr, g.col, g.w, 3, g.is_float)
    ref_txt = tmp_path / "ref"
    adds_amd.write_dist_text(ref_txt, ref.view(np.float32) if g.is_float else ref, "float" if g.is_float else "int")
    assert out.read_text() == ref_txt.read_text()
    # a key that is not in the file: defaults, no message
    env["ADDS_TUNE_FILE"] = str(tmp_path / "missing.json")
    r = subprocess.run(f"{os.path.join(BIN, 'sssp_' + kind)} -s 3 {path} 3>/dev/null", shell=True,
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0 and "tuned parameters" not in r.stdout


def test_cli_rejects_out_of_range_source(tmp_path):
    r = subprocess.run([os.path.join(BIN, "sssp_int"), "-s", "999999", os.path.join(GOLDEN, "rmat8_int.gr")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "out of range" in r.stderr


def test_convert_then_solve(tmp_path):
    """A SuiteSparse-style MatrixMarket file (symmetric, negative values) -> adds_convert -> .gr -> the
    drop-in float binary and the near-far binary: both write the oracle's distance text for the graph the
    converter made."""
    import numpy as np

    import adds_amd
    import oracle

    rng = np.random.default_rng(12)
    n, nnz = 3000, 12000
    i = rng.integers(1, n + 1, nnz)
    j = rng.integers(1, n + 1, nnz)
    keep = i >= j  # lower triangle, as symmetric files store it
    v = rng.normal(0.0, 2.0, nnz)
    mtx = tmp_path / "g.mtx"
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{n} {n} {int(keep.sum())}\n")
        for a, b, x in zip(i[keep], j[keep], v[keep]):
            f.write(f"{a} {b} {x:.6g}\n")
    gr = tmp_path / "g.gr"
    r = subprocess.run([os.path.join(BIN, "adds_convert"), "mtx", "float", str(mtx), str(gr)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    h = adds_amd.HostGraph.read_gr(str(gr), "float")
    ref = tmp_path / "ref.txt"
    adds_amd.write_dist_text(str(ref), oracle.dijkstra(h.row_ptr, h.col, h.w, 0, True), "float")
    for exe in ("sssp_float", "sssp_nf_float"):
        out = tmp_path / f"{exe}.txt"
        r = subprocess.run(f"{os.path.join(BIN, exe)} -s 0 -o {out} {gr} 3>/dev/null", shell=True,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert out.read_text() == ref.read_text(), exe
