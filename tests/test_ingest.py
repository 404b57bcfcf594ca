"""Text graph formats (MatrixMarket coordinate, DIMACS shortest-path) -> host CSR -> .gr, on hand-built
files whose CSR is written out by hand; a DIMACS path checked against its closed-form distances."""
import os
import subprocess

import numpy as np
import pytest

import adds_amd
from adds_amd._lib import AddsError, EIO, EOVERFLOW
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONVERT = os.path.join(ROOT, "adds_amd", "bin", "adds_convert")

MTX_SYM = """%%MatrixMarket matrix coordinate real symmetric
% a comment line
4 4 5
1 1 2.0
2 1 -1.5
3 2 0.25

4 3 3e0
4 1 7
"""
# Mirrored entries follow their original; rows keep file order:
#   (1,1) 0->0;  (2,1) 1->0, 0->1;  (3,2) 2->1, 1->2;  (4,3) 3->2, 2->3;  (4,1) 3->0, 0->3
SYM_ROW = [0, 3, 5, 7, 9]
SYM_COL = [0, 1, 3, 0, 2, 1, 3, 2, 0]
SYM_WF = [2.0, 1.5, 7.0, 1.5, 0.25, 0.25, 3.0, 3.0, 7.0]
SYM_WI = [2, 2, 7, 2, 0, 0, 3, 3, 7]  # nearest integer of |v|, ties to even (1.5 -> 2)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_mtx_symmetric_float_and_int(tmp_path):
    p = _write(tmp_path, "a.mtx", MTX_SYM)
    g = adds_amd.HostGraph.read_mtx(p, "float")
    assert (g.n, g.m) == (4, 9)
    assert g.row_ptr.tolist() == SYM_ROW and g.col.tolist() == SYM_COL
    assert g.w.tolist() == SYM_WF
    gi = adds_amd.HostGraph.read_mtx(p, "int")
    assert gi.col.tolist() == SYM_COL and gi.w.tolist() == SYM_WI


def test_mtx_pattern_general_and_rectangular(tmp_path):
    p = _write(tmp_path, "b.mtx", "%%MatrixMarket matrix coordinate pattern general\n2 3 3\n1 3\n2 1\n1 2\n")
    g = adds_amd.HostGraph.read_mtx(p, "int")
    assert g.n == 3  # max(rows, cols)
    assert g.row_ptr.tolist() == [0, 2, 3, 3]
    assert g.col.tolist() == [2, 1, 0]
    assert g.w.tolist() == [1, 1, 1]


def test_dimacs_path_closed_form(tmp_path):
    p = _write(tmp_path, "c.gr.txt", "c path 1-2-3-4\np sp 4 3\na 1 2 5\nc middle comment\na 2 3 7\na 3 4 1\n")
    g = adds_amd.HostGraph.read_dimacs(p, "int")
    assert g.row_ptr.tolist() == [0, 1, 2, 3, 3] and g.col.tolist() == [1, 2, 3]
    d = oracle.dijkstra(g.row_ptr, g.col, g.w, 0, is_float=False)
    assert d.tolist() == [0, 5, 12, 13]  # prefix sums of the weights
    gf = adds_amd.HostGraph.read_dimacs(p, "float")
    assert gf.w.tolist() == [5.0, 7.0, 1.0]


def test_convert_cli_round_trip(tmp_path):
    p = _write(tmp_path, "a.mtx", MTX_SYM)
    out = str(tmp_path / "a.gr")
    r = subprocess.run([CONVERT, "mtx", "float", p, out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["4", "9"]
    g = adds_amd.HostGraph.read_gr(out, "float")
    assert g.row_ptr.tolist() == SYM_ROW and g.col.tolist() == SYM_COL and g.w.tolist() == SYM_WF
    bad = _write(tmp_path, "bad.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 2 1.0\n")
    r = subprocess.run([CONVERT, "mtx", "int", bad, str(tmp_path / "bad.gr")], capture_output=True, text=True)
    assert r.returncode == 1 and "entry count" in r.stderr


@pytest.mark.parametrize("text,status,what", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", EIO, "coordinate"),
    ("not a header\n", EIO, "header"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", EIO, "index out of range"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 2 3e9\n", EOVERFLOW, "weight"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 2 nan\n", EOVERFLOW, "weight"),
])
def test_mtx_errors(tmp_path, text, status, what):
    p = _write(tmp_path, "e.mtx", text)
    with pytest.raises(AddsError) as ei:
        adds_amd.HostGraph.read_mtx(p, "int")
    assert ei.value.status == status and what in str(ei.value)


@pytest.mark.parametrize("text,what", [
    ("a 1 2 3\np sp 2 1\n", "before the problem line"),
    ("p sp 2 2\na 1 2 3\n", "arc count"),
    ("p sp 2 1\na 1 5 3\n", "out of range"),
    ("p max 2 1\na 1 2 3\n", "problem line"),
])
def test_dimacs_errors(tmp_path, text, what):
    p = _write(tmp_path, "e.txt", text)
    with pytest.raises(AddsError) as ei:
        adds_amd.HostGraph.read_dimacs(p, "int")
    assert ei.value.status == EIO and what in str(ei.value)


def test_float_weights_round_to_nearest(tmp_path):
    p = _write(tmp_path, "f.mtx", "%%MatrixMarket matrix coordinate real general\n1 2 1\n1 2 0.1\n")
    g = adds_amd.HostGraph.read_mtx(p, "float")
    assert g.w[0] == np.float32(0.1)
