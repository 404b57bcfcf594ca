"""CPU tests of the host side of the library: ABI exports, generators, .gr I/O, distance text,
the verify tool and the CLI's argument handling (no GPU calls)."""
import hashlib
import json
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import adds_amd
from adds_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "adds_amd", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    lib = adds_amd.load()
    header = open(os.path.join(ROOT, "include", "adds.h")).read()
    declared = set(re.findall(r"\b(adds_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations parsed"
    for name in sorted(declared | set(L.EXPORTS)):
        assert hasattr(lib, name), f"libadds.so does not export {name}"
    assert declared <= set(L.EXPORTS), f"binding misses {sorted(declared - set(L.EXPORTS))}"


def test_no_device_is_reported_not_crashed():
    # this container has no GPU: device calls must fail with a status code
    st = L._lib().adds_set_device(0)
    if st == L.OK:
        pytest.skip("a GPU is visible")
    assert st == L.ENODEV
    g = adds_amd.HostGraph.rmat(6, 4, seed=1)
    with pytest.raises(adds_amd.AddsError) as ei:
        adds_amd.Graph(g)
    assert ei.value.status == L.ENODEV


def csr_hash(g):
    h = hashlib.sha256()
    for a in (g.row_ptr, g.col, g.w_bits):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def check_simple_symmetric(g, lo=None, hi=None):
    n, row, col = g.n, g.row_ptr.astype(np.int64), g.col.astype(np.int64)
    src = np.repeat(np.arange(n), np.diff(row))
    assert np.all(src != col), "self loop"
    key = src * n + col
    assert np.all(np.diff(key) > 0), "rows not sorted / duplicate edge"
    rkey = col * n + src
    idx = np.searchsorted(key, rkey)
    assert np.all(key[idx] == rkey), "missing reverse edge"
    assert np.array_equal(g.w_bits[idx], g.w_bits), "asymmetric weight"
    if lo is not None:
        assert g.w.min() >= lo and g.w.max() <= hi


def test_generators_simple_symmetric_and_in_range():
    check_simple_symmetric(adds_amd.HostGraph.rmat(10, 8, seed=3, wkind="u32", lo=1, hi=100), 1, 100)
    check_simple_symmetric(adds_amd.HostGraph.rmat(9, 16, seed=3, wkind="f32", lo=0.0, hi=1.0), 0.0, 1.0)
    g = adds_amd.HostGraph.grid(30, 40, 25, seed=2, wkind="f32", lo=1.0, hi=2.0)
    check_simple_symmetric(g, 1.0, 2.0)
    assert g.w.max() < 2.0
    check_simple_symmetric(adds_amd.HostGraph.er(2000, 8, seed=5, wkind="u32", lo=1, hi=1000), 1, 1000)
    gl = adds_amd.HostGraph.grid(20, 20, 0, seed=5, wkind="lognormal", lo=0.5, hi=0.0)
    check_simple_symmetric(gl)
    assert gl.w.min() > 0


def test_symmetric_declaration_matches_exact_check():
    """The generators declare their graphs symmetric (what enables pull phases); the exact check -- the
    sorted (u, v, w) triples equal the sorted (v, u, w) triples -- agrees, and it rejects a graph with one
    reversed pair's weight changed, one edge removed, or arrays not declared symmetric."""
    gens = [adds_amd.HostGraph.rmat(10, 8, seed=3, wkind="u32", lo=1, hi=100),
            adds_amd.HostGraph.rmat(9, 16, seed=3, wkind="f32", lo=0.0, hi=1.0).reweight(seed=9, wkind="f32"),
            adds_amd.HostGraph.grid(30, 40, 25, seed=2, wkind="f32", lo=1.0, hi=2.0),
            adds_amd.HostGraph.er(2000, 8, seed=5, wkind="u32", lo=1, hi=1000)]
    for g in gens:
        assert g.symmetric and g.check_symmetric()
    g = gens[0]
    w = g.w_bits.copy()
    w[5] += 1  # one direction of a pair
    bad = adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, w, "u32")
    assert not bad.symmetric and not bad.check_symmetric()
    # drop the last edge of vertex 0's row: its reverse is left alone
    r = g.row_ptr.astype(np.int64)
    keep = np.ones(g.m, dtype=bool)
    keep[r[1] - 1] = False
    rp = g.row_ptr.copy()
    rp[1:] -= 1
    assert not adds_amd.HostGraph.from_arrays(rp, g.col[keep], g.w_bits[keep], "u32").check_symmetric()
    assert adds_amd.HostGraph.from_arrays(g.row_ptr, g.col, g.w_bits, "u32").check_symmetric()


def test_grid_shape():
    g = adds_amd.HostGraph.grid(2048, 2048, 0, seed=1)
    assert g.n == 2048 * 2048 and g.m == 16769024  # 4-neighbour directed edges
    g = adds_amd.HostGraph.grid(7, 5, 0, seed=1, wkind="u32", lo=1, hi=1)
    d = g.degrees()
    assert d.max() == 4 and d.min() == 2


def test_rmat16_size_matches_config():
    g = adds_amd.HostGraph.rmat(16, 16, seed=1, wkind="u32", lo=1, hi=100)
    assert g.n == 65536
    assert 1.7e6 < g.m < 2.1e6  # ~2M directed edges after symmetrisation and deduplication


def test_generator_determinism_golden_hashes():
    with open(os.path.join(GOLDEN, "gen_hashes.json")) as f:
        exp = json.load(f)
    got = {
        "rmat10_ef8_u32": csr_hash(adds_amd.HostGraph.rmat(10, 8, seed=3, wkind="u32", lo=1, hi=100)),
        "grid30x40_sc25_f32": csr_hash(adds_amd.HostGraph.grid(30, 40, 25, seed=2, wkind="f32", lo=1.0, hi=2.0)),
        "er2000_u32": csr_hash(adds_amd.HostGraph.er(2000, 8, seed=5, wkind="u32", lo=1, hi=1000)),
    }
    assert got == exp


def test_reweight_equals_generation_with_those_weights():
    a = adds_amd.HostGraph.rmat(9, 8, seed=4, wkind="u32", lo=1, hi=65535)
    b = adds_amd.HostGraph.rmat(9, 8, seed=4, wkind="f32", lo=0.0, hi=1.0)
    a.reweight(seed=4, wkind="f32", lo=0.0, hi=1.0)
    assert a.is_float and np.array_equal(a.w_bits, b.w_bits)


def hand_gr_bytes():
    # 3 vertices: 0->1 (w 5), 0->2 (w 7), 2->1 (w 1); odd edge count -> one pad word
    b = struct.pack("<4Q", 1, 4, 3, 3)
    b += struct.pack("<3Q", 2, 2, 3)              # end offsets
    b += struct.pack("<3I", 1, 2, 1)              # destinations
    b += struct.pack("<I", 0)                     # pad
    b += struct.pack("<3I", 5, 7, 1)              # edge data
    return b


def test_gr_known_answer_and_round_trip(tmp_path):
    p = tmp_path / "tiny.gr"
    p.write_bytes(hand_gr_bytes())
    g = adds_amd.HostGraph.read_gr(p, "int")
    assert g.n == 3 and g.m == 3
    assert list(g.row_ptr) == [0, 2, 2, 3]
    assert list(g.col) == [1, 2, 1]
    assert list(g.w) == [5, 7, 1]
    q = tmp_path / "again.gr"
    g.write_gr(q)
    assert q.read_bytes() == hand_gr_bytes()
    big = adds_amd.HostGraph.rmat(10, 8, seed=9, wkind="f32", lo=0.0, hi=1.0)
    big.write_gr(tmp_path / "big.gr")
    back = adds_amd.HostGraph.read_gr(tmp_path / "big.gr", "float")
    assert np.array_equal(back.row_ptr, big.row_ptr) and np.array_equal(back.col, big.col)
    assert np.array_equal(back.w_bits, big.w_bits)


def test_gr_zero_edge_size_gives_zero_weights(tmp_path):
    b = struct.pack("<4Q", 1, 0, 2, 2) + struct.pack("<2Q", 1, 2) + struct.pack("<2I", 1, 0)
    p = tmp_path / "now.gr"
    p.write_bytes(b)
    g = adds_amd.HostGraph.read_gr(p, "int")
    assert list(g.w) == [0, 0]


@pytest.mark.parametrize("mutate", ["version", "truncate", "dst", "offsets"])
def test_gr_malformed_is_rejected(tmp_path, mutate):
    b = bytearray(hand_gr_bytes())
    if mutate == "version":
        b[0] = 2
    elif mutate == "truncate":
        b = b[:-6]
    elif mutate == "dst":
        struct.pack_into("<I", b, 32 + 24, 9)
    elif mutate == "offsets":
        struct.pack_into("<Q", b, 32, 3)   # first end offset larger than the second
    p = tmp_path / "bad.gr"
    p.write_bytes(bytes(b))
    with pytest.raises(adds_amd.AddsError) as ei:
        adds_amd.HostGraph.read_gr(p, "int")
    assert ei.value.status in (L.EINVAL, L.EIO)


def test_missing_file_is_eio(tmp_path):
    with pytest.raises(adds_amd.AddsError) as ei:
        adds_amd.HostGraph.read_gr(tmp_path / "nope.gr", "int")
    assert ei.value.status == L.EIO


def test_distance_text_format(tmp_path):
    p = tmp_path / "d_int"
    adds_amd.write_dist_text(p, np.array([0, 7, L.INF_U32, 2147483647], dtype=np.uint32), "int")
    assert p.read_text() == "0 0\n1 7\n2 INF\n3 2147483647\n"
    f = np.array([0.0, 1.5, np.inf, 1.0 / 3.0], dtype=np.float32)
    q = tmp_path / "d_float"
    adds_amd.write_dist_text(q, f, "float")
    assert q.read_text() == "0 0.000000\n1 1.500000\n2 INF\n3 0.333333\n"


def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, **kw)


def test_verify_tool_semantics(tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir(), b.mkdir()
    (a / "g1").write_text("0 0\n1 5\n")
    (b / "g1").write_text("0 0\n1 5\n")
    (a / "g2").write_text("0 0\n1 5\n")
    (b / "g2").write_text("0 0\n1 6\n")
    (a / "g3").write_text("0 0\n1 5\n2 INF\n")
    (b / "g3").write_text("0 0\n1 5\n")
    lst = tmp_path / "list"
    lst.write_text("g1\t0\n")
    r = run([os.path.join(BIN, "adds_verify"), f"{a}/", f"{b}/", str(lst)])
    assert r.returncode == 0 and r.stdout.strip().endswith("all results match!!!")
    for name, msg in (("g2", "not match"), ("g3", "path1 length not match"), ("g4", "File1 not found")):
        lst.write_text(f"g1\t0\n{name}\t3 \n")
        r = run([os.path.join(BIN, "adds_verify"), f"{a}/", f"{b}/", str(lst)])
        assert r.returncode == 1 and msg in r.stdout and r.stdout.strip().endswith("result mismatch!!!")


def test_cli_argument_handling(tmp_path):
    exe = os.path.join(BIN, "sssp_int")
    r = run([exe])
    assert r.returncode == 1 and "usage" in r.stderr
    r = run([exe, "-g", "x1", "g.gr"])
    assert r.returncode == 1 and "Invalid GPU device" in r.stderr
    r = run([exe, "-s", "0"])
    assert r.returncode == 255 and "usage" in r.stderr


def test_golden_fixtures_reproduced_by_oracle():
    import oracle
    with open(os.path.join(GOLDEN, "fixtures.json")) as f:
        fx = json.load(f)
    for name, meta in fx.items():
        g = adds_amd.HostGraph.read_gr(os.path.join(GOLDEN, name + ".gr"), meta["wtype"])
        d = oracle.dijkstra(g.row_ptr, g.col, g.w, meta["src"], g.is_float)
        path = os.path.join(GOLDEN, name + ".dist")
        got = "".join(
            f"{i} INF\n" if x == (L.INF_F32_BITS if g.is_float else L.INF_U32)
            else (f"{i} {np.float32(np.uint32(x).view(np.float32)):f}\n" if g.is_float else f"{i} {x}\n")
            for i, x in enumerate(d))
        assert got == open(path).read(), name
