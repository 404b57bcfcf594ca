"""The CPU baselines (baselines/: Galois Dijkstra -t 1 and delta-step -t N, restated) reach the oracle's
distances, saturate integer sums like the GPU solver, and count work the way Galois reports it."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import baselines  # noqa: E402
import oracle  # noqa: E402
from test_oracle import rand_graph  # noqa: E402


@pytest.mark.parametrize("is_float", [False, True])
def test_dijkstra_equals_oracle(is_float):
    rng = np.random.default_rng(5)
    row, col, w = rand_graph(rng, 4000, 30000, is_float=is_float, zero_ok=True)
    for src in (0, 17, 3999):
        d, work = baselines.dijkstra(row, col, w, src, is_float)
        ref = oracle.dijkstra(row, col, w, src, is_float)
        assert np.array_equal(d, ref)
        assert work == int(np.count_nonzero(ref != (0x7F800000 if is_float else 0xFFFFFFFF)))


@pytest.mark.parametrize("is_float,delta", [(False, 8.0), (False, 1.0), (True, 0.25), (True, 64.0)])
def test_parallel_delta_step_equals_dijkstra(is_float, delta):
    rng = np.random.default_rng(17)
    row, col, w = rand_graph(rng, 3000, 24000, is_float=is_float, zero_ok=True)
    for src in (0, 1234):
        d = oracle.dijkstra(row, col, w, src, is_float)
        for threads in (1, 4):
            got, work = baselines.delta_step(row, col, w, src, is_float, delta, threads)
            assert np.array_equal(got, d)
            assert work >= int(np.count_nonzero(d != (0x7F800000 if is_float else 0xFFFFFFFF)))


def test_int_sums_saturate():
    """0 -> 1 -> 2 with weights 0xF0000000 and 0x20000000: the second sum passes 2^32-1 and saturates at
    2^32-2, the oracle's saturated solution."""
    row = np.array([0, 1, 2, 2], dtype=np.uint32)
    col = np.array([1, 2], dtype=np.uint32)
    w = np.array([0xF0000000, 0x20000000], dtype=np.uint32)
    ref = oracle.dijkstra(row, col, w, 0, False, saturate=True)
    assert ref[2] == 0xFFFFFFFE
    assert np.array_equal(baselines.dijkstra(row, col, w, 0, False)[0], ref)
    assert np.array_equal(baselines.delta_step(row, col, w, 0, False, 1 << 20, 2)[0], ref)


def test_source_out_of_range():
    row = np.array([0, 1, 1], dtype=np.uint32)
    col = np.array([1], dtype=np.uint32)
    w = np.array([3], dtype=np.uint32)
    with pytest.raises(ValueError):
        baselines.dijkstra(row, col, w, 2, False)
    with pytest.raises(ValueError):
        baselines.delta_step(row, col, w, 5, False, 1.0)
