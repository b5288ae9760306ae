"""GPU: the allocation-free entry point (mrf_solve_noalloc) against mrf_solve.

mrf_solve_noalloc carves every device buffer of a solve from the library's
fixed scratch array instead of allocating; its results must equal mrf_solve's
(the existing entry point, itself bit-exact against the oracle) on two calls
in a row, the second with a smaller input, on a board whose batch fits one
chunk and on one whose labels need several chunks of the array; a request
beyond the array's size must fail with a clear error, after which the entry
point still works.
"""
from __future__ import annotations

import numpy as np
import pytest

import mrfield

pytestmark = pytest.mark.gpu


def same(a: mrfield.PathBatch, b: mrfield.PathBatch):
    assert np.array_equal(a.routed, b.routed)
    assert np.array_equal(a.cost[a.routed != 0], b.cost[b.routed != 0])
    assert np.array_equal(a.offsets, b.offsets)
    assert np.array_equal(a.cells, b.cells)


@pytest.mark.parametrize("config", [1, 2])
def test_noalloc_equals_solve(config):
    print(f"noalloc config {config}", flush=True)
    eng = mrfield.Engine(0)
    try:
        eng.set_tiles(16, 16, 256, 32)
        s = mrfield.synth_config(config, seed=5)
        board = mrfield.Board(eng, s.grid, s.edges, s.pricing())
        try:
            nets = s.nets()[:48] if config == 1 else s.nets()[:24]
            for sub in (nets, nets[: len(nets) // 3]):
                want = board.solve_arrays(sub)
                got = board.solve_noalloc(sub)
                same(got, want)
                assert int(np.count_nonzero(got.routed)) > 0
        finally:
            board.close()
    finally:
        eng.close()


def test_noalloc_rejects_a_field_beyond_the_array():
    """One complete field of a 2048x2048x2 board needs 64 MiB of labels, more
    than the array holds: a clear error, then the entry point still works."""
    print("noalloc limit", flush=True)
    eng = mrfield.Engine(0)
    try:
        s = mrfield.synth_config(3, seed=1)
        board = mrfield.Board(eng, s.grid, s.edges, s.pricing())
        try:
            nets = s.nets()[:1]
            with pytest.raises(mrfield.MrfError) as x:
                board.solve_noalloc(nets)
            assert x.value.code == mrfield.BACKEND_UNAVAILABLE and "scratch array" in str(x.value)
            small = s.nets()[:4]
            wins = [tuple(int(v) for v in w) for w in s.windows[:4]]
            same(board.solve_noalloc(small, windows=wins, retry_full=False),
                 board.solve_arrays(small, windows=wins, retry_full=False))
        finally:
            board.close()
    finally:
        eng.close()
