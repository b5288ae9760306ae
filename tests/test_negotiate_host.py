"""CPU: the negotiation loop's host logic (metal-route_amd/negotiate.py).

  * net_window equals the oracle's restatement of negotiated.rs:568-594 on
    random endpoints and pads;
  * over-use across groups on hand cases (negotiated.rs:1924-1945);
  * the loop driven by the oracle's route_negotiated: both branches (Jacobi
    with fused and term-by-term snapshots, Gauss-Seidel for small boards)
    converge to paths that share no cell across groups, history grows by
    SCALE exactly on the cells over-used in each iteration, the dirty sets
    follow the incremental rule, and reruns are identical.
"""
from __future__ import annotations

import random

import numpy as np
import pytest

import mrcpu
import mrfield
import negotiate as ng
from cases import oracle_negotiation_router


def test_net_window_matches_oracle():
    rnd = random.Random(5)
    for _ in range(400):
        w, h, L = rnd.randint(1, 300), rnd.randint(1, 300), rnd.randint(1, 4)
        n = w * h * L
        src, dst = rnd.randrange(n), rnd.randrange(n)
        pads = [rnd.randrange(n) for _ in range(rnd.randint(0, 3))]
        og = mrcpu.Grid.filled(w, h, L, 1)
        assert ng.net_window(w, h, src, dst, pads) == mrcpu.net_window(og, src, dst, pads)


def test_overuse_across_groups():
    P = lambda *c: np.asarray(c, dtype=np.int64)
    # two nets of different groups share cell 5; nets of one group may overlap
    assert ng._overuse([P(1, 5, 9), P(4, 5, 6)], np.array([0, 1])).tolist() == [5]
    assert ng._overuse([P(1, 5, 9), P(4, 5, 6)], np.array([3, 3])).tolist() == []
    assert ng._overuse([P(1, 2), P(2, 3), P(3, 2)], np.array([0, 0, 1])).tolist() == [2, 3]
    assert ng._overuse([P(), P(7)], np.array([0, 1])).tolist() == []


def run_loop(s, nets, groups=None, max_iters=ng.MAX_ITERS):
    g = s.grid
    wins = [ng.net_window(g.w, g.h, n.src, n.dst, n.pads) for n in nets]
    route = oracle_negotiation_router(s, nets, wins)
    return ng.negotiate(route, g.n, len(nets), groups, max_iters, keep_trace=True)


def check_trace(res, n_nets, groups):
    grp = list(groups) if groups is not None else list(range(n_nets))
    hist = np.zeros_like(res.history, dtype=np.int64)
    prev_over = set()
    for k, t in enumerate(res.trace):
        assert t.pfac == k + 1
        # history grows by SCALE exactly on this iteration's over-used cells
        hist[t.overused_cells] += ng.SCALE
        # over-used cells are those held by two groups
        owners = {}
        for i, p in enumerate(t.paths):
            for c in p:
                owners.setdefault(c, set()).add(grp[i])
        assert sorted(c for c, o in owners.items() if len(o) > 1) == t.overused_cells
        if n_nets > ng.PARALLEL_NEGOTIATION_THRESHOLD and k > 0:
            prev_paths = res.trace[k - 1].paths
            want = [i for i in range(n_nets) if not prev_paths[i] or prev_over & set(prev_paths[i])]
            assert t.dirty == want
            assert t.fused == (len(want) >= ng.FUSED_JACOBI_MIN_DIRTY_NETS)
        prev_over = set(t.overused_cells)
    assert np.array_equal(hist, res.history.astype(np.int64))
    if res.converged:
        assert not res.trace[-1].any_overuse


@pytest.mark.parametrize("side,layers,n,seed", [(40, 1, 40, 3), (32, 1, 20, 4), (24, 2, 12, 5)])
def test_negotiation_loop_with_oracle_search(side, layers, n, seed):
    s = mrfield.synth_board(side, layers, n, seed)
    nets = s.nets()
    res = run_loop(s, nets)
    check_trace(res, len(nets), None)
    assert res.iterations > 1  # the fixture has contention
    again = run_loop(s, nets)
    assert again.paths == res.paths and again.iterations == res.iterations


def test_negotiation_groups_may_share_cells():
    s = mrfield.synth_board(32, 1, 20, 4)
    nets = s.nets()
    groups = [i // 2 for i in range(len(nets))]
    res = run_loop(s, nets, groups)
    check_trace(res, len(nets), groups)


def test_negotiation_converges_to_disjoint_paths():
    s = mrfield.synth_board(64, 4, 24, 11)
    nets = s.nets()
    res = run_loop(s, nets)
    check_trace(res, len(nets), None)
    assert res.converged and res.iterations > 2
    seen = set()
    for p in res.paths:
        assert p and not (seen & set(p))
        seen |= set(p)


def test_settle_matches_numpy_statement():
    """mrf_negotiation_settle (host code of the library, no device) against
    the numpy statement of the same step: present = paths per cell, over-use
    across groups, saturating history bump, per-net touches."""
    import mrfield

    rng = np.random.default_rng(5)
    for case in range(60):
        n_cells = int(rng.integers(1, 400))
        n_nets = int(rng.integers(0, 40))
        paths = []
        for _ in range(n_nets):
            k = int(rng.integers(0, min(n_cells, 30) + 1))
            paths.append(rng.choice(n_cells, size=k, replace=False).astype(np.int64))
        groups = (rng.integers(0, max(1, n_nets // 2 + 1), size=n_nets) if case % 2
                  else np.arange(n_nets)).astype(np.int64)
        if case % 7 == 3 and n_nets:
            groups[0] = -1  # any int64 group id
        history = rng.integers(0, 2**32 - 1, size=n_cells, dtype=np.uint64).astype(np.uint32)
        history[: n_cells // 3] = 2**32 - 1 - rng.integers(0, 20, size=n_cells // 3).astype(np.uint32)
        present = rng.integers(0, 9, size=n_cells).astype(np.uint32)  # stale values are overwritten
        off, cells = ng._flat(paths)
        h2 = history.copy()
        over, touches = mrfield.negotiation_settle(off, cells, groups, ng.SCALE, present, h2)
        cnt = np.bincount(np.concatenate(paths) if paths and off[-1] else np.zeros(0, dtype=np.int64),
                          minlength=n_cells)
        assert np.array_equal(present, cnt.astype(np.uint32)), case
        # the numpy statement packs (cell, group) pairs with non-negative
        # groups: shifting the ids keeps which nets share a group
        ref_over = ng._overuse(paths, groups - groups.min()) if n_nets else np.zeros(0, dtype=np.int64)
        assert over.tolist() == ref_over.tolist(), case
        want = history.astype(np.int64)
        want[ref_over] = np.minimum(want[ref_over] + ng.SCALE, 2**32 - 1)
        assert np.array_equal(h2, want.astype(np.uint32)), case
        mask = np.zeros(n_cells, dtype=bool)
        mask[ref_over] = True
        ref_touch = ng._touches(paths, mask) if n_nets else np.zeros(0, dtype=bool)
        assert touches.tolist() == ref_touch.tolist(), case
    # malformed input fails closed
    with pytest.raises(mrfield.MrfError):
        mrfield.negotiation_settle(np.array([0, 2], dtype=np.uint64), np.array([1, 99], dtype=np.uint32),
                                   np.array([0]), 16, np.zeros(10, dtype=np.uint32), np.zeros(10, dtype=np.uint32))
