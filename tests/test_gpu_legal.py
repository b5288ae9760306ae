"""GPU: legalization searches (mrf_route_legal_batch) against the oracle's
route_legal_once restatement (negotiated.rs:2885-3052 with committed copper
owners and clearance halo, blocks :2937-2970).

Snapshots: half of a board's nets are committed (their alone paths become
copper of their group, radius-1 planar footprints the halo, -2 where two
groups' footprints meet); then every net, committed or not, searches against
that snapshot with its own group, in its net window with a full-board retry.
Paths and costs must equal the oracle's byte for byte.
"""
from __future__ import annotations

import numpy as np
import pytest

import mrcpu
import mrfield
from cases import oracle_parts

pytestmark = pytest.mark.gpu


def snapshot(s, paths, groups, committed):
    g = s.grid
    n = g.n
    plane = g.w * g.h
    owner = np.full(n, -1, dtype=np.int64)
    halo = np.full(n, -1, dtype=np.int64)
    for i in committed:
        if paths[i] is None:
            continue
        gi = groups[i]
        for c in paths[i][0]:
            owner[c] = gi
        for c in paths[i][0]:
            l, r = divmod(c, plane)
            cy, cx = divmod(r, g.w)
            for y in range(max(0, cy - 1), min(g.h, cy + 2)):
                for x in range(max(0, cx - 1), min(g.w, cx + 2)):
                    k = l * plane + y * g.w + x
                    if g.cost[k] == mrfield.MAX:
                        continue
                    halo[k] = gi if halo[k] in (-1, gi) else -2
    return owner, halo


CASES = [(64, 2, 24, 41, (16, 16, 256, 32)), (96, 4, 20, 42, (16, 16, 256, 32)), (72, 3, 16, 43, (8, 4, 64, 32))]


@pytest.mark.parametrize("mode", [mrfield.MODE_PLAIN, mrfield.MODE_QUEUE])
@pytest.mark.parametrize("side,layers,n,seed,tiles", CASES)
def test_legal_batch_matches_oracle(mode, side, layers, n, seed, tiles):
    s = mrfield.synth_board(side, layers, n, seed)
    nets = s.nets()
    groups = [i // 2 for i in range(len(nets))]  # pairs of nets share a group
    eng = mrfield.Engine(0)
    try:
        eng.set_tiles(*tiles)
        eng.set_mode(mode)
        board = mrfield.Board(eng, s.grid, s.edges, None)
        try:
            alone = board.solve(nets)
            owner, halo = snapshot(s, alone, groups, committed=range(0, len(nets), 2))
            og, _ = oracle_parts(s)
            wins = [tuple(mrcpu.net_window(og, x.src, x.dst)) for x in nets]
            got = board.route_legal(nets, owner, halo, groups, windows=wins)
            only_owner = board.route_legal(nets, owner, None, groups, windows=wins)
        finally:
            board.close()
    finally:
        eng.close()
    og, oe = oracle_parts(s)
    for o, h, res in ((owner, halo, got), (owner, None, only_owner)):
        tmpl = mrcpu.Spec(og, oe, legal_owner=o, legal_halo=h)
        onets = [mrcpu.Net(x.src, x.dst, own_group=gi) for x, gi in zip(nets, groups)]
        paths, costs = mrcpu.alone_batch(tmpl, onets, wins, retry_full=True, threads=8)
        assert res == [None if p is None else (p, c) for p, c in zip(paths, costs)]
    # the snapshot changes the answers, and most nets still route
    assert any(a != b for a, b in zip(got, alone)) and sum(p is not None for p in got) >= len(nets) // 3


def test_legal_batch_board_rules():
    s = mrfield.synth_board(32, 2, 6, 44)
    nets = s.nets()
    eng = mrfield.Engine(0)
    try:
        board = mrfield.Board(eng, s.grid, s.edges, None)
        try:
            n = s.grid.n
            free = np.full(n, -1, dtype=np.int64)
            base = board.solve(nets)
            # no owners: the alone search
            assert board.route_legal(nets, free, free, list(range(len(nets)))) == base
            # an additive plane must be dropped first
            board.set_additive(np.ones(n, dtype=np.uint32))
            with pytest.raises(mrfield.MrfError):
                board.route_legal(nets, free, free, list(range(len(nets))))
            board.set_additive(None)
            assert board.route_legal(nets, free, free, list(range(len(nets)))) == base
            # the planes apply to the call only
            assert board.solve(nets) == base
            with pytest.raises(mrfield.MrfError):
                board.route_legal(nets, free[:-1], None, list(range(len(nets))))
        finally:
            board.close()
    finally:
        eng.close()
