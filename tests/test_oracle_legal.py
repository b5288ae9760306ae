"""Pins the oracle's legal-search blocks (route_legal_once with committed copper
owners and clearance halo, negotiated.rs:2885-3052, blocks :2937-2970).

  * hand cases on a corridor for the ordering of the rules: foreign copper
    blocks even an endpoint, an endpoint inside a foreign halo stays
    reachable, the group's own copper is exempt from a foreign or mixed halo,
    an own pad inside a foreign halo is blocked, owners below -1 are free;
  * on random tiny boards (masks, vias, pads, windows), the legal search
    equals the plain alone search (pinned by the reference vectors) on the
    same board with every blocked cell turned into a board trace node and the
    endpoint exemption applied by hand, and the complete-field search agrees
    with A* under the blocks.
"""
from __future__ import annotations

import random

import numpy as np

import mrcpu
from cases import MAX


def corridor(n=5):
    g = mrcpu.Grid.filled(n, 1, 1, 1)
    return g, mrcpu.Edges.uniform(n, 1, 1, planar=16, via=160)


def legal(g, e, owner=None, halo=None, group=0, pads=()):
    return mrcpu.Spec(g, e, legal_owner=owner, legal_halo=halo, own_group=group, pads=pads)


def test_block_order_on_a_corridor():
    g, e = corridor()
    free = [-1] * 5
    path = [0, 1, 2, 3, 4]
    assert mrcpu.astar(legal(g, e, free, free), 0, 4) == (path, 64)
    # foreign copper in the middle, or on the target itself
    assert mrcpu.astar(legal(g, e, [-1, -1, 7, -1, -1], free), 0, 4) is None
    assert mrcpu.astar(legal(g, e, [-1, -1, -1, -1, 7], free), 0, 4) is None
    # own copper is passable
    assert mrcpu.astar(legal(g, e, [-1, -1, 0, -1, -1], free), 0, 4) == (path, 64)
    # a foreign halo blocks the corridor but not the endpoints
    assert mrcpu.astar(legal(g, e, free, [-1, -1, 7, -1, -1]), 0, 4) is None
    assert mrcpu.astar(legal(g, e, free, [7, -1, -1, -1, 7]), 0, 4) == (path, 64)
    # mixed halo (-2) is foreign to every group
    assert mrcpu.astar(legal(g, e, free, [-1, -1, -2, -1, -1]), 0, 4) is None
    # ... unless the cell is this group's own copper
    assert mrcpu.astar(legal(g, e, [-1, -1, 0, -1, -1], [-1, -1, -2, -1, -1]), 0, 4) == (path, 64)
    assert mrcpu.astar(legal(g, e, [-1, -1, 0, -1, -1], [-1, -1, 3, -1, -1]), 0, 4) == (path, 64)
    # the group's own halo is free; owner values below -1 are free too
    assert mrcpu.astar(legal(g, e, free, [-1, -1, 0, -1, -1]), 0, 4) == (path, 64)
    assert mrcpu.astar(legal(g, e, [-1, -1, -5, -1, -1], free), 0, 4) == (path, 64)
    # an own pad (obstacle) inside a foreign halo is blocked; outside it is a pad
    cost = np.ones(5, dtype=np.uint32)
    cost[2] = MAX
    g2 = mrcpu.Grid(5, 1, 1, cost)
    assert mrcpu.astar(legal(g2, e, free, free, pads=[2]), 0, 4) == (path, 48 + 16)
    assert mrcpu.astar(legal(g2, e, free, [-1, -1, 4, -1, -1], pads=[2]), 0, 4) is None


def as_trace_nodes(g, src, dst, owner, halo, group):
    """The same board with every cell the legal search blocks made a board
    trace node (a trace node blocks even the endpoints)."""
    board = np.zeros(g.n, dtype=np.uint8) if g.board is None else np.array(g.board, dtype=np.uint8)
    for c in range(g.n):
        foreign_copper = owner[c] >= 0 and owner[c] != group
        foreign_halo = (halo[c] == -2 or (halo[c] >= 0 and halo[c] != group)) and owner[c] != group
        if foreign_copper or (foreign_halo and c not in (src, dst)):
            board[c] |= mrcpu.BOARD_TRACE_NODE
    return mrcpu.Grid(g.w, g.h, g.layers, g.cost, via_forbidden=g.via_forbidden, board=board)


def test_legal_blocks_equal_trace_nodes_random():
    rnd = random.Random(11)
    routed = 0
    for case in range(300):
        w, h, L = rnd.randint(2, 6), rnd.randint(2, 6), rnd.randint(1, 3)
        n = w * h * L
        cost = np.array([MAX if rnd.random() < 0.12 else rnd.choice([1, 1, 2, 5, 17]) for _ in range(n)],
                        dtype=np.uint32)
        board = np.array([mrcpu.BOARD_VIA_NODE if rnd.random() < 0.05 else 0 for _ in range(n)], dtype=np.uint8)
        g = mrcpu.Grid(w, h, L, cost, board=board)
        e = mrcpu.Edges([rnd.randint(1, 40) for _ in range(w - 1)], [rnd.randint(1, 40) for _ in range(h - 1)],
                        [rnd.choice([None, rnd.randint(1, 90)]) for _ in range(L - 1)])
        src, dst = rnd.randrange(n), rnd.randrange(n)
        pads = sorted({c for c in range(n) if cost[c] == MAX and rnd.random() < 0.3} | {
            c for c in (src, dst) if cost[c] == MAX})
        group = rnd.randint(0, 3)
        owner = [rnd.choice([0, 1, 2, 3, -7]) if rnd.random() < 0.25 else -1 for _ in range(n)]
        halo = [rnd.choice([-2, 0, 1, 2, 3]) if rnd.random() < 0.25 else -1 for _ in range(n)]
        for c in pads:  # pads inside halos are the hand case above
            if c not in (src, dst):
                halo[c] = -1
        x0, y0 = rnd.randint(0, w - 1), rnd.randint(0, h - 1)
        win = (0, 0, w - 1, h - 1) if case % 2 else (min(x0, src % (w * h) % w, dst % (w * h) % w),
                                                     min(y0, src % (w * h) // w, dst % (w * h) // w), w - 1, h - 1)
        got = mrcpu.astar(mrcpu.Spec(g, e, pads=pads, window=win, legal_owner=owner, legal_halo=halo,
                                     own_group=group), src, dst)
        ref = mrcpu.astar(mrcpu.Spec(as_trace_nodes(g, src, dst, owner, halo, group), e, pads=pads, window=win),
                          src, dst)
        assert got == ref, case
        spec = mrcpu.Spec(g, e, pads=pads, window=win, legal_owner=owner, legal_halo=halo, own_group=group)
        dist, hops = mrcpu.field_lex(spec, [src], forced=[src, dst])
        r = mrcpu.reconstruct(spec, dist, hops, dst)
        if got is None:
            assert r is None or int(dist[dst]) >= MAX, case
        else:
            routed += 1
            assert r == got[0] and int(dist[dst]) == got[1], case
    assert routed > 80
