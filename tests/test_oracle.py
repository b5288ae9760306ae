"""Pins the CPU oracle (oracle/mrcpu.c) before anything is compared against it.

Checks, per function:
  * every known-answer vector transcribed from the reference's tests
    (tests/golden/reference_vectors.json, file:line per case);
  * brute-force enumeration of all simple paths on tiny grids (the canonical
    path is the lexicographically smallest reversed sequence among the
    (cost, hops)-optimal paths);
  * two independent algorithms agreeing: the A* restatement (label-correcting,
    early stop, optional heuristic) and the complete lexicographic Dijkstra
    field plus backward reconstruction;
  * closed forms and invariants (sweep == bfs on all 3x3 masks, multi-source
    = per-cell minimum, open-grid Manhattan distances).
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import mrcpu
from cases import (MAX, edges_for_case, grid_from_case, isolated_stress_case, load_vectors,
                   weighted_stress, xorshift64)

VEC = load_vectors()


def endpoint_valid(g, src, dst, pads):
    """mr-cpu lib.rs Router validation / provider_alone_paths (negotiated.rs:1396-1407)."""
    for c in (src, dst):
        if c is None or c >= g.n:
            return False
        if g.board is not None and g.board[c] & mrcpu.BOARD_TRACE_NODE:
            return False
        if g.cost[c] == MAX and c not in pads:
            return False
    return True


@pytest.mark.parametrize("case", VEC["cases"], ids=[c["name"] for c in VEC["cases"]])
def test_reference_vector(case):
    g, src, dst = grid_from_case(case)
    pads = case.get("pads", [])
    if case.get("expect_invalid_endpoint"):
        assert not endpoint_valid(g, src, dst, pads)
        return
    assert endpoint_valid(g, src, dst, pads)
    if case["mode"] == "lee":
        spec = mrcpu.Spec(g, mode=mrcpu.MODE_LEE, pads=pads)
    else:
        edges = edges_for_case(case, g)
        spec = mrcpu.Spec(g, edges, pads=pads)
    got = mrcpu.astar(spec, src, dst)
    if case.get("expect_unrouted"):
        assert got is None
        return
    assert got is not None, case["name"]
    path, cost = got
    assert path[0] == src and path[-1] == dst
    if "expect_path" in case:
        assert path == case["expect_path"]
    if "expect_cost" in case:
        assert cost == case["expect_cost"]
    if "expect_path_len" in case:
        assert len(path) == case["expect_path_len"]
    if "expect_search_cost" in case:
        assert cost == case["expect_search_cost"]
    if "expect_route_cost" in case:
        assert mrcpu.grid_path_cost(g, pads, path) == case["expect_route_cost"]
    if "expect_touches_layers" in case:
        plane = g.w * g.h
        assert sorted({c // plane for c in path}) == case["expect_touches_layers"]


def test_reference_edge_prices():
    for c in VEC["edge_price_cases"]:
        assert mrcpu.axis_edge_costs(c["x_lines"], c["w"]) == c["expect_x"]
        assert mrcpu.axis_edge_costs(c["y_lines"], c["h"]) == c["expect_y"]
        allowed = {tuple(p) for p in c["via_allowed_pairs"]}
        via = [c["via_step_cost"] if (k, k + 1) in allowed else None for k in range(c["layers"] - 1)]
        assert via == c["expect_via"]


def test_edge_cost_rounding_closed_forms():
    assert mrcpu.edge_cost(1.0) == 16
    assert mrcpu.edge_cost(0.5 / 16) == 1          # half away from zero
    assert mrcpu.edge_cost(0.49 / 16) == 0
    assert mrcpu.edge_cost(-3.0) == 0
    assert mrcpu.edge_cost(float("nan")) == 0
    assert mrcpu.edge_cost(float("inf")) == 0
    assert mrcpu.edge_cost(1e12) == MAX
    assert mrcpu.axis_edge_costs([0.0, 0.2, 0.7, 1.9], 4) == [3, 8, 19]


def test_net_window_margin_rule():
    g = mrcpu.Grid.filled(200, 100, 1, 1)
    # span 10 -> margin max(16, ceil(3)) = 16, clamped at the board
    assert mrcpu.net_window(g, g.idx3(5, 5), g.idx3(15, 8)) == (0, 0, 31, 24)
    # span 100 -> margin ceil(30) = 30
    assert mrcpu.net_window(g, g.idx3(50, 40), g.idx3(150, 45)) == (20, 10, 180, 75)
    # span 101 -> margin ceil(30.3) = 31
    assert mrcpu.net_window(g, g.idx3(50, 40), g.idx3(151, 45)) == (19, 9, 182, 76)
    # pads extend the box; layer is ignored
    g3 = mrcpu.Grid.filled(200, 100, 2, 1)
    assert mrcpu.net_window(g3, g3.idx3(100, 50, 1), g3.idx3(100, 50, 0), [g3.idx3(120, 50, 1)]) == \
        (84, 34, 136, 66)


@pytest.mark.parametrize("case", VEC["field_cases"], ids=[c["name"] for c in VEC["field_cases"]])
def test_reference_field_vectors(case):
    g, src, dst = grid_from_case(case)
    bfs = mrcpu.bfs_distance_field(g, src)
    sw = mrcpu.sweep_distance_field(g, src)
    assert np.array_equal(bfs, sw)
    if case.get("expect_all_max"):
        assert (bfs == MAX).all()
        return
    for i, v in case.get("expect_at", []):
        assert int(bfs[i]) == v
    if "expect_path" in case:
        assert mrcpu.path_from_field(g, sw, src, dst) == case["expect_path"]


def test_sweep_equals_bfs_and_lee_on_every_3x3_mask():
    """sweep.rs:302-330: exhaustive over obstacle masks and endpoint pairs."""
    w = h = 3
    for mask in range(1 << 9):
        cost = np.array([MAX if mask >> c & 1 else 1 for c in range(9)], dtype=np.uint32)
        g = mrcpu.Grid(w, h, 1, cost)
        lee = mrcpu.Spec(g, mode=mrcpu.MODE_LEE)
        buf = mrcpu.SearchBuf(9)
        for src in range(9):
            if cost[src] == MAX:
                continue
            bfs = mrcpu.bfs_distance_field(g, src)
            assert np.array_equal(bfs, mrcpu.sweep_distance_field(g, src))
            for dst in range(9):
                if cost[dst] == MAX:
                    continue
                got = mrcpu.path_from_field(g, bfs, src, dst)
                ref = mrcpu.astar(lee, src, dst, buf)
                assert got == (ref[0] if ref else None), (mask, src, dst)


# ---- brute force ------------------------------------------------------------

def _neighbors(g, c):
    plane = g.w * g.h
    l, r = divmod(c, plane)
    y, x = divmod(r, g.w)
    out = []
    if y > 0: out.append(c - g.w)
    if x > 0: out.append(c - 1)
    if x + 1 < g.w: out.append(c + 1)
    if y + 1 < g.h: out.append(c + g.w)
    if l > 0: out.append(c - plane)
    if l + 1 < g.layers: out.append(c + plane)
    return out


def brute_force(spec, g, src, dst, blocked):
    """Lexicographic min over all simple paths of (cost, hops, reversed path)."""
    best = None
    stack = [(src, [src], 0)]
    while stack:
        c, path, cost = stack.pop()
        if c == dst:
            key = (cost, len(path) - 1, tuple(reversed(path)))
            if best is None or key < best:
                best = key
            continue
        for v in _neighbors(g, c):
            if v in path or blocked(v):
                continue
            s = mrcpu.step_cost(spec, c, v)
            if s == MAX:
                continue
            nc = cost + s
            if nc >= MAX:
                continue
            stack.append((v, path + [v], nc))
    if best is None:
        return None
    return list(reversed(best[2])), best[0]


def _random_tiny(seed, w, h, L, zero_frac=0.2, obst_frac=0.2):
    rng = np.random.default_rng(seed)
    n = w * h * L
    cost = rng.integers(1, 6, size=n).astype(np.uint32)
    cost[rng.random(n) < zero_frac] = 0
    cost[rng.random(n) < obst_frac] = MAX
    board = None
    vf = None
    if seed % 3 == 0:
        board = (rng.integers(0, 64, size=n) & rng.integers(0, 64, size=n)).astype(np.uint8)
        board &= np.uint8(0x3E)  # keep nodes mostly open; edges and via nodes masked
    if seed % 4 == 1:
        vf = (rng.random(n) < 0.3).astype(np.uint8)
    return mrcpu.Grid(w, h, L, cost, via_forbidden=vf, board=board)


@pytest.mark.parametrize("seed", range(40))
def test_astar_equals_brute_force_on_tiny_grids(seed):
    w, h, L = [(3, 3, 1), (3, 2, 2), (2, 2, 3), (4, 2, 1)][seed % 4]
    g = _random_tiny(seed, w, h, L)
    rng = np.random.default_rng(1000 + seed)
    ex = [int(v) for v in rng.integers(0, 4, size=w - 1)]
    ey = [int(v) for v in rng.integers(0, 4, size=h - 1)]
    via = [None if rng.random() < 0.25 else int(rng.integers(0, 6)) for _ in range(L - 1)]
    edges = mrcpu.Edges(ex, ey, via)
    add = rng.integers(0, 3, size=g.n).astype(np.uint32) if seed % 5 == 0 else None
    for src in range(g.n):
        for dst in range(g.n):
            pads = [c for c in (src, dst) if g.cost[c] == MAX]
            if not endpoint_valid(g, src, dst, pads):
                continue
            vpads = pads[:1] if seed % 2 else []
            spec = mrcpu.Spec(g, edges, pads=pads, via_pads=vpads, additive=add)
            forced = {src, dst}

            def blocked(c):
                if g.board is not None and g.board[c] & 1:
                    return True
                if c in forced:
                    return False
                return g.cost[c] == MAX and c not in pads

            want = brute_force(spec, g, src, dst, blocked)
            got = mrcpu.astar(spec, src, dst)
            assert got == want, (src, dst, got, want)
            # the complete-field algorithm must agree as well
            dist, hops = mrcpu.field_lex(spec, [src], [src, dst])
            rec = mrcpu.reconstruct(spec, dist, hops, dst)
            assert rec == (None if want is None else want[0]), (src, dst)


def test_brute_force_plants_catch_mistakes():
    """The brute-force enumeration itself distinguishes the tie-break variants."""
    g = mrcpu.Grid.filled(2, 2, 1, 1)
    spec = mrcpu.Spec(g, mrcpu.Edges([16], [16], []))
    want = brute_force(spec, g, 0, 3, lambda c: False)
    assert want == ([0, 1, 3], 32)          # lower predecessor of 3 is 1, not 2
    # minimum hops beats lower predecessor (mr-metal lib.rs:1811-1843)
    g2 = mrcpu.Grid(3, 2, 1, [1, 1, 1, 1, 3, 1])
    spec2 = mrcpu.Spec(g2, mrcpu.Edges([16, 16], [16], []))
    assert brute_force(spec2, g2, 3, 5, lambda c: False)[0] == [3, 4, 5]


# ---- two independent algorithms on the reference stress batteries ------------

def test_weighted_stress_lee_field_and_paths():
    g, sources, nets = weighted_stress()
    lee = mrcpu.Spec(g, mode=mrcpu.MODE_LEE)
    buf = mrcpu.SearchBuf(g.n)
    for s in sources:
        bfs = mrcpu.bfs_distance_field(g, s)
        dist, hops = mrcpu.field_lex(lee, [s])
        # planar-only Lee fields equal the bfs field on the source layer
        assert np.array_equal(dist, bfs)
    for (s, d) in nets:
        ref = mrcpu.astar(lee, s, d, buf)
        dist, hops = mrcpu.field_lex(lee, [s])
        rec = mrcpu.reconstruct(lee, dist, hops, d)
        assert rec == (ref[0] if ref else None)
        if ref:
            assert int(dist[d]) == ref[1]


@pytest.mark.parametrize("case", range(12))
def test_isolated_stress_astar_equals_field_reconstruction(case):
    g, nets, edges = isolated_stress_case(case)
    has_zero = bool((g.cost == 0).any())
    for net in nets:
        base = dict(pads=net.pads, via_pads=net.via_pads)
        win = mrcpu.net_window(g, net.src, net.dst, net.pads)
        for window in (win, None):
            s0 = mrcpu.Spec(g, edges, window=window, **base)
            s1 = mrcpu.Spec(g, edges, window=window, use_heuristic=not has_zero,
                            min_via_cost=min([v for v in edges.via if v is not None], default=0), **base)
            a0 = mrcpu.astar(s0, net.src, net.dst)
            a1 = mrcpu.astar(s1, net.src, net.dst)
            assert a0 == a1
            dist, hops = mrcpu.field_lex(s0, [net.src], [net.src, net.dst])
            rec = mrcpu.reconstruct(s0, dist, hops, net.dst)
            assert rec == (a0[0] if a0 else None)
            if a0:
                assert int(dist[net.dst]) == a0[1]
                assert mrcpu.provider_path_is_valid(g, net, edges, a0[0])


def test_multi_source_field_is_per_cell_minimum():
    rng = np.random.default_rng(7)
    w, h, L = 12, 9, 3
    n = w * h * L
    cost = rng.integers(1, 9, size=n).astype(np.uint32)
    cost[rng.random(n) < 0.2] = MAX
    g = mrcpu.Grid(w, h, L, cost)
    edges = mrcpu.Edges([10] * (w - 1), [10] * (h - 1), [50, 50])
    spec = mrcpu.Spec(g, edges)
    free = [int(c) for c in np.flatnonzero(cost != MAX)]
    srcs = [free[3], free[40], free[77], free[150]]
    dist, hops = mrcpu.field_lex(spec, srcs)
    singles = [mrcpu.field_lex(spec, [s]) for s in srcs]
    D = np.stack([s[0] for s in singles]).astype(np.uint64)
    H = np.stack([s[1] for s in singles]).astype(np.uint64)
    dmin = D.min(axis=0)
    hmin = np.where(D == dmin, H, np.uint64(MAX)).min(axis=0)
    assert np.array_equal(dist.astype(np.uint64), dmin)
    assert np.array_equal(hops.astype(np.uint64), hmin)


def test_open_grid_closed_form():
    w, h, L = 9, 7, 3
    g = mrcpu.Grid.filled(w, h, L, 1)
    E, Ev = 10, 50
    spec = mrcpu.Spec(g, mrcpu.Edges([E] * (w - 1), [E] * (h - 1), [Ev] * (L - 1)))
    src = g.idx3(2, 3, 1)
    dist, hops = mrcpu.field_lex(spec, [src])
    for l in range(L):
        for y in range(h):
            for x in range(w):
                m = abs(x - 2) + abs(y - 3)
                assert dist[g.idx3(x, y, l)] == E * m + Ev * abs(l - 1)
                assert hops[g.idx3(x, y, l)] == m + abs(l - 1)


def test_extreme_via_into_zero_weight_costs_zero():
    """mr-metal lib.rs:1792-1806: a MAX-priced legal via into weight 0 is free."""
    g = mrcpu.Grid(1, 1, 3, [1, 0, 1])
    spec = mrcpu.Spec(g, mrcpu.Edges([], [], [MAX, MAX]))
    assert mrcpu.step_cost(spec, 0, 1) == 0
    assert mrcpu.step_cost(spec, 1, 2) == MAX - 1
    spec_none = mrcpu.Spec(g, mrcpu.Edges([], [], [None, MAX]))
    assert mrcpu.step_cost(spec_none, 0, 1) == MAX   # forbidden, distinct from a zero price


def test_provider_validation_rejects_illegal_paths():
    """Cases of negotiated.rs:4486-4517 (non-canonical / illegal provider paths)."""
    g = mrcpu.Grid.filled(3, 2, 2, 1)
    cost = g.cost.copy()
    cost[g.idx3(1, 0, 0)] = MAX
    g = mrcpu.Grid(3, 2, 2, cost, via_forbidden=np.eye(1, 12, g.idx3(2, 1, 0), dtype=np.uint8)[0])
    e = mrcpu.Edges([16, 16], [16], [160])
    net = mrcpu.Net(g.idx3(0, 0, 0), g.idx3(2, 0, 0))
    good = [0, 3, 4, 5, 2]
    assert mrcpu.provider_path_is_valid(g, net, e, good)
    assert not mrcpu.provider_path_is_valid(g, net, e, [0, 1, 2])            # through obstacle
    assert not mrcpu.provider_path_is_valid(g, net, e, [0, 4, 5, 2])         # diagonal
    assert not mrcpu.provider_path_is_valid(g, net, e, [0, 3, 0, 3, 4, 5, 2])  # revisits
    assert not mrcpu.provider_path_is_valid(g, net, e, [3, 4, 5, 2])         # wrong start
    assert not mrcpu.provider_path_is_valid(g, net, e, [])                   # empty
    # via through a via-forbidden cell without an exemption is illegal ...
    via_path = [0, 3, 4, 5, 11, 8]
    net2 = mrcpu.Net(0, 8)
    assert not mrcpu.provider_path_is_valid(g, net2, e, via_path)
    # ... and legal once that cell is a via-exempt own pad
    net3 = mrcpu.Net(0, 8, pads=[5], via_pads=[5])
    assert mrcpu.provider_path_is_valid(g, net3, e, via_path)
    e_none = mrcpu.Edges([16, 16], [16], [None])
    assert not mrcpu.provider_path_is_valid(g, net3, e_none, via_path)


def test_congestion_from_reference_vectors():
    """mr-core/src/lib.rs:856-891 (congestion_counts_overlaps,
    congestion_counts_each_route_once_per_cell)."""
    assert mrcpu.congestion_from((3, 1, 1), [[0, 1], [1, 2]]).tolist() == [1, 2, 1]
    assert mrcpu.congestion_from((4, 1, 1), [[0, 1, 2, 1, 0, 1], [1, 2, 3, 2]]).tolist() == [1, 2, 2, 1]
