"""GPU parity of the one-wavefront-per-tile form against the CPU oracle.

The wave form (tiles 16 x 16, 64 threads; two wavefronts per tile on boards
of five to eight layers) runs for batches without pads, overrides or legal
planes on boards without a weight plane whose step prices fit 15 bits; these
cases are built to qualify, and every one asserts that the wave kernel is
what ran.  Boards
whose every open step costs at least 1 run it on packed u32 labels
(dist << 16 | hops), certified after the solve and solved again on u64
labels when a distance comes near the 16-bit limit; the cases assert which
label width the result came from.  Integer work: dist and hops of every
cell, paths and search costs, bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import mrcpu
import mrfield
from cases import MAX

pytestmark = pytest.mark.gpu

WAVE = (16, 16, 64, 32)


@pytest.fixture(scope="module")
def eng():
    e = mrfield.Engine(0)
    e.set_tiles(*WAVE)
    e.set_mode(mrfield.MODE_PLAIN)
    yield e
    e.close()


def unit_board(seed, w, h, L):
    """Unit enter weights (1 or obstacle), trace nodes, one-way planar edges
    and via-forbidden cells."""
    rng = np.random.default_rng(seed)
    n = w * h * L
    cost = np.ones(n, dtype=np.uint32)
    cost[rng.random(n) < 0.2] = MAX
    board = np.zeros(n, dtype=np.uint8)
    r = rng.random(n)
    board[r < 0.04] |= 1
    board[(r >= 0.04) & (r < 0.09)] |= 2
    for bit in (4, 8, 16, 32):
        board[rng.random(n) < 0.04] |= bit
    vf = (rng.random(n) < 0.1).astype(np.uint8)
    return mrcpu.Grid(w, h, L, cost, via_forbidden=vf, board=board)


def free_cells(g):
    ok = (g.cost != MAX) & ((g.board & 3) == 0)
    return np.flatnonzero(ok)


def as_dev(g):
    return mrfield.Grid(g.w, g.h, g.layers, g.cost, g.via_forbidden, g.board)


def assert_wave_ran(eng, label_bytes=None):
    st = eng.stats()
    assert st["relax_launches"] > 0
    assert st["wave_launches"] == st["relax_launches"], st
    if label_bytes is not None:
        assert st["label_bytes"] == label_bytes, st


def lean_fields_case(L, seed, packed):
    """Multi-source fields on boards with every lean feature: trace nodes,
    obstacles, one-way edges, via-forbidden cells, per-gap x / y prices,
    vias off between some layers, an additive plane; full and partial
    windows (ragged edge tiles).  Packed: every price at least 1; wide: one
    zero-price gap (a zero-cost step rules the packed labels out).  Returns
    the engine call's arguments and the oracle's fields."""
    rng = np.random.default_rng(1000 * L + seed + (7 if packed else 0))
    w, h = int(rng.integers(20, 70)), int(rng.integers(20, 70))
    g = unit_board(seed + 10 * L, w, h, L)
    lo = 1 if packed else 0
    ex = [int(v) for v in rng.integers(lo, 30, size=w - 1)]
    ey = [int(v) for v in rng.integers(lo, 30, size=h - 1)]
    if not packed:
        ex[int(rng.integers(0, w - 1))] = 0
    via = [None if rng.random() < 0.25 else int(rng.integers(lo, 60)) for _ in range(L - 1)]
    edges = mrcpu.Edges(ex, ey, via)
    add = rng.integers(0, 40, size=g.n).astype(np.uint32)
    free = free_cells(g)
    sources, dsts, wins = [], [], []
    for i in range(10):
        k = 1 + int(rng.integers(0, 3))
        sources.append([int(c) for c in rng.choice(free, size=k, replace=False)])
        dsts.append(int(rng.choice(free)))
        if i % 3 == 0:
            wins.append((0, 0, w - 1, h - 1))
        else:
            x0, y0 = int(rng.integers(0, w // 2)), int(rng.integers(0, h // 2))
            wins.append((x0, y0, int(rng.integers(x0, w)), int(rng.integers(y0, h))))
    nets = [mrcpu.Net(s[0], d) for s, d in zip(sources, dsts)]
    ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(g, edges, additive=add), nets, sources, wins)
    dedges = mrfield.Edges(list(edges.x), list(edges.y), list(edges.via))
    args = (as_dev(g), dedges, sources)
    kw = dict(dsts=dsts, windows=wins, pricing=mrfield.Pricing(additive=add))
    return args, kw, ref_d, ref_h


@pytest.mark.parametrize("packed", [True, False])
@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 6, 8])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_wave_random_fields(eng, L, seed, packed):
    """The lean cases in in the plain form (one launch per round)."""
    args, kw, ref_d, ref_h = lean_fields_case(L, seed, packed)
    got_d, got_h = eng.fields(*args, **kw)
    assert_wave_ran(eng, 4 if packed else 8)
    for i in range(len(ref_d)):
        assert np.array_equal(got_d[i], ref_d[i]), (i, np.flatnonzero(got_d[i] != ref_d[i])[:10])
        assert np.array_equal(got_h[i], ref_h[i]), (i, np.flatnonzero(got_h[i] != ref_h[i])[:10])


@pytest.fixture(scope="module")
def qeng():
    e = mrfield.Engine(0)
    e.set_tiles(*WAVE)
    e.set_mode(mrfield.MODE_QUEUE)
    yield e
    e.close()


@pytest.mark.parametrize("packed", [True, False])
@pytest.mark.parametrize("L", [1, 2, 4, 5, 8])
@pytest.mark.parametrize("seed", [1, 2])
def test_wave_queue_random_fields(qeng, L, seed, packed):
    """The same cases in the persistent queue form: one launch of the wave
    visit's queue kernel (two wavefronts per visit from five layers), on
    packed and on u64 labels."""
    args, kw, ref_d, ref_h = lean_fields_case(L, seed, packed)
    got_d, got_h = qeng.fields(*args, **kw)
    assert_wave_ran(qeng, 4 if packed else 8)
    assert qeng.stats()["relax_launches"] == 1
    for i in range(len(ref_d)):
        assert np.array_equal(got_d[i], ref_d[i]), (i, np.flatnonzero(got_d[i] != ref_d[i])[:10])
        assert np.array_equal(got_h[i], ref_h[i]), (i, np.flatnonzero(got_h[i] != ref_h[i])[:10])


@pytest.mark.parametrize("L", [3, 6])
@pytest.mark.parametrize("seed", [1, 2])
def test_wave_halo_owner_fields(eng, L, seed):
    """Boards with a halo-owner plane: a cell owned by another net (or by
    several, -2) costs the penalty on top of its additive price, the net's
    own cells do not.  Never packed: u64 labels."""
    rng = np.random.default_rng(500 + 10 * L + seed)
    w, h = int(rng.integers(30, 70)), int(rng.integers(30, 70))
    g = unit_board(seed + 40, w, h, L)
    edges = mrcpu.Edges([int(v) for v in rng.integers(1, 30, size=w - 1)],
                        [int(v) for v in rng.integers(1, 30, size=h - 1)],
                        [int(rng.integers(1, 60)) for _ in range(L - 1)])
    add = rng.integers(1, 40, size=g.n).astype(np.uint32)
    halo = np.full(g.n, -1, dtype=np.int32)
    sel = rng.random(g.n) < 0.15
    halo[sel] = rng.integers(0, 4, size=int(sel.sum()))
    halo[rng.random(g.n) < 0.03] = -2
    free = free_cells(g)
    nets, sources, dsts, wins = [], [], [], []
    for i in range(8):
        src = [int(c) for c in rng.choice(free, size=1 + i % 3, replace=False)]
        d = int(rng.choice(free))
        nets.append(mrcpu.Net(src[0], d, own_id=i % 4))
        sources.append(src)
        dsts.append(d)
        wins.append((0, 0, w - 1, h - 1) if i % 2 else (2, 3, w - 4, h - 2))
    ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(g, edges, additive=add, halo_owner=halo, halo_penalty=13),
                                      nets, sources, wins)
    got_d, got_h = eng.fields(as_dev(g), mrfield.Edges(list(edges.x), list(edges.y), list(edges.via)), sources,
                              dsts=dsts, windows=wins,
                              pricing=mrfield.Pricing(add, halo, 13, [n.own_id for n in nets]))
    assert_wave_ran(eng, 8)
    for i in range(len(nets)):
        assert np.array_equal(got_d[i], ref_d[i]), (i, np.flatnonzero(got_d[i] != ref_d[i])[:10])
        assert np.array_equal(got_h[i], ref_h[i]), (i, np.flatnonzero(got_h[i] != ref_h[i])[:10])


@pytest.mark.parametrize("L", [2, 4, 6])
@pytest.mark.parametrize("seed", [1, 2])
def test_wave_pads_and_via_pads(eng, L, seed):
    """Endpoints on obstacle cells made passable by the net's own pads (a
    through-hole footprint: the endpoint's cell and its planar neighbours on
    every layer), every footprint cell also via-exempt (via_forbidden lifted
    for this net only); through the provider seam, window then full-board
    retry.  The wave form applies pads per tile and prices the halo with
    lower bounds (a border next to a pad only activates more); paths and
    search costs equal the oracle's."""
    rng = np.random.default_rng(700 + 10 * L + seed)
    w, h = int(rng.integers(30, 60)), int(rng.integers(30, 60))
    g = unit_board(seed + 60, w, h, L)
    plane = w * h
    edges = mrcpu.Edges([int(v) for v in rng.integers(1, 25, size=w - 1)],
                        [int(v) for v in rng.integers(1, 25, size=h - 1)],
                        [int(rng.integers(1, 40)) for _ in range(L - 1)])
    free = free_cells(g)
    blocked = np.flatnonzero((g.cost == MAX) & ((g.board & 1) == 0))

    def footprint(c):
        x, y = c % w, (c % plane) // w
        out = set()
        for l in range(L):
            for dx, dy in ((0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)):
                if 0 <= x + dx < w and 0 <= y + dy < h:
                    out.add(l * plane + (y + dy) * w + x + dx)
        return out

    nets = []
    for i in range(10):
        src = int(rng.choice(free))
        dst = int(rng.choice(blocked)) if i % 2 == 0 else int(rng.choice(free))
        pads = sorted(footprint(src) | footprint(dst)) if i % 3 != 2 else []
        vpads = pads if i % 4 != 1 else pads[: len(pads) // 2]
        if i % 2 == 0 and not pads:
            dst = int(rng.choice(free))
        nets.append(mrcpu.Net(src, dst, pads, vpads))
    og = g
    wins = [mrcpu.net_window(og, n.src, n.dst, n.pads) for n in nets]
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, edges), nets, wins, retry_full=True)
    got = eng.route_isolated_batch(as_dev(g), [mrfield.Net(n.src, n.dst, n.pads, n.via_pads) for n in nets], wins,
                                   mrfield.Edges(list(edges.x), list(edges.y), list(edges.via)))
    assert_wave_ran(eng, 4)
    routed = 0
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), i
        if r is not None:
            routed += 1
            assert r[0] == ref_p[i] and r[1] == ref_c[i], i
    assert routed >= 5


@pytest.mark.parametrize("seed", [4, 5])
def test_wave_random_paths(eng, seed):
    """Two-pin nets through the provider seam: window pass with full-board
    retry, paths and search costs."""
    rng = np.random.default_rng(seed)
    w, h, L = 61, 47, 4
    g = unit_board(seed, w, h, L)
    edges = mrcpu.Edges([int(v) for v in rng.integers(1, 25, size=w - 1)],
                        [int(v) for v in rng.integers(1, 25, size=h - 1)], [30, None, 45])
    free = free_cells(g)
    nets = [mrcpu.Net(int(rng.choice(free)), int(rng.choice(free))) for _ in range(24)]
    wins = [mrcpu.net_window(g, n.src, n.dst, n.pads) for n in nets]
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(g, edges), nets, wins, retry_full=True)
    got = eng.route_isolated_batch(as_dev(g), [mrfield.Net(n.src, n.dst) for n in nets], wins,
                                   mrfield.Edges(list(edges.x), list(edges.y), list(edges.via)))
    assert_wave_ran(eng)
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), i
        if r is not None:
            assert r[0] == ref_p[i] and r[1] == ref_c[i], i


def test_wave_config1_complete_fields_and_paths(eng):
    """128x128x2, 64 nets: every field cell and every path."""
    s = mrfield.synth_config(1, seed=1)
    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    nets = [mrcpu.Net(int(a), int(b)) for a, b in zip(s.src, s.dst)]
    ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(og, oe), nets, [[n.src] for n in nets], threads=8)
    got_d, got_h = eng.fields(g, s.edges, [[int(a)] for a in s.src], dsts=[int(b) for b in s.dst])
    assert_wave_ran(eng, 4)
    assert np.array_equal(got_d, ref_d)
    assert np.array_equal(got_h, ref_h)
    full = [(0, 0, g.w - 1, g.h - 1)] * len(nets)
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, oe), nets, full, threads=8)
    got = eng.route_isolated_batch(g, s.nets(), full, s.edges)
    assert_wave_ran(eng)
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), i
        if r is not None:
            assert r[0] == ref_p[i] and r[1] == ref_c[i], i


def test_wave_round_cap_reactivates(eng):
    """A round cap of one in-LDS round: visits end before their tile's
    fixpoint and re-activate themselves; the fields are unchanged."""
    s = mrfield.synth_config(1, seed=2, scale_nets=16)
    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    nets = [mrcpu.Net(int(a), int(b)) for a, b in zip(s.src, s.dst)]
    ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(og, oe), nets, [[n.src] for n in nets], threads=8)
    eng.set_tiles(16, 16, 64, 1)
    try:
        got_d, got_h = eng.fields(g, s.edges, [[int(a)] for a in s.src], dsts=[int(b) for b in s.dst])
        assert_wave_ran(eng)
    finally:
        eng.set_tiles(*WAVE)
    assert np.array_equal(got_d, ref_d)
    assert np.array_equal(got_h, ref_h)


def test_wave_equals_workgroup_form_with_weights_falling_back(eng):
    """A board with a weight plane does not qualify: the wave shape runs the
    workgroup kernel (at 256 threads) and still matches the oracle."""
    rng = np.random.default_rng(9)
    w, h, L = 40, 33, 2
    g = unit_board(9, w, h, L)
    g.cost[g.cost != MAX] = rng.integers(1, 5, size=int((g.cost != MAX).sum())).astype(np.uint32)
    edges = mrcpu.Edges([10] * (w - 1), [10] * (h - 1), [50])
    free = free_cells(g)
    sources = [[int(rng.choice(free))] for _ in range(6)]
    nets = [mrcpu.Net(s[0], MAX) for s in sources]
    ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(g, edges), nets, sources)
    got_d, got_h = eng.fields(as_dev(g), mrfield.Edges(list(edges.x), list(edges.y), list(edges.via)), sources)
    st = eng.stats()
    assert st["relax_launches"] > 0 and st["wave_launches"] == 0
    assert np.array_equal(got_d, ref_d)
    assert np.array_equal(got_h, ref_h)


def test_wave_same_round_neighbours_after_a_layout_change(eng):
    """Label memory laid out by an 8 x 4 solve, then the wave form at 16 x 16
    with a round cap of one (short visits: many neighbours first visited in
    the same round, whose border words still hold the other layout's
    labels): the same paths as the workgroup form."""
    s = mrfield.synth_config(2, seed=1)
    board = mrfield.Board(eng, s.grid, s.edges, s.pricing())
    nets = s.nets()
    try:
        eng.set_tiles(16, 16, 256, 32)
        ref = board.solve_arrays(nets)
        for cap in (1, 32):
            eng.set_tiles(8, 4, 128, 4)
            board.solve_arrays(nets)
            eng.set_tiles(16, 16, 64, cap)
            got = board.solve_arrays(nets)
            assert_wave_ran(eng)
            for k in ("routed", "cost", "offsets", "cells"):
                assert np.array_equal(getattr(got, k), getattr(ref, k)), (cap, k)
    finally:
        board.close()
        eng.set_tiles(*WAVE)


def test_wave_packed_certificate_falls_back_to_wide_labels(eng):
    """Distances past the 16-bit packed limit (gap prices of 700 over a
    96 x 80 board): the certificate rejects the packed solve and the batch is
    solved again on u64 labels; fields and paths equal the oracle's.  The
    same board with prices of 7 stays packed."""
    w, h, L = 96, 80, 2
    g = unit_board(31, w, h, L)
    free = free_cells(g)
    rng = np.random.default_rng(31)
    sources = [[int(rng.choice(free))] for _ in range(5)]
    dsts = [int(rng.choice(free)) for _ in range(5)]
    nets = [mrcpu.Net(s[0], d) for s, d in zip(sources, dsts)]
    for price, width in ((700, 8), (7, 4)):
        edges = mrcpu.Edges([price] * (w - 1), [price] * (h - 1), [3 * price])
        ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(g, edges), nets, sources)
        if width == 8:
            assert ref_d[ref_d != MAX].max() > 0xFFFF  # the case is past the limit
        got_d, got_h = eng.fields(as_dev(g), mrfield.Edges(list(edges.x), list(edges.y), list(edges.via)), sources,
                                  dsts=dsts)
        assert_wave_ran(eng, width)
        assert np.array_equal(got_d, ref_d)
        assert np.array_equal(got_h, ref_h)
        full = [(0, 0, w - 1, h - 1)] * len(nets)
        ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(g, edges), nets, full)
        got = eng.route_isolated_batch(as_dev(g), [mrfield.Net(n.src, n.dst) for n in nets], full,
                                       mrfield.Edges(list(edges.x), list(edges.y), list(edges.via)))
        assert_wave_ran(eng, width)
        for i, r in enumerate(got):
            assert (r is None) == (ref_p[i] is None), (price, i)
            if r is not None:
                assert r[0] == ref_p[i] and r[1] == ref_c[i], (price, i)


@pytest.mark.parametrize("L", [2, 4])
def test_wave_resident_board_additive_updates(eng, L):
    """A resident board's step price planes follow its additive plane: solve,
    replace the plane (held as u32 from then on), solve, drop it, solve; every
    field equals the oracle's under the plane in force, and the lean wave
    kernel ran each time."""
    rng = np.random.default_rng(90 + L)
    w, h = 53, 41
    g = unit_board(77 + L, w, h, L)
    edges = mrcpu.Edges([int(v) for v in rng.integers(1, 30, size=w - 1)],
                        [int(v) for v in rng.integers(1, 30, size=h - 1)],
                        [int(rng.integers(1, 60)) for _ in range(L - 1)])
    dedges = mrfield.Edges(list(edges.x), list(edges.y), list(edges.via))
    free = free_cells(g)
    sources = [[int(c) for c in rng.choice(free, size=1 + i % 2, replace=False)] for i in range(6)]
    dsts = [int(rng.choice(free)) for _ in range(6)]
    nets = [mrcpu.Net(s[0], d) for s, d in zip(sources, dsts)]
    wins = [(0, 0, w - 1, h - 1)] * len(nets)
    planes = [rng.integers(0, 40, size=g.n).astype(np.uint32),
              rng.integers(7_000, 9_000, size=g.n).astype(np.uint32),
              None]
    board = mrfield.Board(eng, as_dev(g), dedges, mrfield.Pricing(additive=planes[0]))
    try:
        for k, add in enumerate(planes):
            if k:
                board.set_additive(add)
            board.solve_fields(sources, dsts=dsts)
            assert_wave_ran(eng)
            spec = mrcpu.Spec(g, edges, additive=add) if add is not None else mrcpu.Spec(g, edges)
            ref_d, ref_h = mrcpu.fields_batch(spec, nets, sources, wins)
            for i in range(len(nets)):
                got_d, got_h = board.read_field(i)
                assert np.array_equal(got_d, ref_d[i]), (k, i, np.flatnonzero(got_d != ref_d[i])[:10])
                assert np.array_equal(got_h, ref_h[i]), (k, i)
    finally:
        board.close()


def test_wave_formed_prices_on_a_large_lean_board(eng):
    """The config-2 board without its history plane through the provider
    entry: its step planes would not fit an XCD's L2, so the lean wave visit
    forms its prices from the masks instead of loading them.  32 nets' paths
    and search costs equal the oracle's alone search."""
    s = mrfield.synth_config(2, seed=1)
    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost, board=g.board)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    idx = list(range(0, 256, 8))
    nets = [mrcpu.Net(int(s.src[i]), int(s.dst[i])) for i in idx]
    full = [(0, 0, g.w - 1, g.h - 1)] * len(nets)
    got = eng.route_isolated_batch(g, [s.nets()[i] for i in idx], full, s.edges)
    assert_wave_ran(eng)
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, oe), nets, full, threads=8)
    for k, r in enumerate(got):
        assert (r is None) == (ref_p[k] is None), idx[k]
        if r is not None:
            assert r == (ref_p[k], ref_c[k]), idx[k]


def test_round_timing_sampled_calls(eng):
    """mrf_set_round_timing(2): the first of every two solving calls brackets
    each round with HIP events (timed_launches = relax_launches), the other
    runs without them; 0 restores the untimed default.  Results never
    change."""
    s = mrfield.synth_config(1, seed=5, scale_nets=8)
    g = s.grid
    full = [(0, 0, g.w - 1, g.h - 1)] * len(s.src)
    try:
        eng.set_round_timing(2)
        seen = []
        for _ in range(4):
            got = eng.route_isolated_batch(g, s.nets(), full, s.edges)
            st = eng.stats()
            assert st["relax_launches"] > 0
            seen.append((got, st["timed_launches"] == st["relax_launches"], st["timed_launches"] == 0))
        assert [t for _, t, _ in seen] == [True, False, True, False]
        assert [u for _, _, u in seen] == [False, True, False, True]
        assert all(got == seen[0][0] for got, _, _ in seen)
        eng.set_round_timing(0)
        assert eng.route_isolated_batch(g, s.nets(), full, s.edges) == seen[0][0]
        assert eng.stats()["timed_launches"] == 0
    finally:
        eng.set_round_timing(0)
