"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Integer work, so the bar is bit-exact: dist and hops labels of every cell,
path cells, and search costs.  Tile shapes are varied so fields span many
tiles (halo exchange, ragged edge tiles), and the solve forms (plain rounds,
hipGraph rounds and the persistent work-queue launch) are checked.
"""
from __future__ import annotations

import numpy as np
import pytest

import mrcpu
import mrfield
from cases import (MAX, grid_from_case, isolated_stress_case, load_vectors, weighted_stress)

pytestmark = pytest.mark.gpu

SHAPES = [(32, 32, 256, 32), (8, 4, 64, 8), (16, 8, 128, 3), (4, 1, 64, 2)]
# The persistent grid-barrier form did not return on MI355X (twice, no fault);
# it is switched off in the library and its cases stay skipped until a cause
# is shown.  The graph form replaces it.
PERSISTENT_SKIP = pytest.param(mrfield.MODE_PERSISTENT, marks=pytest.mark.skip(
    reason="persistent grid-barrier form disabled: did not return on MI355X; cause not yet shown"))
MODES = [mrfield.MODE_PLAIN, PERSISTENT_SKIP, mrfield.MODE_GRAPH, mrfield.MODE_QUEUE]


@pytest.fixture(scope="module")
def eng():
    e = mrfield.Engine(0)
    yield e
    e.close()


def as_dev(g: mrcpu.Grid) -> mrfield.Grid:
    return mrfield.Grid(g.w, g.h, g.layers, g.cost, g.via_forbidden, g.board)


def as_dev_edges(e: mrcpu.Edges) -> mrfield.Edges:
    return mrfield.Edges(list(e.x), list(e.y), list(e.via))


def configure(eng, shape, mode):
    eng.set_tiles(*shape)
    eng.set_mode(mode)


VEC = load_vectors()


@pytest.mark.parametrize("mode", MODES)
def test_reference_vectors(eng, mode):
    configure(eng, (4, 1, 64, 4), mode)
    for case in VEC["cases"]:
        g, src, dst = grid_from_case(case)
        pads = case.get("pads", [])
        net = mrfield.Net(src, dst, pads)
        dg = as_dev(g)
        if case.get("expect_invalid_endpoint"):
            with pytest.raises(mrfield.MrfError) as e:
                eng.router_route(dg, [net])
            assert e.value.code == mrfield.INVALID_ENDPOINT
            continue
        if case["mode"] == "lee":
            got = eng.router_route(dg, [net])[0]
        else:
            via = case.get("via", [160] * (g.layers - 1))
            ed = mrfield.Edges([16] * (g.w - 1), [16] * (g.h - 1), via)
            got = eng.route_isolated_batch(dg, [net], [(0, 0, g.w - 1, g.h - 1)], ed)[0]
        if case.get("expect_unrouted"):
            assert got is None, case["name"]
            continue
        assert got is not None, case["name"]
        path, cost = got
        if "expect_path" in case:
            assert path == case["expect_path"], case["name"]
        if "expect_cost" in case:
            assert cost == case["expect_cost"], case["name"]
        if "expect_search_cost" in case:
            assert cost == case["expect_search_cost"], case["name"]
        if "expect_path_len" in case:
            assert len(path) == case["expect_path_len"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape", SHAPES[:3])
def test_isolated_stress_matches_oracle(eng, mode, shape):
    """mr-metal/src/lib.rs:1882-1933 battery through the provider seam."""
    configure(eng, shape, mode)
    for case in range(12):
        g, nets, edges = isolated_stress_case(case)
        wins = [mrcpu.net_window(g, n.src, n.dst, n.pads) for n in nets]
        ref_paths, ref_costs = mrcpu.alone_batch(mrcpu.Spec(g, edges), nets, wins, retry_full=True)
        dn = [mrfield.Net(n.src, n.dst, n.pads, n.via_pads) for n in nets]
        got = eng.route_isolated_batch(as_dev(g), dn, wins, as_dev_edges(edges))
        for i in range(len(nets)):
            if ref_paths[i] is None:
                assert got[i] is None, (case, i)
            else:
                assert got[i] is not None, (case, i)
                assert got[i][0] == ref_paths[i], (case, i)
                assert got[i][1] == ref_costs[i], (case, i)
        # one pass without retry must report the window-unreachable nets
        ref1, _ = mrcpu.alone_batch(mrcpu.Spec(g, edges), nets, wins, retry_full=False)
        got1 = eng.route_isolated_pass(as_dev(g), dn, wins, as_dev_edges(edges))
        assert [p is None for p in ref1] == [p is None for p in got1]


@pytest.mark.parametrize("mode", MODES)
def test_weighted_stress_fields_and_router(eng, mode):
    """mr-metal/src/lib.rs:1652-1701: sweep fields == bfs, router == Lee."""
    configure(eng, (8, 4, 64, 4), mode)
    g, sources, nets = weighted_stress()
    dg = as_dev(g)
    fields = eng.sweep_fields(dg, sources)
    for s, f in zip(sources, fields):
        assert np.array_equal(f, mrcpu.bfs_distance_field(g, s)), s
    assert np.array_equal(fields[0], eng.sweep_fields(dg, sources[:1])[0])
    got = eng.router_route(dg, [mrfield.Net(s, d) for s, d in nets])
    lee = mrcpu.Spec(g, mode=mrcpu.MODE_LEE)
    for (s, d), r in zip(nets, got):
        ref = mrcpu.astar(lee, s, d)
        assert (r is None) == (ref is None)
        if ref:
            assert r[0] == ref[0] and r[1] == ref[1]


def random_board(seed, w, h, L, masks=True):
    rng = np.random.default_rng(seed)
    n = w * h * L
    cost = rng.integers(1, 7, size=n).astype(np.uint32)
    cost[rng.random(n) < 0.08] = 0
    cost[rng.random(n) < 0.18] = MAX
    board = vf = None
    if masks:
        board = np.zeros(n, dtype=np.uint8)
        r = rng.random(n)
        board[r < 0.04] |= 1
        board[(r >= 0.04) & (r < 0.10)] |= 2
        for bit in (4, 8, 16, 32):
            board[rng.random(n) < 0.05] |= bit
        vf = (rng.random(n) < 0.1).astype(np.uint8)
    return mrcpu.Grid(w, h, L, cost, via_forbidden=vf, board=board)


def free_cells(g):
    ok = g.cost != MAX
    if g.board is not None:
        ok &= (g.board & 1) == 0
    return np.flatnonzero(ok)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("seed", [1, 2])
def test_random_fields_all_features(eng, mode, shape, seed):
    """Every mask, pad, via exemption, additive and halo feature at once."""
    configure(eng, shape, mode)
    rng = np.random.default_rng(100 + seed)
    w, h, L = 37, 23, 3
    g = random_board(seed, w, h, L)
    ex = [int(v) for v in rng.integers(0, 30, size=w - 1)]
    ey = [int(v) for v in rng.integers(0, 30, size=h - 1)]
    edges = mrcpu.Edges(ex, ey, [7, None] if seed % 2 else [0, MAX])
    add = rng.integers(0, 20, size=g.n).astype(np.uint32)
    halo = np.full(g.n, -1, dtype=np.int32)
    sel = rng.random(g.n) < 0.1
    halo[sel] = rng.integers(0, 4, size=int(sel.sum()))
    halo[rng.random(g.n) < 0.03] = -2
    free = free_cells(g)
    obst = np.flatnonzero(g.cost == MAX)
    nets, sources, dsts, wins = [], [], [], []
    for i in range(6):
        s = int(rng.choice(free))
        d = int(rng.choice(obst)) if i % 3 == 0 else int(rng.choice(free))
        pads = sorted(set([d] + [int(c) for c in rng.choice(obst, size=4)])) if i % 3 == 0 else []
        vpads = pads[:2]
        nets.append(mrcpu.Net(s, d, pads, vpads, own_id=i % 4))
        sources.append([s] + ([int(rng.choice(free))] if i % 2 else []))
        dsts.append(d)
        wins.append(None if i < 3 else (3, 2, w - 5, h - 3))
    for k in range(len(nets)):
        if wins[k] is None:
            wins[k] = (0, 0, w - 1, h - 1)
    tmpl = mrcpu.Spec(g, edges, additive=add, halo_owner=halo, halo_penalty=13)
    ref_d, ref_h = mrcpu.fields_batch(tmpl, nets, sources, wins)
    dg = as_dev(g)
    pr = mrfield.Pricing(add, halo, 13, [n.own_id for n in nets])
    # the engine takes pads through nets only on the provider seam; fields use
    # the pricing and sources, so compare the pad-free nets here and pads below
    pad_free = [i for i, n in enumerate(nets) if not n.pads]
    got_d, got_h = eng.fields(dg, as_dev_edges(edges), [sources[i] for i in pad_free],
                              dsts=[dsts[i] for i in pad_free], windows=[wins[i] for i in pad_free],
                              pricing=mrfield.Pricing(add, halo, 13, [nets[i].own_id for i in pad_free]))
    for j, i in enumerate(pad_free):
        assert np.array_equal(got_d[j], ref_d[i]), (i, np.flatnonzero(got_d[j] != ref_d[i])[:10])
        assert np.array_equal(got_h[j], ref_h[i]), i


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape", SHAPES[:3])
def test_random_paths_with_pads_and_masks(eng, mode, shape):
    configure(eng, shape, mode)
    for seed in range(4):
        rng = np.random.default_rng(500 + seed)
        w, h, L = 29, 19, 3
        g = random_board(seed + 10, w, h, L)
        edges = mrcpu.Edges([int(v) for v in rng.integers(1, 20, size=w - 1)],
                            [int(v) for v in rng.integers(1, 20, size=h - 1)],
                            [int(rng.integers(0, 40)), None if seed % 2 else 5])
        free = free_cells(g)
        obst = np.flatnonzero(g.cost == MAX)
        nets = []
        for i in range(10):
            s = int(rng.choice(free))
            d = int(rng.choice(obst)) if i % 2 else int(rng.choice(free))
            extra = [int(c) for c in rng.choice(obst, size=6)]
            pads = sorted(set(([d] if g.cost[d] == MAX else []) + extra))
            if g.board is not None and g.board[d] & 1:
                continue
            nets.append(mrcpu.Net(s, d, pads, pads[:3]))
        wins = [mrcpu.net_window(g, n.src, n.dst, n.pads) for n in nets]
        ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(g, edges), nets, wins, retry_full=True)
        got = eng.route_isolated_batch(as_dev(g), [mrfield.Net(n.src, n.dst, n.pads, n.via_pads) for n in nets],
                                       wins, as_dev_edges(edges))
        for i, r in enumerate(got):
            assert (r is None) == (ref_p[i] is None), (seed, i)
            if r is not None:
                assert r[0] == ref_p[i] and r[1] == ref_c[i], (seed, i)


@pytest.mark.parametrize("mode", MODES)
def test_config1_complete_fields_and_paths(eng, mode):
    """128x128x2, 64 nets: every field cell and every path bit-exact."""
    configure(eng, (32, 32, 256, 32), mode)
    s = mrfield.synth_config(1, seed=1)
    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    nets = [mrcpu.Net(int(a), int(b)) for a, b in zip(s.src, s.dst)]
    ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(og, oe), nets, [[n.src] for n in nets], threads=8)
    got_d, got_h = eng.fields(g, s.edges, [[int(a)] for a in s.src], dsts=[int(b) for b in s.dst])
    assert np.array_equal(got_d, ref_d)
    assert np.array_equal(got_h, ref_h)
    full = [(0, 0, g.w - 1, g.h - 1)] * len(nets)
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, oe), nets, full, threads=8)
    got = eng.route_isolated_batch(g, s.nets(), full, s.edges)
    n_routed = 0
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), i
        if r is not None:
            n_routed += 1
            assert r[0] == ref_p[i] and r[1] == ref_c[i], i
    assert n_routed > 32
    flat = eng.route_isolated_batch_arrays(g, s.nets(), full, s.edges)
    assert len(flat) == len(nets)
    for i in range(len(nets)):
        assert flat.path(i) == got[i], i


@pytest.mark.parametrize("delta", [1, 60, 2000])
def test_delta_stepping_matches_oracle(eng, delta):
    """Plain form with tile delta stepping (rounds process only the tiles whose
    activating cost is within delta of the smallest pending one): fields and
    paths unchanged, including windowed nets with the full-board retry."""
    configure(eng, (16, 16, 256, 32), mrfield.MODE_PLAIN)
    eng.set_delta(delta)
    try:
        s = mrfield.synth_config(1, seed=2)
        g = s.grid
        og = mrcpu.Grid(g.w, g.h, g.layers, g.cost)
        oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
        nets = [mrcpu.Net(int(a), int(b)) for a, b in zip(s.src, s.dst)]
        ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(og, oe), nets, [[n.src] for n in nets], threads=8)
        got_d, got_h = eng.fields(g, s.edges, [[int(a)] for a in s.src], dsts=[int(b) for b in s.dst])
        assert np.array_equal(got_d, ref_d)
        assert np.array_equal(got_h, ref_h)
        for case in range(4):
            g2, nets2, edges2 = isolated_stress_case(case)
            wins = [mrcpu.net_window(g2, n.src, n.dst, n.pads) for n in nets2]
            ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(g2, edges2), nets2, wins, retry_full=True)
            dn = [mrfield.Net(n.src, n.dst, n.pads, n.via_pads) for n in nets2]
            got = eng.route_isolated_batch(as_dev(g2), dn, wins, as_dev_edges(edges2))
            for i in range(len(nets2)):
                assert (got[i] is None) == (ref_p[i] is None), (case, i)
                if got[i] is not None:
                    assert got[i][0] == ref_p[i] and got[i][1] == ref_c[i], (case, i)
    finally:
        eng.set_delta(MAX)


@pytest.mark.skip(reason="persistent grid-barrier form disabled: did not return on MI355X; cause not yet shown")
def test_persistent_form_agrees_with_plain(eng):
    s = mrfield.synth_config(1, seed=3, scale_nets=24)
    g = s.grid
    configure(eng, (16, 16, 256, 8), mrfield.MODE_PLAIN)
    ref = eng.fields(g, s.edges, [[int(a)] for a in s.src])
    configure(eng, (16, 16, 256, 8), mrfield.MODE_PERSISTENT)
    got = eng.fields(g, s.edges, [[int(a)] for a in s.src])
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])


def test_persistent_form_is_refused_cleanly(eng):
    configure(eng, (16, 16, 128, 8), mrfield.MODE_PERSISTENT)
    s = mrfield.synth_config(1, seed=3, scale_nets=2)
    with pytest.raises(mrfield.MrfError) as e:
        eng.fields(s.grid, s.edges, [[int(a)] for a in s.src])
    assert e.value.code == mrfield.BACKEND_UNAVAILABLE
    configure(eng, (16, 16, 128, 8), mrfield.MODE_GRAPH)


def test_plain_and_graph_agree_and_are_deterministic(eng):
    s = mrfield.synth_config(1, seed=3, scale_nets=24)
    g = s.grid
    out = []
    for mode in (mrfield.MODE_PLAIN, mrfield.MODE_GRAPH):
        for _ in range(2):
            configure(eng, (16, 16, 256, 8), mode)
            out.append(eng.fields(g, s.edges, [[int(a)] for a in s.src]))
    for d, h in out[1:]:
        assert np.array_equal(d, out[0][0]) and np.array_equal(h, out[0][1])


def test_alone_equals_mixed(eng):
    """A net's result is the same solved alone and inside a larger batch."""
    configure(eng, (16, 8, 128, 8), mrfield.MODE_GRAPH)
    s = mrfield.synth_config(1, seed=5, scale_nets=40)
    g = s.grid
    full = [(0, 0, g.w - 1, g.h - 1)]
    mixed = eng.route_isolated_batch(g, s.nets(), full * 40, s.edges)
    for i in (0, 7, 21, 39):
        alone = eng.route_isolated_batch(g, [s.nets()[i]], full, s.edges)[0]
        assert alone == mixed[i]


def test_fail_closed_contracts(eng):
    configure(eng, (32, 32, 256, 32), mrfield.MODE_GRAPH)
    g = mrfield.Grid(3, 2, 1, np.ones(6, dtype=np.uint32))
    e = mrfield.Edges([16, 16], [16], [])
    bad = mrfield.Grid(3, 2, 1, np.ones(5, dtype=np.uint32))
    with pytest.raises(mrfield.MrfError) as x:
        eng.route_isolated_batch(bad, [mrfield.Net(0, 1)], [(0, 0, 2, 1)], e)
    assert x.value.code == mrfield.MALFORMED_GRID
    with pytest.raises(mrfield.MrfError) as x:
        eng.route_isolated_batch(g, [mrfield.Net(0, 99)], [(0, 0, 2, 1)], e)
    assert x.value.code == mrfield.INVALID_ENDPOINT
    with pytest.raises(mrfield.MrfError) as x:
        eng.route_isolated_batch(g, [mrfield.Net(0, 1)], [(0, 0, 3, 1)], e)
    assert x.value.code == mrfield.BACKEND_UNAVAILABLE
    with pytest.raises(mrfield.MrfError) as x:
        eng.route_isolated_batch(g, [mrfield.Net(0, 1)], [(0, 0, 2, 1)], mrfield.Edges([16], [16], []))
    assert x.value.code == mrfield.BACKEND_UNAVAILABLE
    assert eng.route_isolated_batch(g, [], [], e) == []


def test_same_round_neighbours_after_a_layout_change(eng):
    """A tile whose neighbour is first visited in the same round must not take
    that neighbour's border labels: its state and labels are loaded together,
    unordered, and the label memory still holds the previous solve's values
    (here laid out for 16x16 tiles, so they index other cells).  Solving the
    same batch at 16x16, then at 8x4 tiles with 128 threads and a round cap of
    4 (short visits, many neighbours first visited in one round), gives the
    same paths in both forms."""
    s = mrfield.synth_config(2, seed=1)
    board = mrfield.Board(eng, s.grid, s.edges, s.pricing())
    nets = s.nets()
    try:
        configure(eng, (16, 16, 256, 32), mrfield.MODE_PLAIN)
        ref = board.solve_arrays(nets)
        for mode in (mrfield.MODE_PLAIN, mrfield.MODE_GRAPH):
            configure(eng, (16, 16, 256, 32), mrfield.MODE_PLAIN)
            board.solve_arrays(nets)
            configure(eng, (8, 4, 128, 4), mode)
            got = board.solve_arrays(nets)
            for k in ("routed", "cost", "offsets", "cells"):
                assert np.array_equal(getattr(got, k), getattr(ref, k)), (mode, k)
    finally:
        board.close()
        configure(eng, (16, 16, 256, 32), mrfield.MODE_PLAIN)


def test_stamp_wait_falls_back_to_a_stream_synchronisation(monkeypatch):
    """The plain form waits for each launch's stamp in host-mapped memory and
    hands over to a stream synchronisation after MRF_STAMP_SPIN_US (read when
    a context opens); with 0 every round takes the hand-over, and the paths
    are the same."""
    s = mrfield.synth_config(1, seed=5)
    wins = [(0, 0, s.grid.w - 1, s.grid.h - 1)] * len(s.nets())
    e = mrfield.Engine(0)
    try:
        configure(e, (16, 16, 256, 32), mrfield.MODE_PLAIN)
        ref = e.route_isolated_batch(s.grid, s.nets(), wins, s.edges)
    finally:
        e.close()
    monkeypatch.setenv("MRF_STAMP_SPIN_US", "0")
    e = mrfield.Engine(0)
    monkeypatch.delenv("MRF_STAMP_SPIN_US")
    try:
        configure(e, (16, 16, 256, 32), mrfield.MODE_PLAIN)
        got = e.route_isolated_batch(s.grid, s.nets(), wins, s.edges)
        assert got == ref and e.stats()["relax_launches"] > 1
    finally:
        e.close()


@pytest.mark.parametrize("mode,message", [(1, "injected backend failure"), (2, "re-validation")])
def test_host_revalidation_fails_closed(monkeypatch, mode, message):
    """A device result the host cannot re-validate fails the whole call (the
    provider's fail-closed contract, mr-metal lib.rs:902-974).  The library's
    MRF_INJECT_FAULT hook, read when a context opens, either raises after the
    device solve (1) or flips one returned cell on the host before
    re-validation (2)."""
    s = mrfield.synth_config(1, seed=3, scale_nets=8)
    monkeypatch.setenv("MRF_INJECT_FAULT", str(mode))
    e = mrfield.Engine(0)
    monkeypatch.delenv("MRF_INJECT_FAULT")
    try:
        with pytest.raises(mrfield.MrfError) as x:
            e.route_isolated_batch(s.grid, s.nets(), [(0, 0, s.grid.w - 1, s.grid.h - 1)] * len(s.nets()), s.edges)
        assert x.value.code == mrfield.BACKEND_UNAVAILABLE and message in str(x.value)
    finally:
        e.close()
    # a context opened without the hook routes the same nets
    e = mrfield.Engine(0)
    try:
        got = e.route_isolated_batch(s.grid, s.nets(), [(0, 0, s.grid.w - 1, s.grid.h - 1)] * len(s.nets()), s.edges)
        assert sum(r is not None for r in got) > 0
    finally:
        e.close()


def test_autotune_keeps_results(eng):
    """The autotuner only changes time: paths after tuning equal the oracle's."""
    s = mrfield.synth_config(3, seed=2, scale_nets=48)
    g = s.grid
    board = mrfield.Board(eng, g, s.edges, s.pricing())
    try:
        # the context's own configuration is a finalist too: start from a known one
        configure(eng, (32, 32, 256, 32), mrfield.MODE_PLAIN)
        t = board.autotune(s.nets(), windows=s.windows, sample=16)
        assert t["tried"] > 10 and t["sample_ms"] > 0 and t["default_ms"] > 0
        assert (t["tile_w"], t["tile_h"]) in [(32, 32), (32, 16), (16, 16), (16, 8), (8, 4)]
        got = board.solve(s.nets(), windows=s.windows, retry_full=True)
    finally:
        board.close()
        configure(eng, (16, 16, 256, 32), mrfield.MODE_PLAIN)
        eng.set_delta(MAX)
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    spec = mrcpu.Spec(og, oe, additive=s.additive)
    nets = [mrcpu.Net(int(a), int(b)) for a, b in zip(s.src, s.dst)]
    ref_p, ref_c = mrcpu.alone_batch(spec, nets, s.windows, retry_full=True, threads=8)
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), i
        if r is not None:
            assert r[0] == ref_p[i] and r[1] == ref_c[i], i
