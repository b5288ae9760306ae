"""GPU: the persistent work-queue form (MODE_QUEUE) against the oracle.

One launch relaxes a whole batch: workgroups pop tiles from device work
queues and push the neighbours their visits improve.  Results must be the
oracle's, bit for bit, like every other form's: the reference vectors, the
stress batteries, random boards with every feature, complete fields of config
1, the windowed batch of config 3 and the config-2 board.  A wait limit of 0
(set from the host when the context opens) makes the launch stop with an
error instead of waiting, and the call fails closed.
"""
from __future__ import annotations

import numpy as np
import pytest

import mrcpu
import mrfield
from cases import grid_from_case, isolated_stress_case, load_vectors, oracle_parts
from test_gpu_parity import SHAPES, as_dev, as_dev_edges, free_cells, random_board

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = mrfield.Engine(0)
    e.set_mode(mrfield.MODE_QUEUE)
    yield e
    e.close()


def test_queue_reference_vectors(eng):
    eng.set_tiles(4, 1, 64, 4)
    for case in load_vectors()["cases"]:
        if case.get("expect_invalid_endpoint"):
            continue
        g, src, dst = grid_from_case(case)
        net = mrfield.Net(src, dst, case.get("pads", []))
        if case["mode"] == "lee":
            got = eng.router_route(as_dev(g), [net])[0]
        else:
            via = case.get("via", [160] * (g.layers - 1))
            ed = mrfield.Edges([16] * (g.w - 1), [16] * (g.h - 1), via)
            got = eng.route_isolated_batch(as_dev(g), [net], [(0, 0, g.w - 1, g.h - 1)], ed)[0]
        if case.get("expect_unrouted"):
            assert got is None, case["name"]
            continue
        assert got is not None, case["name"]
        if "expect_path" in case:
            assert got[0] == case["expect_path"], case["name"]
        for key in ("expect_cost", "expect_search_cost"):
            if key in case:
                assert got[1] == case[key], case["name"]


@pytest.mark.parametrize("shape", SHAPES[:3])
def test_queue_isolated_stress(eng, shape):
    eng.set_tiles(*shape)
    for case in range(12):
        g, nets, edges = isolated_stress_case(case)
        wins = [mrcpu.net_window(g, n.src, n.dst, n.pads) for n in nets]
        ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(g, edges), nets, wins, retry_full=True)
        got = eng.route_isolated_batch(as_dev(g), [mrfield.Net(n.src, n.dst, n.pads, n.via_pads) for n in nets],
                                       wins, as_dev_edges(edges))
        for i in range(len(nets)):
            assert (got[i] is None) == (ref_p[i] is None), (case, i)
            if got[i] is not None:
                assert got[i] == (ref_p[i], ref_c[i]), (case, i)


# workgroup visits (relax_queue) and the wave visit's queue kernel
QSHAPES = [(16, 16, 256, 32), (16, 16, 64, 32)]


def assert_queue_kernel(eng, shape):
    st = eng.stats()
    assert st["relax_launches"] == 1
    assert st["wave_launches"] == (1 if shape[2] == 64 else 0), st


@pytest.mark.parametrize("shape", QSHAPES)
def test_queue_config1_fields_and_paths(eng, shape):
    eng.set_tiles(*shape)
    s = mrfield.synth_config(1, seed=1)
    g = s.grid
    og, oe = oracle_parts(s)
    nets = [mrcpu.Net(int(a), int(b)) for a, b in zip(s.src, s.dst)]
    ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(og, oe), nets, [[n.src] for n in nets], threads=8)
    got_d, got_h = eng.fields(g, s.edges, [[int(a)] for a in s.src], dsts=[int(b) for b in s.dst])
    assert np.array_equal(got_d, ref_d) and np.array_equal(got_h, ref_h)
    assert_queue_kernel(eng, shape)
    assert eng.stats()["tile_visits"] > 0
    full = [(0, 0, g.w - 1, g.h - 1)] * len(nets)
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, oe), nets, full, threads=8)
    got = eng.route_isolated_batch(g, s.nets(), full, s.edges)
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), i
        if r is not None:
            assert r == (ref_p[i], ref_c[i]), i


@pytest.mark.parametrize("shape", [(32, 32, 256, 32), (8, 4, 64, 8), (16, 8, 128, 3)])
def test_queue_random_fields_all_features(eng, shape):
    """Pads, via exemptions, halo pricing, additive planes and wide steps."""
    eng.set_tiles(*shape)
    rng = np.random.default_rng(7)
    w, h, L = 37, 23, 3
    g = random_board(3, w, h, L)
    ex = [int(v) for v in rng.integers(0, 30, size=w - 1)]
    ey = [int(v) for v in rng.integers(0, 30, size=h - 1)]
    edges = mrcpu.Edges(ex, ey, [0, mrfield.MAX])
    add = rng.integers(0, 20, size=g.n).astype(np.uint32)
    halo = np.full(g.n, -1, dtype=np.int32)
    sel = rng.random(g.n) < 0.1
    halo[sel] = rng.integers(0, 4, size=int(sel.sum()))
    free = free_cells(g)
    nets, sources, dsts, wins = [], [], [], []
    for i in range(6):
        s0 = int(rng.choice(free))
        d0 = int(rng.choice(free))
        nets.append(mrcpu.Net(s0, d0, own_id=i % 4))
        sources.append([s0] + ([int(rng.choice(free))] if i % 2 else []))
        dsts.append(d0)
        wins.append((0, 0, w - 1, h - 1) if i < 3 else (3, 2, w - 5, h - 3))
    tmpl = mrcpu.Spec(g, edges, additive=add, halo_owner=halo, halo_penalty=13)
    ref_d, ref_h = mrcpu.fields_batch(tmpl, nets, sources, wins)
    got_d, got_h = eng.fields(as_dev(g), as_dev_edges(edges), sources, dsts=dsts, windows=wins,
                              pricing=mrfield.Pricing(add, halo, 13, [n.own_id for n in nets]))
    for i in range(len(nets)):
        assert np.array_equal(got_d[i], ref_d[i]), i
        assert np.array_equal(got_h[i], ref_h[i]), i
    # pads and via pads on the provider seam
    obst = np.flatnonzero(g.cost == mrfield.MAX)
    pn = []
    for i in range(6):
        s0 = int(rng.choice(free))
        d0 = int(rng.choice(obst))
        if g.board is not None and g.board[d0] & 1:
            continue
        pads = sorted(set([d0] + [int(c) for c in rng.choice(obst, size=5)]))
        pn.append(mrcpu.Net(s0, d0, pads, pads[:2]))
    pw = [mrcpu.net_window(g, n.src, n.dst, n.pads) for n in pn]
    e2 = mrcpu.Edges(ex, ey, [7, None])
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(g, e2), pn, pw, retry_full=True)
    got = eng.route_isolated_batch(as_dev(g), [mrfield.Net(n.src, n.dst, n.pads, n.via_pads) for n in pn], pw,
                                   as_dev_edges(e2))
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), i
        if r is not None:
            assert r == (ref_p[i], ref_c[i]), i


@pytest.mark.parametrize("shape", QSHAPES)
def test_queue_config3_windowed_batch(eng, shape):
    eng.set_tiles(*shape)
    s = mrfield.synth_config(3, seed=1)
    board = mrfield.Board(eng, s.grid, s.edges, s.pricing())
    try:
        got = board.solve_arrays(s.nets(), windows=s.windows, retry_full=True)
        assert eng.stats()["relax_launches"] >= 1
    finally:
        board.close()
    idx = list(range(0, 1024, 8))
    og, oe = oracle_parts(s)
    spec = mrcpu.Spec(og, oe, additive=s.additive, use_heuristic=True, min_via_cost=80)
    nets = [mrcpu.Net(int(s.src[i]), int(s.dst[i])) for i in idx]
    ref_p, ref_c = mrcpu.alone_batch(spec, nets, [s.windows[i] for i in idx], retry_full=True, threads=8)
    for k, i in enumerate(idx):
        gp = got.path(i)
        assert (gp is None) == (ref_p[k] is None), i
        if gp is not None:
            assert gp == (ref_p[k], ref_c[k]), i


@pytest.mark.parametrize("history", [True, False])
@pytest.mark.parametrize("shape", QSHAPES)
def test_queue_config2_equals_plain(eng, shape, history):
    """The config-2 board with its history plane (the wave visits load their
    step prices) and without (a 12.6 MB step plane would not fit L2: the
    visits form their prices), queue form against plain form."""
    s = mrfield.synth_config(2, seed=1)
    board = mrfield.Board(eng, s.grid, s.edges, s.pricing() if history else None)
    try:
        eng.set_tiles(*shape)
        got = board.solve_arrays(s.nets())
        assert_queue_kernel(eng, shape)
        eng.set_mode(mrfield.MODE_PLAIN)
        ref = board.solve_arrays(s.nets())
    finally:
        eng.set_mode(mrfield.MODE_QUEUE)
        board.close()
    for k in ("routed", "cost", "offsets", "cells"):
        assert np.array_equal(getattr(got, k), getattr(ref, k)), k


def test_queue_wait_limit_fails_closed(monkeypatch):
    """MRF_QUEUE_WAIT_MS=0: a workgroup that finds no work while other work is
    pending gives up at once with an error flag, every workgroup exits, and
    the call fails with BACKEND_UNAVAILABLE; a context opened without the
    limit solves the same batch."""
    s = mrfield.synth_config(1, seed=3, scale_nets=8)
    full = [(0, 0, s.grid.w - 1, s.grid.h - 1)] * len(s.src)
    monkeypatch.setenv("MRF_QUEUE_WAIT_MS", "0")
    e = mrfield.Engine(0)
    monkeypatch.delenv("MRF_QUEUE_WAIT_MS")
    try:
        e.set_mode(mrfield.MODE_QUEUE)
        with pytest.raises(mrfield.MrfError) as x:
            e.route_isolated_batch(s.grid, s.nets(), full, s.edges)
        assert x.value.code == mrfield.BACKEND_UNAVAILABLE
    finally:
        e.close()
    e = mrfield.Engine(0)
    try:
        e.set_mode(mrfield.MODE_QUEUE)
        got = e.route_isolated_batch(s.grid, s.nets(), full, s.edges)
        e.set_mode(mrfield.MODE_PLAIN)
        assert got == e.route_isolated_batch(s.grid, s.nets(), full, s.edges)
    finally:
        e.close()
