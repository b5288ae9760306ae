"""GPU: host-side contracts of the library around the kernels.

* A batch whose labels exceed the context's label budget is split into
  device batches that fit (mr-metal's chunking by MAX_BATCH_CELLS /
  fields_per_batch, crates/mr-metal/src/lib.rs:251-296, :626-678); the
  results do not depend on the split.
* The Router entry's congestion plane is BoardRoute::congestion_from over its
  paths (crates/mr-core/src/lib.rs:575-593, built by MetalRouter::route at
  crates/mr-metal/src/lib.rs:1227).
* A call that finds the context's GPU lane busy returns BACKEND_UNAVAILABLE at
  once instead of waiting (crates/mr-cli/src/lib.rs:499, :620-622, tested by
  metal_provider_falls_back_instead_of_waiting_for_a_busy_gpu_lane,
  :2184-2203).
"""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

import mrcpu
import mrfield
from cases import oracle_parts, weighted_stress

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = mrfield.Engine(0)
    e.set_tiles(16, 16, 256, 32)
    e.set_mode(mrfield.MODE_PLAIN)
    yield e
    e.close()


def field_bytes(w, h, layers, tw=16, th=16):
    tx, ty = -(-w // tw), -(-h // th)
    return tx * ty * (layers * th * tw + 2 * layers * th) * 8


def test_label_budget_splits_batches(eng):
    s = mrfield.synth_config(1, seed=4, scale_nets=24)
    g = s.grid
    full = [(0, 0, g.w - 1, g.h - 1)] * len(s.src)
    default = eng.label_budget()
    assert default > 1 << 30
    ref = eng.route_isolated_batch(g, s.nets(), full, s.edges)
    ref_f = eng.fields(g, s.edges, [[int(a)] for a in s.src[:9]], dsts=[int(b) for b in s.dst[:9]])
    try:
        # room for three fields per device batch: 8 batches of nets, 3 of fields
        eng.set_label_budget(3 * field_bytes(g.w, g.h, g.layers) + 64)
        assert eng.label_budget() < default
        got = eng.route_isolated_batch(g, s.nets(), full, s.edges)
        assert eng.stats()["relax_launches"] > 0
        got_f = eng.fields(g, s.edges, [[int(a)] for a in s.src[:9]], dsts=[int(b) for b in s.dst[:9]])
        # a budget below one field still runs that field alone
        eng.set_label_budget(1024)
        one = eng.route_isolated_batch(g, s.nets()[:3], full[:3], s.edges)
    finally:
        eng.set_label_budget(0)
    assert eng.label_budget() == default
    assert got == ref
    assert np.array_equal(got_f[0], ref_f[0]) and np.array_equal(got_f[1], ref_f[1])
    assert one == ref[:3]
    og, oe = oracle_parts(s)
    nets = [mrcpu.Net(int(a), int(b)) for a, b in zip(s.src, s.dst)]
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, oe), nets, full, threads=8)
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), i
        if r is not None:
            assert r == (ref_p[i], ref_c[i]), i


def test_router_congestion_equals_congestion_from(eng):
    g, sources, nets = weighted_stress()
    dg = mrfield.Grid(g.w, g.h, g.layers, g.cost)
    got = eng.router_route(dg, [mrfield.Net(s, d) for s, d in nets])
    lee = mrcpu.Spec(g, mode=mrcpu.MODE_LEE)
    ref = [mrcpu.astar(lee, s, d) for s, d in nets]
    ref_paths = [r[0] if r else None for r in ref]
    cong = mrfield.congestion_from((g.w, g.h, g.layers), got)
    assert np.array_equal(cong, mrcpu.congestion_from((g.w, g.h, g.layers), ref_paths))
    assert cong.max() >= 2  # paths overlap on this board
    assert int(cong.sum()) == sum(len(set(p)) for p in ref_paths if p)


def test_busy_lane_fails_fast(eng):
    s = mrfield.synth_config(2, seed=1)
    board = mrfield.Board(eng, s.grid, s.edges, s.pricing())
    tiny = mrfield.Grid(2, 1, 1, np.ones(2, dtype=np.uint32))
    te = mrfield.Edges([16], [], [])
    stop = threading.Event()
    errors = []

    def hold():
        while not stop.is_set():
            try:
                board.solve_arrays(s.nets())
            except mrfield.MrfError as exc:
                if "busy" not in str(exc):  # the probe below may hold the lane for a moment
                    errors.append(exc)
                    return

    t = threading.Thread(target=hold)
    t.start()
    busy, quick = 0, []
    try:
        t_end = time.perf_counter() + 3.0
        while time.perf_counter() < t_end and busy < 5:
            t0 = time.perf_counter()
            try:
                eng.route_isolated_batch(tiny, [mrfield.Net(0, 1)], [(0, 0, 1, 0)], te)
            except mrfield.MrfError as e:
                assert e.code == mrfield.BACKEND_UNAVAILABLE and "busy" in str(e)
                busy += 1
                quick.append(time.perf_counter() - t0)
    finally:
        stop.set()
        t.join()
        board.close()
    assert not errors, errors
    assert busy >= 1
    # the refusal does not wait for the running solve (about 11 ms)
    assert min(quick) < 0.005
    # once the lane is free the same call runs
    assert eng.route_isolated_batch(tiny, [mrfield.Net(0, 1)], [(0, 0, 1, 0)], te)[0][0] == [0, 1]
