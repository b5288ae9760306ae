"""Pins the Bellman certificate (tests/certify.py) on the CPU: it accepts the
oracle's fields, rejects planted mistakes, and its step prices equal the
oracle's mrc_step_cost on every edge of small boards."""
from __future__ import annotations

import numpy as np
import pytest

import certify
import mrcpu
from cases import MAX


def board(seed, w=13, h=9, L=3):
    rng = np.random.default_rng(seed)
    n = w * h * L
    cost = rng.integers(0, 9, size=n).astype(np.uint32)
    cost[rng.random(n) < 0.2] = MAX
    b = np.zeros(n, dtype=np.uint8)
    b[rng.random(n) < 0.05] |= 1
    b[rng.random(n) < 0.1] |= 2
    for bit in (4, 8, 16, 32):
        b[rng.random(n) < 0.06] |= bit
    vf = (rng.random(n) < 0.15).astype(np.uint8)
    add = rng.integers(0, 50, size=n).astype(np.uint32)
    halo = np.where(rng.random(n) < 0.2, rng.integers(-2, 3, size=n), -1).astype(np.int32)
    x = [int(v) for v in rng.integers(0, 40, size=w - 1)]
    y = [int(v) for v in rng.integers(0, 40, size=h - 1)]
    via = [int(rng.integers(0, 60)), None]
    g = mrcpu.Grid(w, h, L, cost, via_forbidden=vf, board=b)
    e = mrcpu.Edges(x, y, via)
    p = certify.Problem(w, h, L, cost, b, vf, x, y, via, add, halo, 11)
    return g, e, p, add, halo, rng


@pytest.mark.parametrize("seed", range(6))
def test_certificate_accepts_oracle_fields_and_rejects_plants(seed):
    g, e, p, add, halo, rng = board(seed)
    free = np.flatnonzero((g.cost != MAX) & ((g.board & 1) == 0))
    obst = np.flatnonzero(g.cost == MAX)
    pads = sorted(int(c) for c in rng.choice(obst, size=5, replace=False))
    vpads = pads[:2]
    srcs = [int(c) for c in rng.choice(free, size=2, replace=False)]
    win = (1, 1, g.w - 2, g.h - 1)
    spec = mrcpu.Spec(g, e, pads=pads, via_pads=vpads, window=win, additive=add, halo_owner=halo,
                      halo_penalty=11, own_id=1)
    dist, hops = mrcpu.field_lex(spec, srcs, srcs)
    args = dict(window=win, sources=srcs, pads=pads, via_pads=vpads, forced=srcs, own_id=1)
    assert certify.certify(p, dist, hops, **args) is None
    reached = np.flatnonzero((dist != MAX) & (dist > 0))
    assert len(reached) > 10
    # planted mistakes: one label too small, one too large, a wrong hop count,
    # an unreachable cell given a label, a dropped additive term
    for plant in range(5):
        d2, h2 = dist.copy(), hops.copy()
        c = int(reached[plant * 7 % len(reached)])
        if plant == 0:
            d2[c] -= 1
        elif plant == 1:
            d2[c] += 1
        elif plant == 2:
            h2[c] += 1
        elif plant == 3:
            u = int(np.flatnonzero(dist == MAX)[0])
            d2[u], h2[u] = 5, 1
        else:
            p2 = certify.Problem(p.w, p.h, p.L, g.cost, g.board, g.via_forbidden, e.x, e.y, e.via,
                                 None, halo, 11)
            assert certify.certify(p2, dist, hops, **args) is not None
            continue
        assert certify.certify(p, d2, h2, **args) is not None, plant


@pytest.mark.parametrize("seed", range(3))
def test_step_prices_equal_oracle_step_cost(seed):
    g, e, p, add, halo, rng = board(seed, 7, 5, 3)
    vpads = [int(c) for c in rng.choice(g.n, size=6, replace=False)]
    pads = sorted(set(vpads + [int(c) for c in np.flatnonzero(g.cost == MAX)[:4]]))
    spec = mrcpu.Spec(g, e, pads=pads, via_pads=vpads, additive=add, halo_owner=halo, halo_penalty=11, own_id=2)
    full = (0, 0, g.w - 1, g.h - 1)
    passable, wgt, a, vp = certify.field_terms(p, full, pads, vpads, (), 2)
    passable[:] = True  # compare prices and edge legality, not node passability
    shape = (g.layers, g.h, g.w)
    idx = np.arange(g.n).reshape(shape)
    seen = 0
    for sl_u, sl_v, legal, step in certify.in_edges(p, passable, wgt, a, vp):
        us, vs = idx[sl_u].ravel(), idx[sl_v].ravel()
        for u, v, ok, s in zip(us, vs, legal.ravel(), step.ravel()):
            ref = mrcpu.step_cost(spec, int(u), int(v))
            assert (ref != MAX) == bool(ok), (u, v)
            if ok:
                assert ref == int(s), (u, v)
            seen += 1
    assert seen > 300
