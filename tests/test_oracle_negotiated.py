"""Pins the oracle's negotiated pricing (route_negotiated, negotiated.rs:2746-2856).

Checks:
  * the counted-halo fixture of negotiated.rs:6535-6660: a net whose own
    copper and own halo are the whole occupancy snapshot routes exactly as on
    an empty snapshot, in the term-by-term and in the fused form, and the
    negative control (own halo not subtracted) costs more;
  * the randomized fused == term-by-term comparison of
    negotiated.rs:6665-6880 (random grids, windows, pads, via models, counts
    up to u32::MAX), with the complete-field search agreeing with A*;
  * hand-computed closed forms on a three-cell corridor (history, own copper,
    own halo multiplicity, saturation to an unreachable price).
"""
from __future__ import annotations

import numpy as np

import mrcpu
from cases import MAX

U32MAX = 0xFFFFFFFF


def neg(n, history=None, present=None, halo=None, pfac=1, fused=False):
    z = np.zeros(n, dtype=np.uint32)
    return mrcpu.Negotiation(
        history=z.copy() if history is None else np.asarray(history, dtype=np.uint32),
        present=z.copy() if present is None else np.asarray(present, dtype=np.uint32),
        present_halo=None if halo is None else np.asarray(halo, dtype=np.uint32),
        pfac=pfac, fused=fused)


def radius1_halo(w, h, path):
    """for_each_halo_cell at clearance 1 on a uniform obstacle-free layer: the
    3x3 Chebyshev box of every path cell, each visit counted."""
    out = []
    for c in path:
        cx, cy = c % w, c // w
        for y in range(max(0, cy - 1), min(h, cy + 2)):
            for x in range(max(0, cx - 1), min(w, cx + 2)):
                out.append(y * w + x)
    return out


def test_counted_halo_self_cost_is_zero():
    w, h = 7, 5
    g = mrcpu.Grid.filled(w, h, 1, 1)
    e = mrcpu.Edges.uniform(w, h, 1, planar=16, via=160)
    src, dst = 2 * w + 1, 2 * w + 5
    old_path = [2 * w + x for x in range(1, 6)]
    present = np.zeros(g.n, dtype=np.uint32)
    present[old_path] += 1
    halo_cells = radius1_halo(w, h, old_path)
    halo = np.bincount(halo_cells, minlength=g.n).astype(np.uint32)
    assert halo[2 * w + 3] > 1  # the fixture needs overlapping footprints
    pfac = 7

    unfused = mrcpu.Spec(g, e, negotiation=neg(g.n, present=present, halo=halo, pfac=pfac),
                         own_path=old_path, own_halo=halo_cells, use_heuristic=True)
    fused = mrcpu.Spec(g, e, negotiation=neg(g.n, present=present, halo=halo, pfac=pfac, fused=True),
                       own_halo=halo_cells + old_path, use_heuristic=True)
    empty = mrcpu.Spec(g, e, negotiation=neg(g.n, pfac=pfac), use_heuristic=True)
    a = mrcpu.astar(unfused, src, dst)
    b = mrcpu.astar(fused, src, dst)
    c = mrcpu.astar(empty, src, dst)
    assert a == b == c
    assert c == (old_path, 4 * 16)
    # negative control: the own halo left in the snapshot is paid for
    leak = mrcpu.Spec(g, e, negotiation=neg(g.n, present=present, halo=halo, pfac=pfac),
                      own_path=old_path, use_heuristic=True)
    assert mrcpu.astar(leak, src, dst)[1] > 4 * 16


def test_negotiated_closed_forms():
    # corridor src - mid - dst, cost 1, planar price 16
    g = mrcpu.Grid.filled(3, 1, 1, 1)
    e = mrcpu.Edges.uniform(3, 1, 1, planar=16, via=160)
    hist = [0, 5, 0]
    # term by term, own copper on mid: 16 + 5 + 3*16*(2-1), then 16
    s = mrcpu.Spec(g, e, negotiation=neg(3, history=hist, present=[0, 2, 0], pfac=3), own_path=[1])
    assert mrcpu.astar(s, 0, 2) == ([0, 1, 2], 85)
    # without the own path the full present count is paid: 16 + 5 + 96 + 16
    s = mrcpu.Spec(g, e, negotiation=neg(3, history=hist, present=[0, 2, 0], pfac=3))
    assert mrcpu.astar(s, 0, 2) == ([0, 1, 2], 133)
    # fused, own copper + own halo counted twice on mid:
    # shared = 48*2 + 48*3 = 240, foreign = 240 - 2*48 = 144 -> 16 + 5 + 144 + 16
    s = mrcpu.Spec(g, e, negotiation=neg(3, history=hist, present=[0, 2, 0], halo=[0, 3, 0], pfac=3, fused=True),
                   own_halo=[1, 1])
    assert mrcpu.astar(s, 0, 2) == ([0, 1, 2], 181)
    # the same state term by term: present 2-0 (no own path) and halo 3-2
    s = mrcpu.Spec(g, e, negotiation=neg(3, history=hist, present=[0, 2, 0], halo=[0, 3, 0], pfac=3),
                   own_halo=[1, 1])
    assert mrcpu.astar(s, 0, 2) == ([0, 1, 2], 16 + 5 + 96 + 48 + 16)
    # a saturated price (MAX-1) on the only way through cannot be completed
    s = mrcpu.Spec(g, e, negotiation=neg(3, history=[0, U32MAX, 0], pfac=1))
    assert mrcpu.astar(s, 0, 2) is None
    # history is never consumed by self-subtraction (negotiated.rs:1647-1652)
    s = mrcpu.Spec(g, e, negotiation=neg(3, history=hist, present=[0, 1, 0], pfac=2, fused=True),
                   own_halo=[1, 1, 1])
    assert mrcpu.astar(s, 0, 2) == ([0, 1, 2], 16 + 5 + 16)


class Lcg:
    """The randomized test's generator (negotiated.rs:6667-6681)."""

    def __init__(self, seed):
        self.s = seed

    def next(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return self.s

    def below(self, bound):
        return ((self.next() >> 32) & 0xFFFFFFFF) % max(bound, 1)


def random_case(rng, case):
    """One case of the randomized comparison (negotiated.rs:6697-6860)."""
    w, h, L = 2 + rng.below(6), 2 + rng.below(6), 1 + rng.below(4)
    n = w * h * L
    cost = np.ones(n, dtype=np.uint32)
    for c in range(n):
        k = rng.below(12)
        cost[c] = MAX if k < 2 else (0 if k == 2 else (1 if k < 5 else 2 + rng.below(2000)))
    g = mrcpu.Grid(w, h, L, cost)
    src, dst = rng.below(n), rng.below(n)
    plane = w * h
    sx, sy, dx, dy = src % plane % w, src % plane // w, dst % plane % w, dst % plane // w
    x0, y0 = rng.below(min(sx, dx) + 1), rng.below(min(sy, dy) + 1)
    x1 = max(sx, dx) + rng.below(w - max(sx, dx))
    y1 = max(sy, dy) + rng.below(h - max(sy, dy))
    pads = []
    for c in range(n):
        cx, cy = c % plane % w, c % plane // w
        inside = x0 <= cx <= x1 and y0 <= cy <= y1
        if inside and (cost[c] == MAX or rng.below(19) == 0) and rng.below(5) == 0:
            pads.append(c)
    for c in (src, dst):
        if cost[c] == MAX:
            pads.append(c)
    if case % 3 == 0:
        xe, ye = [16] * (w - 1), [16] * (h - 1)
    else:
        xe = [1 + rng.below(64) for _ in range(w - 1)]
        ye = [1 + rng.below(64) for _ in range(h - 1)]
    step = 1 + rng.below(160 * 3)
    via = [step if (case % 2 or rng.below(2) == 0) else None for _ in range(L - 1)]
    e = mrcpu.Edges(xe, ye, via)
    pfac = 1 + rng.below(60)
    present = np.zeros(n, dtype=np.uint32)
    halo = np.zeros(n, dtype=np.uint32)
    history = np.zeros(n, dtype=np.uint32)
    own, hc = [], {}
    for c in range(n):
        sp = 1 if rng.below(7) == 0 else 0
        if sp:
            own.append(c)
        sh = 1_000_000 + rng.below(1_000_000) if rng.below(31) == 0 else rng.below(9)
        if sh:
            hc[c] = sh
        fp = (U32MAX - sp - rng.below(4096)) if rng.below(23) == 0 else rng.below(32)
        fh = (U32MAX - sh - rng.below(4096)) if rng.below(23) == 0 else rng.below(64)
        present[c] = fp + sp
        halo[c] = fh + sh
        history[c] = (U32MAX - rng.below(4096)) if rng.below(29) == 0 else rng.below(100_000)
    return g, e, src, dst, pads, (x0, y0, x1, y1), pfac, present, halo, history, own, hc


def test_fused_matches_unfused_randomized():
    rng = Lcg(0x4D595DF4D0F33173)
    seen = {"routed": 0, "none": 0, "saturated": 0}
    for case in range(160):
        g, e, src, dst, pads, win, pfac, present, halo, history, own, hc = random_case(rng, case)
        combined = dict(hc)
        for c in own:
            combined[c] = combined.get(c, 0) + 1
        heur = not bool((g.cost == 0).any())
        kw = dict(pads=pads, window=win, use_heuristic=heur)
        unf = mrcpu.Spec(g, e, negotiation=neg(g.n, history, present, halo, pfac), own_path=own, own_halo=hc, **kw)
        fus = mrcpu.Spec(g, e, negotiation=neg(g.n, history, present, halo, pfac, fused=True), own_halo=combined,
                         **kw)
        a = mrcpu.astar(unf, src, dst)
        b = mrcpu.astar(fus, src, dst)
        assert a == b, case
        # the complete-field search agrees with A* under the negotiated price
        dist, hops = mrcpu.field_lex(unf, [src], forced=[src, dst])
        r = mrcpu.reconstruct(unf, dist, hops, dst)
        if a is None:
            seen["none"] += 1
            assert r is None or int(dist[dst]) >= MAX, case
        else:
            seen["routed"] += 1
            assert r == a[0] and int(dist[dst]) == a[1], case
            seen["saturated"] += int(any(history[c] > 0xFFFF0000 for c in a[0][1:]))
    assert seen["routed"] > 40 and seen["none"] > 5
