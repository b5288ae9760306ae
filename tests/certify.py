"""Bellman-optimality certificate for complete canonical fields (numpy).

A labeling L over cells (packed (dist << 32) | hops, INF = unreachable) is the
exact lexicographic shortest-path labeling iff
  * every passable source has L == 0, and
  * every other cell has L == min over legal in-edges (u -> v) of
    L(u) + (step(u, v), 1), with INF when no candidate exists.
Every step weight (c, 1) is strictly positive in lexicographic order, so a
finite label chain strictly descends in hops to a source (achievable) and no
edge improves it (optimal); the system has exactly one solution.  The check
is size-independent and runs at the full benchmark sizes where the oracle's
per-field Dijkstra would be slow.  ``step_prices`` is itself pinned against
the oracle's mrc_step_cost on sampled edges (tests/test_certify.py).
"""
from __future__ import annotations

import numpy as np

MAX = 0xFFFFFFFF
INF = np.uint64(0xFFFFFFFFFFFFFFFF)


class Problem:
    def __init__(self, w, h, L, cost, board=None, via_forbidden=None, x=None, y=None, via=None,
                 additive=None, halo=None, pen=0):
        self.w, self.h, self.L = w, h, L
        self.n = w * h * L
        self.cost = np.asarray(cost, dtype=np.uint64).reshape(L, h, w)
        self.board = None if board is None else np.asarray(board, dtype=np.uint8).reshape(L, h, w)
        self.vf = None if via_forbidden is None else np.asarray(via_forbidden, dtype=bool).reshape(L, h, w)
        self.x = np.asarray(x, dtype=np.uint64)
        self.y = np.asarray(y, dtype=np.uint64)
        self.via = [None if v is None else np.uint64(v) for v in via]
        self.add = None if additive is None else np.asarray(additive, dtype=np.uint64).reshape(L, h, w)
        self.halo = None if halo is None else np.asarray(halo, dtype=np.int64).reshape(L, h, w)
        self.pen = np.uint64(pen)


def _cells(idx, shape):
    m = np.zeros(int(np.prod(shape)), dtype=bool)
    if len(idx):
        m[np.asarray(idx, dtype=np.int64)] = True
    return m.reshape(shape)


def field_terms(p: Problem, window, pads=(), via_pads=(), forced=(), own_id=-1):
    """Per-cell passability, enter weight and additive price of one field."""
    shape = (p.L, p.h, p.w)
    obst = p.cost == MAX
    pad = _cells(pads, shape) & obst
    x0, y0, x1, y1 = window
    inwin = np.zeros(shape, dtype=bool)
    inwin[:, y0:y1 + 1, x0:x1 + 1] = True
    trace = (p.board & 1).astype(bool) if p.board is not None else np.zeros(shape, dtype=bool)
    passable = inwin & ~trace & (~obst | pad | _cells(forced, shape))
    wgt = np.where(pad, np.uint64(1), p.cost)
    add = np.zeros(shape, dtype=np.uint64) if p.add is None else p.add.copy()
    if p.halo is not None:
        add = add + np.where((p.halo != -1) & (p.halo != own_id), p.pen, np.uint64(0))
    vp = _cells(via_pads, shape)
    return passable, wgt, add, vp


def step_prices(e, wgt, add):
    """min(e*w + a, MAX-1) on uint64 arrays (e*w + a < 2^64 for u32 inputs)."""
    return np.minimum(e * wgt + add, np.uint64(MAX - 1))


def in_edges(p: Problem, passable, wgt, add, vp):
    """Yield (pred_slices, dest_slices, legal_mask, step) for the six directions."""
    b = p.board
    # planar: dest v at [.., y, x], pred u = v shifted; legality symmetric
    def planar(axis, sign):
        if axis == 2:
            n = p.w
            gaps = p.x
        else:
            n = p.h
            gaps = p.y
        if n < 2:
            return None
        sl_v = [slice(None)] * 3
        sl_u = [slice(None)] * 3
        if sign > 0:   # u at index-1 (coming from the lower side)
            sl_v[axis] = slice(1, n)
            sl_u[axis] = slice(0, n - 1)
            gap = gaps[: n - 1]
        else:          # u at index+1
            sl_v[axis] = slice(0, n - 1)
            sl_u[axis] = slice(1, n)
            gap = gaps[: n - 1]
        sl_v, sl_u = tuple(sl_v), tuple(sl_u)
        legal = passable[sl_v] & passable[sl_u]
        if b is not None:
            bu, bv = b[sl_u], b[sl_v]
            if axis == 2:
                sb, db = (0x10, 0x08) if sign > 0 else (0x08, 0x10)
            else:
                sb, db = (0x20, 0x04) if sign > 0 else (0x04, 0x20)
            legal &= ((bu & sb) == 0) & ((bv & db) == 0)
        shape_e = [1, 1, 1]
        shape_e[axis] = n - 1
        e = gap.reshape(shape_e)
        step = step_prices(e, wgt[sl_v], add[sl_v])
        return sl_u, sl_v, legal, step

    out = [planar(2, +1), planar(2, -1), planar(1, +1), planar(1, -1)]
    for sign in (+1, -1):
        if p.L < 2:
            continue
        if sign > 0:
            sl_v, sl_u = (slice(1, p.L), slice(None), slice(None)), (slice(0, p.L - 1), slice(None), slice(None))
        else:
            sl_v, sl_u = (slice(0, p.L - 1), slice(None), slice(None)), (slice(1, p.L), slice(None), slice(None))
        allowed = np.array([v is not None for v in p.via], dtype=bool).reshape(-1, 1, 1)
        price = np.array([0 if v is None else int(v) for v in p.via], dtype=np.uint64).reshape(-1, 1, 1)
        legal = passable[sl_v] & passable[sl_u] & allowed
        if b is not None:
            legal &= ((b[sl_u] & 2) == 0) & ((b[sl_v] & 2) == 0)
        if p.vf is not None:
            legal &= (~p.vf[sl_u] | vp[sl_u]) & (~p.vf[sl_v] | vp[sl_v])
        step = step_prices(price, wgt[sl_v], add[sl_v])
        out.append((sl_u, sl_v, legal, step))
    return [o for o in out if o is not None]


def certify(p: Problem, dist, hops, window, sources, pads=(), via_pads=(), forced=(), own_id=-1):
    """Return None when (dist, hops) is the exact canonical field, else a message."""
    shape = (p.L, p.h, p.w)
    d = np.asarray(dist, dtype=np.uint64).reshape(shape)
    h = np.asarray(hops, dtype=np.uint64).reshape(shape)
    lab = np.where(d == MAX, INF, (d << np.uint64(32)) | h)
    passable, wgt, add, vp = field_terms(p, window, pads, via_pads, forced, own_id)
    best = np.full(shape, INF, dtype=np.uint64)
    for sl_u, sl_v, legal, step in in_edges(p, passable, wgt, add, vp):
        lu = lab[sl_u]
        du = lu >> np.uint64(32)
        ok = legal & (lu != INF) & (du + step < np.uint64(MAX))
        cand = np.where(ok, lu + (step << np.uint64(32)) + np.uint64(1), INF)
        best[sl_v] = np.minimum(best[sl_v], cand)
    src = _cells([s for s in sources if s < p.n], shape) & passable
    expect = np.where(src, np.uint64(0), best)
    bad = np.flatnonzero(expect.ravel() != lab.ravel())
    if len(bad):
        i = int(bad[0])
        return (f"{len(bad)} cells violate the Bellman certificate; first cell {i}: "
                f"label {int(lab.ravel()[i]):#x} expected {int(expect.ravel()[i]):#x}")
    return None
