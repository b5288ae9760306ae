"""Negotiated-congestion routing loop over the field engine.

Host-side restatement of NegotiatedRouter's negotiation loop
(crates/mr-cpu/src/negotiated.rs:1610-1990) for boards without clearance
(clearance_mm == 0, no via keepout or via spacing, so `present_halo` stays
zero, :1551-1560): PathFinder-style iterations with a growing present factor
`pfac = 1 + iter`, per-cell history bumped by SCALE on over-used cells, and
convergence when no cell is used by two connection groups.

  * more than 16 nets (PARALLEL_NEGOTIATION_THRESHOLD, :118, :1588-1592):
    Jacobi iterations.  The dirty nets (iteration 0, unrouted, or touching a
    cell over-used in the previous iteration, :1627-1638) are routed together
    against a read-only snapshot of present/history, each net pricing its own
    previous copper apart; with 32 or more dirty nets (FUSED_JACOBI_MIN_DIRTY_NETS,
    :122) the snapshot is fused into one shared congestion term (:1647-1667).
    One device batch (mrf_route_negotiated_batch) routes all of them.  The
    results are merged in net-index order (:1779-1823).
  * otherwise: Gauss-Seidel, one net at a time with its old path removed
    from `present` first (:1812-1920), each search a batch of one.

Every search is the window search with a full-board retry
(`route_negotiated(..).or_else(full)`, :1752-1776).  Over-use is counted
across groups (:1924-1945); history += SCALE on over-used cells (:1947-1950);
incremental skipping when more than 8 nets (:1575-1576, :1972-1981).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

import mrfield

SCALE = 16                              # negotiated.rs:71 (COST_SCALE)
MAX_ITERS = 60                          # negotiated.rs:108
PARALLEL_NEGOTIATION_THRESHOLD = 16     # negotiated.rs:118
FUSED_JACOBI_MIN_DIRTY_NETS = 32        # negotiated.rs:122
U32MAX = 0xFFFFFFFF

# route(idx, history, present, pfac, fused, own_paths, own_halos) -> one path
# (sequence of cells) or None per net index in idx; own_paths is None when the
# batch passes no own path, own_halos None when it passes no own halo.
RouteBatch = Callable[[Sequence[int], np.ndarray, np.ndarray, int, bool, Optional[list], Optional[list]],
                      List[Optional[Sequence[int]]]]


@dataclass
class IterTrace:
    """One iteration's settled state (the IterSnapshot of negotiated.rs:1952-1962)."""
    iter: int
    pfac: int
    dirty: List[int]
    fused: bool
    paths: List[List[int]]
    overused_cells: List[int]
    any_overuse: bool


@dataclass
class Negotiated:
    paths: List[List[int]]              # empty == unrouted
    history: np.ndarray
    present: np.ndarray
    iterations: int
    converged: bool
    trace: List[IterTrace] = field(default_factory=list)
    routed_searches: int = 0            # searches issued over all iterations


def _concat(paths: List[np.ndarray]):
    """All path cells in net order and the net of each."""
    lens = np.fromiter((len(p) for p in paths), dtype=np.int64, count=len(paths))
    if not lens.sum():
        return np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64), lens
    return np.concatenate([p for p in paths if len(p)]), np.repeat(np.arange(len(paths)), lens), lens


def _overuse(paths: List[np.ndarray], groups: np.ndarray, present: Optional[np.ndarray] = None, flat=None):
    """The settle step's over-use in numpy (the loop runs the library's
    mrf_negotiation_settle; tests hold the two equal).

    Cells occupied by two or more distinct groups, ascending.  `present`
    (the per-cell path count) narrows the candidates to cells held twice;
    when every net is its own group those are exactly the over-used cells.
    flat: _concat(paths) when the caller has it."""
    if present is not None and len(np.unique(groups)) == len(groups):
        return np.flatnonzero(present >= 2)
    cells, net, _ = flat if flat is not None else _concat(paths)
    if not len(cells):
        return np.zeros(0, dtype=np.int64)
    if present is not None:
        keep = present[cells] >= 2
        cells, net = cells[keep], net[keep]
    # distinct (cell, group) pairs, then cells seen with two groups
    pairs = np.unique(cells * (int(groups.max()) + 1) + groups[net])
    c = pairs // (int(groups.max()) + 1)
    return np.unique(c[1:][c[1:] == c[:-1]])


def _touches(paths: List[np.ndarray], mask: np.ndarray, flat=None) -> np.ndarray:
    """Per path: whether any of its cells is set in mask.  flat: _concat(paths)
    when the caller has it."""
    cells, net, lens = flat if flat is not None else _concat(paths)
    hit = np.zeros(len(paths), dtype=bool)
    if len(cells):
        hit[np.unique(net[mask[cells]])] = True
    return hit


def negotiate(route: RouteBatch, n_cells: int, n_nets: int, groups: Optional[Sequence[int]] = None,
              max_iters: int = MAX_ITERS, keep_trace: bool = False, history0=None,
              on_iter: Optional[Callable[[int, int], None]] = None) -> Negotiated:
    """Run the negotiation loop with `route` as the search (see RouteBatch).

    history0: starting history plane (zeros, as the reference starts, when
    None).  on_iter(iteration, searches) is called after each iteration."""
    grp = np.asarray(groups if groups is not None else range(n_nets), dtype=np.int64)
    history = (np.array(history0, dtype=np.uint32) if history0 is not None
               else np.zeros(n_cells, dtype=np.uint32))
    present = np.zeros(n_cells, dtype=np.uint32)
    empty = np.zeros(0, dtype=np.int64)
    paths: List[np.ndarray] = [empty for _ in range(n_nets)]
    touch = np.zeros(n_nets, dtype=bool)  # net's settled path holds an over-used cell
    incremental = n_nets > 8
    use_parallel = n_nets > PARALLEL_NEGOTIATION_THRESHOLD
    out = Negotiated([], history, present, 0, False)

    for it in range(max_iters):
        pfac = 1 + it
        dirty: List[int] = []
        fused = False
        if use_parallel:
            if it == 0:
                dirty = list(range(n_nets))
            else:
                dirty = [i for i in range(n_nets) if not len(paths[i]) or touch[i]]
            fused = len(dirty) >= FUSED_JACOBI_MIN_DIRTY_NETS
            if dirty:
                if fused:
                    # own copper counted once in the own halo stamp (:1709-1716)
                    own_paths, own_halos = None, [paths[i] for i in dirty]
                else:
                    own_paths, own_halos = [paths[i] for i in dirty], None
                got = route(dirty, history, present, pfac, fused, own_paths, own_halos)
                out.routed_searches += len(dirty)
                # Merge (:1779-1823).  Every path is counted in present and
                # holds each cell once, so present after the reference's
                # in-order saturating updates (remove the old path, add the
                # new one, net by net) is the per-cell count of the settled
                # paths, which the settle step below recounts.
                for i, p in zip(dirty, got):
                    paths[i] = np.asarray(p, dtype=np.int64) if p is not None else empty
        else:
            for i in range(n_nets):
                if incremental and it > 0 and len(paths[i]) and not touch[i]:
                    continue
                dirty.append(i)
                if len(paths[i]):
                    present[paths[i]] -= 1
                paths[i] = empty
                got = route([i], history, present, pfac, False, None, None)[0]
                out.routed_searches += 1
                if got is not None:
                    paths[i] = np.asarray(got, dtype=np.int64)
                    present[paths[i]] += 1
        # Settle (:1924-1950) in the library: present recounted, cells held by
        # two or more groups, history += SCALE on them, and which nets touch
        # one (the next iteration's dirty set).
        off, cells = _flat(paths)
        over, touched = mrfield.negotiation_settle(off, cells, grp, SCALE, present, history)
        any_overuse = len(over) > 0
        if keep_trace:
            out.trace.append(IterTrace(it, pfac, dirty, fused, [p.tolist() for p in paths], over.tolist(),
                                       any_overuse))
        if incremental:
            touch = touched
        out.iterations = it + 1
        if on_iter is not None:
            on_iter(it, len(dirty))
        if not any_overuse:
            out.converged = True
            break
    out.paths = [p.tolist() for p in paths]
    return out


def net_window(w: int, h: int, src: int, dst: int, pads: Sequence[int] = ()):
    """Search window of a net (negotiated.rs:568-594): the bounding box of its
    endpoints and own pads grown by max(16, ceil(0.3 * span)) on each side,
    clipped to the board; (x0, y0, x1, y1) inclusive."""
    plane = w * h

    def xy(c):
        r = c % plane
        return r % w, r // w

    (sx, sy), (dx, dy) = xy(src), xy(dst)
    x0, y0, x1, y1 = min(sx, dx), min(sy, dy), max(sx, dx), max(sy, dy)
    for p in pads:
        px, py = xy(p)
        x0, y0, x1, y1 = min(x0, px), min(y0, py), max(x1, px), max(y1, py)
    span = max(x1 - x0, y1 - y0)
    margin = max(16, (3 * span + 9) // 10)
    return (max(x0 - margin, 0), max(y0 - margin, 0), min(x1 + margin, max(w - 1, 0)),
            min(y1 + margin, max(h - 1, 0)))


def negotiate_board(board, nets, groups=None, max_iters: int = MAX_ITERS, keep_trace: bool = False,
                    history0=None, on_iter=None) -> Negotiated:
    """The negotiation loop on a resident board, each net searched in its
    net_window with a full-board retry."""
    g = board.grid
    wins = [net_window(g.w, g.h, n.src, n.dst, getattr(n, "pads", ()) or ()) for n in nets]
    return negotiate(board_router(board, nets, wins), g.n, len(nets), groups, max_iters, keep_trace, history0,
                     on_iter)


def _flat(per_net):
    """Per-net cell sequences as one (offsets, cells) pair."""
    off = np.zeros(len(per_net) + 1, dtype=np.uint64)
    np.cumsum(np.fromiter((len(p) for p in per_net), dtype=np.uint64, count=len(per_net)), out=off[1:])
    cells = (np.concatenate([np.asarray(p, dtype=np.int64) for p in per_net]) if len(per_net) and off[-1]
             else np.zeros(0, dtype=np.int64))
    return off, cells


def board_router(board, nets, windows=None) -> RouteBatch:
    """RouteBatch over a resident board (mrfield.Board.route_negotiated)."""

    wall = np.asarray(windows, dtype=np.uint32).reshape(-1, 4) if windows is not None else None

    def route(idx, history, present, pfac, fused, own_paths, own_halos):
        sub = [nets[i] for i in idx]
        wins = wall[np.asarray(idx, dtype=np.int64)] if wall is not None else None
        flat = lambda per_net: None if per_net is None else _flat(per_net)
        res = board.route_negotiated(sub, history, present, pfac, fused=fused, own_paths=flat(own_paths),
                                     own_halos=flat(own_halos), windows=wins, retry_full=True, arrays=True)
        cells, off = res.cells, res.offsets
        return [cells[off[k]:off[k + 1]] if res.routed[k] else None for k in range(len(sub))]

    return route
