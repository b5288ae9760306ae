"""Shared builders for the transcribed reference vectors and seeded batteries."""
from __future__ import annotations

import json
import os

import numpy as np

import mrcpu

HERE = os.path.dirname(os.path.abspath(__file__))
MAX = 0xFFFFFFFF


def load_vectors():
    with open(os.path.join(HERE, "golden", "reference_vectors.json")) as f:
        return json.load(f)


def xorshift64(seed: int):
    """Reference stress-test generator (mr-metal/src/lib.rs:1655-1661)."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state ^= (state << 13) & 0xFFFFFFFFFFFFFFFF
        state ^= state >> 7
        state ^= (state << 17) & 0xFFFFFFFFFFFFFFFF
        yield state


def grid_from_case(c):
    """Build (grid, src, dst) from one reference_vectors.json case."""
    if "ascii" in c:
        rows = c["ascii"]
        h, w = len(rows), max(len(r) for r in rows)
        cost = np.ones(w * h, dtype=np.uint32)
        src = dst = None
        for y, row in enumerate(rows):
            for x, ch in enumerate(row):
                i = y * w + x
                if ch == "#":
                    cost[i] = MAX
                elif ch == "S":
                    src = i
                elif ch == "T":
                    dst = i
        return mrcpu.Grid(w, h, 1, cost), src, dst
    w, h, L = c["w"], c["h"], c["layers"]
    if "cost" in c:
        cost = np.array(c["cost"], dtype=np.uint32)
    else:
        cost = np.full(w * h * L, c["fill"], dtype=np.uint32)
    for (x0, y0, x1, y1, l) in c.get("obstacle_rects", []):
        for y in range(y0, y1 + 1):
            for x in range(x0, x1 + 1):
                cost[(l * h + y) * w + x] = MAX
    for i, v in c.get("cost_set", []):
        cost[i] = v
    board = None
    if "board_set" in c:
        board = np.zeros(w * h * L, dtype=np.uint8)
        for i, bits in c["board_set"]:
            board[i] |= bits
    g = mrcpu.Grid(w, h, L, cost, board=board)

    def cell(key_xyz, key_idx):
        if key_xyz in c:
            x, y, l = c[key_xyz]
            return (l * h + y) * w + x
        return c.get(key_idx)

    return g, cell("src", "src_idx"), cell("dst", "dst_idx")


def edges_for_case(c, g):
    via = c.get("via", [160] * (g.layers - 1))
    return mrcpu.Edges([16] * (g.w - 1), [16] * (g.h - 1), via)


def weighted_stress():
    """mr-metal/src/lib.rs:1652-1701: 17x11x2 grid, 24 sources, 24 nets."""
    w, h, L = 17, 11, 2
    n = w * h * L
    rng = xorshift64(0xD1B54A32D192ED03)
    cost = np.zeros(n, dtype=np.uint32)
    for i in range(n):
        r = next(rng)
        cost[i] = MAX if r % 13 == 0 else 1 + (r % 31)
    sources = []
    while len(sources) < 24:
        s = next(rng) % n
        if cost[s] != MAX:
            sources.append(int(s))
    nets = [(sources[i], sources[(i * 7 + 5) % 24]) for i in range(24)]
    return mrcpu.Grid(w, h, L, cost), sources, nets


HANAN_X = [0.0, 0.1, 0.8, 0.85, 2.5, 4.0, 4.2]
HANAN_Y = [0.0, 0.4, 0.45, 3.0, 3.2]


def isolated_stress_case(case: int):
    """mr-metal/src/lib.rs:1882-1933: 7x5x3 Hanan grid, 8 nets, restricted vias."""
    w, h, L = 7, 5, 3
    n = w * h * L
    rng = xorshift64(0xA0761D6478BD642F ^ case)
    cost = np.zeros(n, dtype=np.uint32)
    for i in range(n):
        r = next(rng)
        cost[i] = MAX if r % 11 == 0 else (0 if r % 7 == 0 else 1 + (r % 19))
    nets = []
    for _ in range(8):
        src = int(next(rng) % n)
        dst = int(next(rng) % n)
        pads = []
        if cost[src] == MAX:
            pads.append(src)
        if cost[dst] == MAX and dst != src:
            pads.append(dst)
        nets.append(mrcpu.Net(src, dst, pads))
    ex = mrcpu.axis_edge_costs(HANAN_X, w)
    ey = mrcpu.axis_edge_costs(HANAN_Y, h)
    if case % 2 == 0:
        via = [23, 23]      # with_allowed_steps(3, 23, [(0,1),(1,2)])
    else:
        via = [None, 0]     # with_allowed_steps(3, 0, [(1,2)])
    return mrcpu.Grid(w, h, L, cost), nets, mrcpu.Edges(ex, ey, via)


def oracle_parts(s):
    """Oracle grid and edge prices of a synthetic board (mrfield.Synth)."""
    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost, board=g.board)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    return og, oe


def oracle_negotiation_router(s, nets, windows, threads=8):
    """The negotiation loop's search (RouteBatch of metal-route_amd/negotiate.py)
    answered by the oracle's route_negotiated restatement."""
    og, oe = oracle_parts(s)

    def route(idx, history, present, pfac, fused, own_paths, own_halos):
        ng = mrcpu.Negotiation(history=np.array(history, dtype=np.uint32), present=np.array(present, dtype=np.uint32),
                               present_halo=None, pfac=pfac, fused=fused)
        tmpl = mrcpu.Spec(og, oe, negotiation=ng)
        onets = []
        for k, i in enumerate(idx):
            onets.append(mrcpu.Net(nets[i].src, nets[i].dst, pads=list(nets[i].pads), via_pads=list(nets[i].via_pads),
                                   own_path=own_paths[k] if own_paths is not None else None,
                                   own_halo=own_halos[k] if own_halos is not None else []))
        paths, _ = mrcpu.alone_batch(tmpl, onets, [windows[i] for i in idx], retry_full=True, threads=threads)
        return paths

    return route
