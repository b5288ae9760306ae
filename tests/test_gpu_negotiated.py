"""GPU: negotiated-congestion batches (mrf_route_negotiated_batch) against the
CPU oracle's route_negotiated restatement (negotiated.rs:2746-2856).

Each case builds a negotiation snapshot on a synthetic board: the nets'
previous paths make `present`, radius-1 footprints around them make
`present_halo` and the own halo multisets, and `history` is the board's
congestion-history plane plus bumps.  Paths and search costs must be equal,
byte for byte, to the oracle's window-then-full-board searches with the same
pricing.  Tiles and both narrow / wide step planes are exercised (saturating
snapshots force the wide planes).
"""
from __future__ import annotations

import numpy as np
import pytest

import mrcpu
import mrfield

pytestmark = pytest.mark.gpu

U32MAX = 0xFFFFFFFF


def oracle_parts(s):
    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost, board=g.board)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    return og, oe


def footprint(g, path):
    """Planar radius-1 boxes around the path cells on their own layers (each
    visit counted), skipping obstacles: the clearance halo at pitch 1."""
    plane = g.w * g.h
    out = {}
    for c in path:
        l, r = divmod(c, plane)
        cy, cx = divmod(r, g.w)
        for y in range(max(0, cy - 1), min(g.h, cy + 2)):
            for x in range(max(0, cx - 1), min(g.w, cx + 2)):
                k = l * plane + y * g.w + x
                if g.cost[k] != mrfield.MAX:
                    out[k] = out.get(k, 0) + 1
    return out


def snapshot(s, eng, nets, seed, with_halo, saturate=False):
    """(history, present, present_halo, own paths, own halos) from alone paths."""
    g = s.grid
    board = mrfield.Board(eng, g, s.edges, None)
    try:
        prev = board.solve(nets)
    finally:
        board.close()
    rng = np.random.default_rng(seed)
    n = g.n
    history = (np.asarray(s.additive, dtype=np.uint64) if s.additive is not None
               else np.zeros(n, dtype=np.uint64))
    bump = rng.random(n) < 0.05
    history[bump] += (16 * rng.integers(1, 6, size=int(bump.sum()))).astype(np.uint64)
    if saturate:
        hot = rng.random(n) < 0.01
        history[hot] = (U32MAX - rng.integers(0, 4096, size=int(hot.sum()))).astype(np.uint64)
    history = np.minimum(history, U32MAX).astype(np.uint32)
    present = np.zeros(n, dtype=np.uint32)
    halo = np.zeros(n, dtype=np.uint32)
    own_paths, own_halos = [], []
    for p in prev:
        cells = p[0] if p is not None else []
        own_paths.append(cells)
        for c in cells:
            present[c] += 1
        h = footprint(g, cells) if with_halo else {}
        own_halos.append(h)
        for c, k in h.items():
            halo[c] += k
    if saturate:
        hot = rng.random(n) < 0.01
        present[hot] = (U32MAX - rng.integers(0, 4096, size=int(hot.sum()))).astype(np.uint32)
    return history, present, (halo if with_halo else None), own_paths, own_halos


def windows_for(s, nets):
    g = s.grid
    og, _ = oracle_parts(s)
    return [tuple(mrcpu.net_window(og, n.src, n.dst)) for n in nets]


def oracle_batch(s, nets, wins, history, present, halo, pfac, fused, own_paths, own_halos, use_own_path):
    og, oe = oracle_parts(s)
    ng = mrcpu.Negotiation(history=history, present=present, present_halo=halo, pfac=pfac, fused=fused)
    tmpl = mrcpu.Spec(og, oe, negotiation=ng)
    onets = []
    for n, p, h in zip(nets, own_paths, own_halos):
        onets.append(mrcpu.Net(n.src, n.dst, own_path=(p if use_own_path else None), own_halo=h))
    paths, costs = mrcpu.alone_batch(tmpl, onets, wins, retry_full=True, threads=8)
    return [None if p is None else (p, c) for p, c in zip(paths, costs)]


CASES = [
    # (side, layers, nets, seed, fused, with_halo, saturate, pfac, tiles)
    (64, 2, 24, 31, False, False, False, 3, (16, 16, 256, 32)),
    (96, 4, 16, 32, True, True, False, 5, (16, 16, 256, 32)),
    (80, 3, 20, 33, False, True, False, 9, (8, 4, 64, 32)),
    (64, 2, 16, 34, True, True, True, 60, (16, 16, 256, 32)),
    (72, 4, 12, 35, False, False, True, 2, (32, 16, 256, 32)),
    # the wave form (one wavefront per tile; overrides applied per tile)
    (64, 2, 24, 31, False, False, False, 3, (16, 16, 64, 32)),
    (96, 4, 16, 32, True, True, False, 5, (16, 16, 64, 32)),
    (80, 6, 20, 33, False, True, False, 9, (16, 16, 64, 32)),
]


@pytest.mark.parametrize("mode", [mrfield.MODE_PLAIN, mrfield.MODE_QUEUE])
@pytest.mark.parametrize("side,layers,n,seed,fused,with_halo,saturate,pfac,tiles", CASES)
def test_negotiated_batch_matches_oracle(mode, side, layers, n, seed, fused, with_halo, saturate, pfac, tiles):
    s = mrfield.synth_board(side, layers, n, seed)
    nets = s.nets()
    eng = mrfield.Engine(0)
    try:
        eng.set_tiles(*tiles)
        eng.set_mode(mode)
        history, present, halo, own_paths, own_halos = snapshot(s, eng, nets, seed, with_halo, saturate)
        # the fused batch passes no own path and counts the copper in the own halo
        if fused:
            halos = []
            for p, h in zip(own_paths, own_halos):
                hh = dict(h)
                for c in p:
                    hh[c] = hh.get(c, 0) + 1
                halos.append(hh)
        else:
            halos = own_halos
        wins = windows_for(s, nets)
        board = mrfield.Board(eng, s.grid, s.edges, None)
        try:
            got = board.route_negotiated(nets, history, present, pfac, present_halo=halo, fused=fused,
                                         own_paths=None if fused else own_paths, own_halos=halos, windows=wins)
        finally:
            board.close()
        exp = oracle_batch(s, nets, wins, history, present, halo, pfac, fused, own_paths, halos,
                           use_own_path=not fused)
        assert got == exp
        assert sum(p is not None for p in got) >= n // 2
    finally:
        eng.close()


def test_self_occupancy_prices_like_an_empty_snapshot():
    """negotiated.rs:6535-6660 on the device: a net alone on the board whose
    own copper and own halo are the whole snapshot routes as with no
    congestion at all."""
    s = mrfield.synth_board(64, 2, 8, 36)
    nets = s.nets()
    eng = mrfield.Engine(0)
    try:
        eng.set_tiles(16, 16, 256, 32)
        board = mrfield.Board(eng, s.grid, s.edges, None)
        try:
            alone = board.solve(nets)
            n = s.grid.n
            hist = np.zeros(n, dtype=np.uint32)
            for i, net in enumerate(nets):
                if alone[i] is None:
                    continue
                path = alone[i][0]
                present = np.zeros(n, dtype=np.uint32)
                present[path] += 1
                h = footprint(s.grid, path)
                halo = np.zeros(n, dtype=np.uint32)
                for c, k in h.items():
                    halo[c] += k
                empty = board.route_negotiated([net], hist, np.zeros(n, dtype=np.uint32), 7)
                own = board.route_negotiated([net], hist, present, 7, present_halo=halo, own_paths=[path],
                                             own_halos=[h])
                fusedp = dict(h)
                for c in path:
                    fusedp[c] = fusedp.get(c, 0) + 1
                fused = board.route_negotiated([net], hist, present, 7, present_halo=halo, fused=True,
                                               own_halos=[fusedp])
                assert own == fused == empty, i
                assert empty[0] == alone[i]  # zero history: the alone price
        finally:
            board.close()
    finally:
        eng.close()


def test_negotiated_batch_rejects_bad_planes():
    s = mrfield.synth_board(32, 2, 4, 37)
    eng = mrfield.Engine(0)
    try:
        board = mrfield.Board(eng, s.grid, s.edges, None)
        try:
            n = s.grid.n
            with pytest.raises(mrfield.MrfError):
                board.route_negotiated(s.nets(), np.zeros(n - 1, dtype=np.uint32), np.zeros(n, dtype=np.uint32), 1)
            with pytest.raises(mrfield.MrfError):
                board.route_negotiated(s.nets()[:1], np.zeros(n, dtype=np.uint32), np.zeros(n, dtype=np.uint32), 1,
                                       own_paths=[[n + 5]])
        finally:
            board.close()
    finally:
        eng.close()


@pytest.mark.parametrize("threads", [256, 64], ids=["workgroup", "wave"])
@pytest.mark.parametrize("side,layers,n,seed,iters", [(64, 4, 24, 11, 60), (40, 1, 40, 3, 6), (24, 2, 12, 5, 10)])
def test_negotiation_loop_matches_oracle_loop(side, layers, n, seed, iters, threads):
    """The whole loop (negotiate.py) with device batches equals the loop with
    the oracle's searches, iteration by iteration: Jacobi with term-by-term
    snapshots (converging), Jacobi with fused snapshots, and Gauss-Seidel."""
    import negotiate as ng
    from cases import oracle_negotiation_router

    s = mrfield.synth_board(side, layers, n, seed)
    nets = s.nets()
    g = s.grid
    wins = [ng.net_window(g.w, g.h, x.src, x.dst, x.pads) for x in nets]
    eng = mrfield.Engine(0)
    try:
        eng.set_tiles(16, 16, threads, 32)
        board = mrfield.Board(eng, g, s.edges, None)
        try:
            gpu = ng.negotiate(ng.board_router(board, nets, wins), g.n, len(nets), max_iters=iters, keep_trace=True)
        finally:
            board.close()
    finally:
        eng.close()
    cpu = ng.negotiate(oracle_negotiation_router(s, nets, wins), g.n, len(nets), max_iters=iters, keep_trace=True)
    assert (gpu.iterations, gpu.converged) == (cpu.iterations, cpu.converged)
    for a, b in zip(gpu.trace, cpu.trace):
        assert a == b, a.iter
    assert gpu.paths == cpu.paths
    assert np.array_equal(gpu.history, cpu.history)
