"""Serving workload: Poisson arrivals of solve requests over resident boards.

Boards: sides {128, 256, 512, 1024} (each size at least once), 2..8 layers,
built by the 512x512x4 generator (history plane, rectangles with clearance
halos, convex outline) at that size.  A request is one board and 16..512 of
its two-pin nets.  Requests arrive as a Poisson process at rate `lam`.

Two forms serve the same arrival trace:
  plain  one request at a time (FIFO), each solved by mrf_solve on its
         resident board; a request waits for the one before it.  Its service
         times are measured one by one and the queue is replayed on them.
  serve  the serving loop (mrf_server_*): requests are submitted at their
         arrival times from a separate thread, admitted by the label budget
         and relaxed together with whatever else is on the device.

Both report nets/s over the trace (first arrival to last completion) and the
p50 / p99 of submit-to-paths latency.
"""
from __future__ import annotations

import threading
import time
from typing import List, Tuple

import numpy as np

import mrfield

SIDES = (128, 256, 512, 1024)


def make_workload(seed: int, n_boards: int, n_req: int, pool: int = 512, sides=SIDES):
    rng = np.random.default_rng(seed)
    boards = []
    for b in range(n_boards):
        side = sides[b % len(sides)]
        layers = int(rng.integers(2, 9))
        boards.append(mrfield.synth_board(side, layers, pool, seed * 1000 + b))
    nets = [s.nets() for s in boards]
    reqs = []
    for _ in range(n_req):
        b = int(rng.integers(0, n_boards))
        k = int(rng.integers(16, pool + 1))
        idx = np.sort(rng.choice(pool, size=k, replace=False))
        reqs.append((b, [nets[b][i] for i in idx]))
    return boards, reqs


def arrivals(seed: int, n: int, lam: float) -> np.ndarray:
    rng = np.random.default_rng(seed + 7)
    return np.cumsum(rng.exponential(1.0 / lam, size=n))


def plain_service_times(eng, boards, reqs) -> Tuple[List[float], list]:
    """Solve every request alone (resident boards); wall time of each solve."""
    dev = [mrfield.Board(eng, s.grid, s.edges, s.pricing()) for s in boards]
    times, results = [], []
    dev[reqs[0][0]].solve_arrays(reqs[0][1])  # warm-up
    for b, nets in reqs:
        t0 = time.perf_counter()
        r = dev[b].solve_arrays(nets)
        times.append(time.perf_counter() - t0)
        results.append(r)
    for d in dev:
        d.close()
    return times, results


def replay_fifo(arr: np.ndarray, service: List[float]):
    """Single FIFO server: latency of each request and the makespan."""
    t = 0.0
    lat = []
    for a, s in zip(arr, service):
        start = max(t, a)
        t = start + s
        lat.append(t - a)
    return np.asarray(lat), t - arr[0]


def run_server(srv, board_ids, reqs, arr: np.ndarray):
    """Submit at the arrival times from a thread; returns latencies (s), makespan (s), results."""
    rids = [None] * len(reqs)
    t_sub = [0.0] * len(reqs)
    t0 = time.perf_counter()

    def submitter():
        for i, (b, nets) in enumerate(reqs):
            delay = t0 + arr[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            t_sub[i] = time.perf_counter()
            rids[i] = srv.submit(board_ids[b], nets)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    lat, done, results, peers = [], [], [], []
    for i, rid in enumerate(rids):
        paths, st = srv.wait(rid, arrays=True)
        results.append(paths)
        lat.append(st["total_ms"] * 1e-3)
        done.append(t_sub[i] + st["total_ms"] * 1e-3)
        peers.append(st["peers_max"])
    makespan = max(done) - (t0 + arr[0])
    return np.asarray(lat), makespan, results, peers


def same_paths(a, b) -> bool:
    return (np.array_equal(a.routed, b.routed) and np.array_equal(a.cost, b.cost)
            and np.array_equal(a.offsets, b.offsets) and np.array_equal(a.cells, b.cells))


def run(seed: int = 1, n_boards: int = 8, n_req: int = 96, loads=(0.5, 1.0, 2.0, 4.0),
        tiles=(16, 16, 256, 32), label_gib: int = 96, admit_below_tiles: int = 0, admit_small_tiles: int = 0,
        sides=SIDES):
    boards, reqs = make_workload(seed, n_boards, n_req, sides=tuple(sides))
    total_nets = sum(len(n) for _, n in reqs)
    eng = mrfield.Engine(0)
    eng.set_tiles(*tiles)
    eng.set_mode(mrfield.MODE_PLAIN)
    service, plain_res = plain_service_times(eng, boards, reqs)
    eng.close()
    mu = 1.0 / float(np.mean(service))  # plain form's capacity, requests/s
    srv = mrfield.Server(0, label_bytes=label_gib << 30, max_layers=8, tile_w=tiles[0], tile_h=tiles[1],
                         admit_below_tiles=admit_below_tiles, admit_small_tiles=admit_small_tiles,
                         threads=tiles[2])
    ids = [srv.add_board(s.grid, s.edges, s.pricing()) for s in boards]
    # warm-up: one request through the server
    srv.wait(srv.submit(ids[reqs[0][0]], reqs[0][1]))
    rows = []
    equal = True
    for load in loads:
        lam = load * mu
        arr = arrivals(seed, len(reqs), lam)
        p_lat, p_span = replay_fifo(arr, service)
        s_lat, s_span, s_res, peers = run_server(srv, ids, reqs, arr)
        equal = equal and all(same_paths(x, y) for x, y in zip(plain_res, s_res))
        classes = {}
        for name, sides in (("small_128_256", (128, 256)), ("large_512_1024", (512, 1024))):
            sel = [i for i, (b, _) in enumerate(reqs) if boards[b].grid.w in sides]
            if not sel:
                continue
            classes[name] = {
                "requests": len(sel),
                "nets": int(sum(len(reqs[i][1]) for i in sel)),
                "plain_p50_ms": round(float(np.percentile(p_lat[sel], 50)) * 1e3, 2),
                "plain_p99_ms": round(float(np.percentile(p_lat[sel], 99)) * 1e3, 2),
                "serve_p50_ms": round(float(np.percentile(s_lat[sel], 50)) * 1e3, 2),
                "serve_p99_ms": round(float(np.percentile(s_lat[sel], 99)) * 1e3, 2),
            }
        rows.append({
            "load": load,
            "by_board_size": classes,
            "lambda_req_s": round(lam, 3),
            "plain": {"nets_s": round(total_nets / p_span, 1),
                      "p50_ms": round(float(np.percentile(p_lat, 50)) * 1e3, 2),
                      "p99_ms": round(float(np.percentile(p_lat, 99)) * 1e3, 2)},
            "serve": {"nets_s": round(total_nets / s_span, 1),
                      "p50_ms": round(float(np.percentile(s_lat, 50)) * 1e3, 2),
                      "p99_ms": round(float(np.percentile(s_lat, 99)) * 1e3, 2),
                      "mean_peers": round(float(np.mean(peers)), 2)},
        })
    st = srv.stats()
    srv.close()
    return {
        "boards": [{"side": s.grid.w, "layers": s.grid.layers} for s in boards],
        "requests": len(reqs),
        "nets": total_nets,
        "plain_mean_service_ms": round(float(np.mean(service)) * 1e3, 3),
        "rows": rows,
        "paths_equal_plain": bool(equal),
        "server": {k: st[k] for k in ("rounds", "tile_visits", "kernel_ms", "label_bytes_peak", "inflight_peak",
                                      "admitted_ahead")},
    }
