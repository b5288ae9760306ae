"""Benchmark: nets/s (complete fields + canonical paths, bit-exact) on the
512x512x4 board with 256 two-pin nets, plus p50/p99 per-batch solve latency.

One step = one batch of 256 nets solved on a board resident in HBM: complete
u64 (dist, hops) label fields over the full board for every net, canonical
path reconstruction, and the host re-validation of every path.  Inputs (board
planes) are uploaded once before the timed region.

Output: one JSON line (see the keys below).  ``roofline`` is per launch of
the relaxation kernel (relax_round in the plain form): algorithmic bytes =
19 B per cell visit (8 B label read + 8 B label write + 2 B history + 1 B
mask) times the cell visits of one launch (device counter of tile loads x
tile cells, averaged over the step's launches), divided by the kernel's
average launch duration from HIP events recorded around every launch on the
engine's stream.  ``traffic`` is the measured HBM bytes per launch from a
committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated
against a copy of known size).  ``cpu_baseline`` times the CPU oracle (a C
restatement of mr-cpu, not the Rust reference) on a bounded sample of the
same nets.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "metal-route_amd"))

METRIC = "nets/sec (fields+paths, bit-exact) at 512²×4L; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_VISIT = 19


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["fields", "serve"], default="fields",
                    help="fields: one batch per step of --config (default, the headline); "
                         "serve: Poisson request mix through the serving loop vs one request at a time")
    ap.add_argument("--config", type=int, default=2, help="synthetic config (2 = 512x512x4, 256 nets)")
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--mode", choices=["graph", "plain"], default="plain")
    ap.add_argument("--tiles", default="16,16,256,32", help="tw,th,threads,max_local")
    ap.add_argument("--cpu-sample", type=int, default=16, help="nets in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (rocprof counter calibration)")
    return ap.parse_args()


def calibrate_copy(mib: int) -> None:
    """A plain HBM copy of known size: under rocprofv3 --pmc its FETCH_SIZE /
    WRITE_SIZE give the factor between the counters and real bytes."""
    import torch

    n = mib * 1024 * 1024 // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "checksum": int(dst[-1].item())}))


def load_traffic(kernel: str, workload: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if any."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get(f"{kernel}@{workload}")
    except (OSError, ValueError):
        return None


def cpu_baseline(s, sample, got_paths):
    """Time the CPU oracle on `sample` nets of the same workload (fields + paths)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mrcpu

    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost, board=g.board)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    spec = mrcpu.Spec(og, oe, additive=s.additive)
    nets = [mrcpu.Net(int(a), int(b)) for a, b in zip(s.src[:sample], s.dst[:sample])]
    threads = min(os.cpu_count() or 1, 16, sample)
    t0 = time.perf_counter()
    dist, hops = mrcpu.fields_batch(spec, nets, [[n.src] for n in nets], threads=threads)
    paths = [mrcpu.reconstruct(spec, dist[i], hops[i], n.dst) for i, n in enumerate(nets)]
    dt = time.perf_counter() - t0
    exact = True
    for i, p in enumerate(paths):
        gp = got_paths.path(i)
        exact = exact and ((p is None and gp is None) or (gp is not None and p == gp[0]))
    return {
        "value": round(sample / dt, 3),
        "unit": "nets/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample} nets of the same board: complete lexicographic (dist,hops) Dijkstra fields "
                  f"+ canonical reconstruction, {threads} threads, {dt:.2f} s",
        "paths_bit_exact_vs_gpu": bool(exact),
    }


def serve_main(a, tiles):
    import serve_bench

    t0 = time.perf_counter()
    r = serve_bench.run(seed=a.seed, n_req=a.serve_requests, tiles=tiles)
    top = r["rows"][-1]
    out = {
        "metric": "serving nets/sec, Poisson request mix on resident boards (sides 128-1024, 2-8 layers, "
                  "16-512 nets); p50/p99 latency",
        "value": top["serve"]["nets_s"],
        "unit": "nets/s",
        "n_gpus": 1,
        "steps": r["requests"],
        "warmup": 1,
        "ms_per_step": round(r["plain_mean_service_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"{r['requests']} requests over {len(r['boards'])} resident boards, "
                               f"{r['nets']} nets, Poisson arrivals at 0.5x-4x the one-request-at-a-time "
                               f"capacity (seed {a.seed})", "boards": r["boards"], "tiles": a.tiles},
        "rows": r["rows"],
        "speedup_vs_plain_at_top_load": round(top["serve"]["nets_s"] / top["plain"]["nets_s"], 3),
        "paths_equal_plain": r["paths_equal_plain"],
        "server": r["server"],
        "wall_s": round(time.perf_counter() - t0, 1),
    }
    print(json.dumps(out))


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np

    import mrfield

    tw, th, threads, ml = (int(v) for v in a.tiles.split(","))
    if a.workload == "serve":
        serve_main(a, (tw, th, threads, ml))
        return
    eng = mrfield.Engine(0)
    eng.set_tiles(tw, th, threads, ml)
    eng.set_mode(mrfield.MODE_GRAPH if a.mode == "graph" else mrfield.MODE_PLAIN)
    s = mrfield.synth_config(a.config, seed=a.seed)
    g = s.grid
    board = mrfield.Board(eng, g, s.edges, s.pricing())
    nets = s.nets()
    wins = s.windows

    lat, relax_ms, path_ms, visits, cells_visited, rounds_l, host_rounds = [], [], [], [], [], [], []
    kern_ms, launches = [], []
    got = None
    for i in range(a.warmup + a.steps):
        if i == a.warmup:
            t_start = time.perf_counter()
        t0 = time.perf_counter()
        got = board.solve_arrays(nets, windows=wins, retry_full=wins is not None)
        t1 = time.perf_counter()
        if i >= a.warmup:
            st = eng.stats()
            lat.append((t1 - t0) * 1e3)
            relax_ms.append(st["relax_ms"])
            path_ms.append(st["path_ms"])
            visits.append(st["tile_visits"])
            cells_visited.append(st["cells_visited"])
            rounds_l.append(st["local_rounds"])
            host_rounds.append(st["host_rounds"])
            kern_ms.append(st["relax_kernel_ms"])
            launches.append(st["relax_launches"])
    wall = time.perf_counter() - t_start
    n_nets = len(nets)
    value = n_nets * a.steps / wall
    routed = int(np.count_nonzero(got.routed))
    # per launch of the relaxation kernel: algorithmic bytes / average duration
    n_launch = float(np.mean(launches))
    cells_per_launch = float(np.mean(cells_visited)) / max(n_launch, 1.0)
    avg_launch_ms = float(np.mean(kern_ms)) / max(n_launch, 1.0)
    achieved = cells_per_launch * BYTES_PER_CELL_VISIT / (avg_launch_ms * 1e-3) / 1e9
    workload = f"c{a.config}-{g.w}x{g.h}x{g.layers}-{n_nets}nets-t{tw}x{th}"
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "nets/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"{g.w}x{g.h}x{g.layers} board, {n_nets} two-pin nets, complete fields + canonical "
                        f"paths per step (synthetic config {a.config}, seed {a.seed})",
            "form": a.mode,
            "tiles": a.tiles,
        },
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "fields_per_s": round(value, 3),
        "routed": routed,
        "relax_phase_ms": round(float(np.mean(relax_ms)), 4),
        "relax_kernel_ms": round(float(np.mean(kern_ms)), 4),
        "relax_launches": int(n_launch),
        "relax_avg_launch_ms": round(avg_launch_ms, 5),
        "path_ms": round(float(np.mean(path_ms)), 4),
        "tile_visits": int(np.mean(visits)),
        "local_rounds_per_visit": round(float(np.mean(rounds_l)) / max(float(np.mean(visits)), 1), 3),
        "host_rounds": int(np.mean(host_rounds)),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic("relax_list" if a.mode == "graph" else "relax_round", workload),
            "kernel": "relax_list" if a.mode == "graph" else "relax_round",
            "bytes_per_cell_visit": BYTES_PER_CELL_VISIT,
            "cell_visits_per_launch": int(cells_per_launch),
        },
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, min(a.cpu_sample, n_nets), got)
    board.close()
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
