"""Benchmark: nets/s (complete fields + canonical paths, bit-exact) on the
512x512x4 board with 256 two-pin nets, plus p50/p99 per-batch solve latency.

One step = one batch of nets solved on a board resident in HBM: complete
u64 (dist, hops) label fields for every net (full board, or the net's window
for --config 3), canonical path reconstruction, and the host re-validation of
every path.  Inputs (board planes) are uploaded once before the timed region.
--config 4 solves multi-source fields (one per multi-pin net, no paths).

Output: one JSON line (see the keys below).  ``roofline`` is per launch of
the relaxation kernel (relax_round in the plain form): algorithmic bytes =
B per cell visit (8 B label read + 8 B label write + the per-cell planes of
the workload: 1 B mask, + 2 B u16 history for configs 2 and 3, + 4 B u32
history for config 4) times the cell visits of one launch (device counter of
tile loads x tile cells, averaged over the step's launches), divided by the
kernel's average launch duration from HIP events recorded around every launch
on the engine's stream.  ``traffic`` is the measured HBM bytes per launch
from a committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated
against a copy of known size).  ``cpu_baseline`` times the CPU oracle (a C
restatement of mr-cpu, not the Rust reference) on a bounded sample of the
same nets on the host cores.

--workload serve runs the serving comparison instead (serve_bench.py).
--workload negotiate runs the negotiated-congestion loop (negotiate.py) on
the same board and nets: one step = one negotiation iteration, its dirty nets
routed in one device batch with per-net self-subtracted prices.
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
BYTES_PER_CELL_VISIT = {0: 17, 1: 17, 2: 19, 3: 19, 4: 21}
CPU_SECONDS = 10.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["fields", "serve", "negotiate"], default="fields",
                    help="fields: one batch per step of --config (default, the headline); "
                         "serve: Poisson request mix through the serving loop vs one request at a time")
    ap.add_argument("--config", type=int, default=2,
                    help="synthetic config: 1 = 128x128x2/64 nets, 2 = 512x512x4/256 nets (headline), "
                         "3 = 2048x2048x2/1024 windowed nets, 4 = 1024x1024x6/512 multi-pin fields")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--mode", choices=["graph", "plain"], default="plain")
    ap.add_argument("--tiles", default="16,16,256,32", help="tw,th,threads,max_local")
    ap.add_argument("--delta", type=int, default=0, help="delta-stepping window of the plain form (0 = off)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick tiles / round cap / form with the library's autotuner (always on for config 3)")
    ap.add_argument("--serve-requests", type=int, default=96)
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


def oracle_spec(s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mrcpu

    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost, board=g.board)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    return mrcpu, mrcpu.Spec(og, oe, additive=s.additive, halo_owner=s.halo_owner, halo_penalty=s.halo_penalty)


def cpu_threads():
    return max(1, min(os.cpu_count() or 1, 16))


def cpu_baseline_paths(s, got):
    """The oracle's alone search (window, then full-board retry) over the
    step's nets in chunks on the host cores until about CPU_SECONDS; the first
    chunk's paths are compared with the GPU's."""
    mrcpu, spec = oracle_spec(s)
    g = s.grid
    nets = s.nets()
    wins = s.windows or [(0, 0, g.w - 1, g.h - 1)] * len(nets)
    threads = cpu_threads()
    chunk = max(threads, 16)
    done, exact, k = 0, True, 0
    t0 = time.perf_counter()
    while True:
        idx = [(k + j) % len(nets) for j in range(chunk)]
        on = [mrcpu.Net(nets[i].src, nets[i].dst) for i in idx]
        paths, _ = mrcpu.alone_batch(spec, on, [wins[i] for i in idx], retry_full=True, threads=threads)
        if k == 0:
            for i, p in zip(idx, paths):
                gp = got.path(i)
                exact = exact and ((p is None and gp is None) or (gp is not None and p == gp[0]))
        done += len(idx)
        k += chunk
        if time.perf_counter() - t0 >= CPU_SECONDS:
            break
    dt = time.perf_counter() - t0
    return {
        "value": round(done / dt, 3),
        "unit": "nets/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{done} nets of the same workload in chunks of {chunk}: alone search (complete "
                  f"lexicographic labels, window then full-board retry) + canonical path, {threads} threads, "
                  f"{dt:.1f} s",
        "paths_bit_exact_vs_gpu": bool(exact),
    }


def cpu_baseline_fields(s, board):
    """The oracle's complete multi-source fields for nets in chunks until about
    CPU_SECONDS; the first chunk's dist and hops planes are compared with the
    GPU's fields."""
    import numpy as np

    mrcpu, spec = oracle_spec(s)
    threads = cpu_threads()
    pins = s.pins
    done, exact, k = 0, True, 0
    t0 = time.perf_counter()
    while True:
        idx = [(k + j) % len(pins) for j in range(threads)]
        nets = [mrcpu.Net(0, 0xFFFFFFFF, own_id=i) for i in idx]  # no target: sources only
        dist, hops = mrcpu.fields_batch(spec, nets, [pins[i] for i in idx], threads=threads)
        if k == 0:
            for j, i in enumerate(idx[:2]):
                gd, gh = board.read_field(i)
                exact = exact and np.array_equal(gd, dist[j]) and np.array_equal(gh, hops[j])
        done += len(idx)
        k += len(idx)
        if time.perf_counter() - t0 >= CPU_SECONDS:
            break
    dt = time.perf_counter() - t0
    return {
        "value": round(done / dt, 3),
        "unit": "fields/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{done} multi-source fields of the same board, {threads} threads, {dt:.1f} s",
        "fields_bit_exact_vs_gpu": bool(exact),
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


def negotiate_main(a, tiles):
    """Negotiation iterations on the config's board: searches/s over the
    timed iterations, per-iteration latency, and the oracle's route_negotiated
    on the first iteration's snapshot as the CPU baseline."""
    import numpy as np

    import mrfield
    import negotiate as ng

    eng = mrfield.Engine(0)
    eng.set_tiles(*tiles)
    eng.set_mode(mrfield.MODE_GRAPH if a.mode == "graph" else mrfield.MODE_PLAIN)
    s = mrfield.synth_config(a.config, seed=a.seed)
    g = s.grid
    board = mrfield.Board(eng, g, s.edges, None)
    nets = s.nets()
    hist0 = s.additive  # the configuration's congestion-history plane
    if a.warmup:
        ng.negotiate_board(board, nets, max_iters=a.warmup, history0=hist0)
    times, dirty = [], []
    last = [time.perf_counter()]

    def on_iter(it, n):
        t = time.perf_counter()
        times.append(t - last[0])
        dirty.append(n)
        last[0] = t

    t0 = time.perf_counter()
    last[0] = t0
    res = ng.negotiate_board(board, nets, max_iters=a.steps, history0=hist0, on_iter=on_iter)
    wall = time.perf_counter() - t0
    searches = int(sum(dirty))
    out = {
        "metric": "negotiation searches/sec (route_negotiated batches, bit-exact) at "
                  f"{g.w}x{g.h}x{g.layers}; per-iteration latency",
        "value": round(searches / wall, 3),
        "unit": "searches/s",
        "n_gpus": 1,
        "steps": res.iterations,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / max(res.iterations, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"{g.w}x{g.h}x{g.layers} board, {len(nets)} two-pin nets, negotiation "
                               f"iterations (synthetic config {a.config}, seed {a.seed}; history seeded from "
                               "the configuration's history plane)", "tiles": a.tiles},
        "dirty_per_iteration": dirty,
        "ms_per_iteration": [round(t * 1e3, 2) for t in times],
        "converged": res.converged,
        "unrouted": int(sum(1 for p in res.paths if not p)),
    }
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import mrcpu

        og = mrcpu.Grid(g.w, g.h, g.layers, g.cost, board=g.board)
        oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
        n = g.n
        tmpl = mrcpu.Spec(og, oe, negotiation=mrcpu.Negotiation(
            history=np.asarray(hist0, dtype=np.uint32), present=np.zeros(n, dtype=np.uint32), present_halo=None,
            pfac=1, fused=True))
        wins = [ng.net_window(g.w, g.h, x.src, x.dst) for x in nets]
        threads = cpu_threads()
        chunk = max(threads, 16)
        done, k = 0, 0
        c0 = time.perf_counter()
        while time.perf_counter() - c0 < CPU_SECONDS:
            idx = [(k + j) % len(nets) for j in range(chunk)]
            on = [mrcpu.Net(nets[i].src, nets[i].dst, own_halo=[]) for i in idx]
            mrcpu.alone_batch(tmpl, on, [wins[i] for i in idx], retry_full=True, threads=threads)
            done += len(idx)
            k += chunk
        dt = time.perf_counter() - c0
        out["cpu_baseline"] = {
            "value": round(done / dt, 3), "unit": "searches/s", "cores": threads, "kind": "port",
            "sample": f"{done} route_negotiated searches of the first iteration (window then full board), "
                      f"{threads} threads, {dt:.1f} s"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np

    import mrfield

    tw, th, threads, ml = (int(v) for v in a.tiles.split(","))
    if a.workload == "serve":
        serve_main(a, (tw, th, threads, ml))
        return
    if a.workload == "negotiate":
        negotiate_main(a, (tw, th, threads, ml))
        return
    eng = mrfield.Engine(0)
    eng.set_tiles(tw, th, threads, ml)
    eng.set_mode(mrfield.MODE_GRAPH if a.mode == "graph" else mrfield.MODE_PLAIN)
    if a.delta:
        eng.set_delta(a.delta)
    s = mrfield.synth_config(a.config, seed=a.seed)
    g = s.grid
    multi = s.pins is not None
    board = mrfield.Board(eng, g, s.edges, s.pricing(list(range(len(s.pins))) if multi else None))
    nets = s.nets()
    wins = s.windows
    tuning = None
    form, tiles = a.mode, a.tiles
    if (a.autotune or a.config == 3) and not multi:
        tuning = board.autotune(nets, windows=wins, sample=64)
        form = "graph" if tuning["mode"] == mrfield.MODE_GRAPH else "plain"
        tw, th = tuning["tile_w"], tuning["tile_h"]
        tiles = f"{tw},{th},{tuning['threads']},{tuning['max_local']}"

    def step():
        if multi:
            board.solve_fields(s.pins, net_ids=list(range(len(s.pins))))
            return None
        return board.solve_arrays(nets, windows=wins, retry_full=wins is not None)

    lat, relax_ms, path_ms, visits, cells_visited, rounds_l, host_rounds = [], [], [], [], [], [], []
    kern_ms, launches, changed, valid_ms, call_ms = [], [], [], [], []
    got = None
    for i in range(a.warmup + a.steps):
        if i == a.warmup:
            t_start = time.perf_counter()
        t0 = time.perf_counter()
        got = step()
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
            changed.append(st["changed_visits"])
            valid_ms.append(st["validate_ms"])
            call_ms.append(st["total_ms"])
            launches.append(st["relax_launches"])
    wall = time.perf_counter() - t_start
    n_nets = len(s.pins) if multi else len(nets)
    value = n_nets * a.steps / wall
    routed = int(np.count_nonzero(got.routed)) if got is not None else None
    bpc = BYTES_PER_CELL_VISIT[a.config]
    # per launch of the relaxation kernel: algorithmic bytes / average duration
    n_launch = float(np.mean(launches))
    cells_per_launch = float(np.mean(cells_visited)) / max(n_launch, 1.0)
    avg_launch_ms = float(np.mean(kern_ms)) / max(n_launch, 1.0)
    achieved = cells_per_launch * bpc / (avg_launch_ms * 1e-3) / 1e9
    workload = f"c{a.config}-{g.w}x{g.h}x{g.layers}-{n_nets}nets-t{tw}x{th}"
    kernel = "relax_list" if form == "graph" else "relax_round"
    if a.config == 2:
        metric, desc = METRIC, "two-pin nets, complete fields + canonical paths per step"
    elif multi:
        metric = f"fields/sec (multi-source, complete, bit-exact) at {g.w}x{g.h}x{g.layers}"
        desc = "multi-pin nets, one complete multi-source field per net per step (no paths)"
    else:
        metric = f"nets/sec (fields+paths, bit-exact) at {g.w}x{g.h}x{g.layers}; p50/p99 solve latency"
        desc = ("two-pin nets in their windows (full-board retry), fields + canonical paths per step"
                if wins is not None else "two-pin nets, complete fields + canonical paths per step")
    out = {
        "metric": metric,
        "value": round(value, 3),
        "unit": "fields/s" if multi else "nets/s",
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
            "workload": f"{g.w}x{g.h}x{g.layers} board, {n_nets} {desc} (synthetic config {a.config}, "
                        f"seed {a.seed})",
            "form": form,
            "tiles": tiles,
            "delta": (tuning["delta"] if tuning is not None else (a.delta or None)),
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
        "validate_ms": round(float(np.mean(valid_ms)), 4),
        "solve_call_ms": round(float(np.mean(call_ms)), 4),
        "tile_visits": int(np.mean(visits)),
        "changed_visits": int(np.mean(changed)),
        "local_rounds_per_visit": round(float(np.mean(rounds_l)) / max(float(np.mean(visits)), 1), 3),
        "host_rounds": int(np.mean(host_rounds)),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(kernel, workload),
            "kernel": kernel,
            "bytes_per_cell_visit": bpc,
            "cell_visits_per_launch": int(cells_per_launch),
        },
        "cpu_baseline": None,
    }
    if tuning is not None:
        out["autotune"] = tuning
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_fields(s, board) if multi else cpu_baseline_paths(s, got)
    board.close()
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
