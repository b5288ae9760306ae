"""Benchmark: nets/s (complete fields + canonical paths, bit-exact) on the
512x512x4 board with 256 two-pin nets, plus p50/p99 per-batch solve latency.

One step = one batch of nets solved on a board resident in HBM: complete
u64 (dist, hops) label fields for every net (full board, or the net's window
for --config 3), canonical path reconstruction, and the host re-validation of
every path.  Inputs (board planes) are uploaded once before the timed region.
--config 4 solves multi-source fields (one per multi-pin net, no paths).

Output: one JSON line (see the keys below).  ``roofline`` is per launch of
the relaxation kernel (relax_wave or relax_round in the plain form): algorithmic bytes =
B per cell visit (8 B label read + 8 B label write + the per-cell planes of
the workload: 1 B mask, + 2 B u16 history for configs 2 and 3, + 4 B u32
history for config 4) times the cell visits of one launch (device counter of
tile loads x tile cells, averaged over the step's launches), divided by the
kernel's average launch duration from HIP events recorded around every launch
on the engine's stream.  ``traffic`` is the measured HBM bytes per launch
from a committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated
against a copy of known size).  ``cpu_baseline`` times the CPU oracle (a C
restatement of mr-cpu, not the Rust reference) on bounded samples of the same
nets, on one thread and on every CPU of the process's affinity mask, in two
modes (targeted A* with mr-cpu's Manhattan heuristic, and complete fields);
``seam`` times the provider entry (mrf_route_isolated_batch) with its
per-call uploads on a fresh context.  --config 0 times the reference's own
GPU benchmark workload (bench_board) through the Router entry.

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
FORMS = {"plain": 0, "graph": 2, "queue": 3}
CPU_SECONDS = 10.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["fields", "serve", "negotiate", "srj"], default="fields",
                    help="fields: one batch per step of --config (default, the headline); "
                         "serve: Poisson request mix through the serving loop vs one request at a time")
    ap.add_argument("--config", type=int, default=2,
                    help="synthetic config: 0 = the reference's bench_board through the Router entry, "
                         "1 = 128x128x2/64 nets, 2 = 512x512x4/256 nets (headline), "
                         "3 = 2048x2048x2/1024 windowed nets, 4 = 1024x1024x6/512 multi-pin fields")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--mode", choices=["graph", "plain", "queue"], default="plain",
                    help="solve form: plain rounds, hipGraph rounds, or one persistent work-queue launch")
    ap.add_argument("--tiles", default="16,16,64,32",
                    help="tw,th,threads,max_local (16,16,64: one wavefront per tile where the batch qualifies)")
    ap.add_argument("--delta", type=int, default=0, help="delta-stepping window of the plain form (0 = off)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick tiles / round cap / form with the library's autotuner (always on for config 3)")
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--serve-sides", default="128,256,512,1024",
                    help="board sides of the serving mix (boards cycle through them)")
    ap.add_argument("--admit-small", type=int, default=0,
                    help="serving: tiles up to which a request may go ahead (0 = library default, 1 = arrival order)")
    ap.add_argument("--time-every", type=int, default=4,
                    help="plain form: HIP events around every round of every N-th timed step (1 = every step); "
                         "the roofline's launch duration comes from those steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-seam", action="store_true", help="skip the provider-seam measurement")
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


def host_info():
    """The host's CPUs: count, the affinity mask this process may use, model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = sorted(os.sched_getaffinity(0))
    return {"nproc": os.cpu_count(), "affinity_cpus": len(aff),
            "affinity": f"{aff[0]}-{aff[-1]}" if aff and aff[-1] - aff[0] + 1 == len(aff) else ",".join(map(str, aff)),
            "model": model}


def cpu_threads():
    """Every CPU this process may run on."""
    return max(1, len(os.sched_getaffinity(0)))


def timed_chunks(fn, items, chunk, budget_s):
    """Run fn over consecutive chunks of items (wrapping) until budget_s has
    passed (at least one chunk); returns (items done, seconds)."""
    done, k = 0, 0
    t0 = time.perf_counter()
    while True:
        idx = [(k + j) % len(items) for j in range(chunk)]
        fn(idx)
        done += len(idx)
        k += chunk
        if time.perf_counter() - t0 >= budget_s:
            break
    return done, time.perf_counter() - t0


def cpu_baseline_paths(s, got, board):
    """The oracle on the step's nets, two modes, one thread and every core:
    (ii) targeted: the alone search (astar_buf with the Manhattan heuristic
    mr-cpu enables when no cell is free, negotiated.rs:3044-3050; window then
    full-board retry) + canonical path, the work NegotiatedRouter runs; (i)
    complete fields: the lexicographic label field of every cell from the
    source (the bfs_distance_field analogue of sweep.rs:143-148).  Every
    sampled net's path is compared with the GPU's, and the first complete
    fields with the GPU's fields."""
    import numpy as np

    mrcpu, spec = oracle_spec(s)
    g = s.grid
    nets = s.nets()
    wins = s.windows or [(0, 0, g.w - 1, g.h - 1)] * len(nets)
    via = [v for v in s.edges.via if v is not None]
    has_zero = bool(np.any(np.asarray(g.cost) == 0))  # grid.cost.contains(&0), negotiated.rs:1501
    tspec = mrcpu.Spec(spec.grid, spec.edges, additive=s.additive, halo_owner=s.halo_owner,
                       halo_penalty=s.halo_penalty, use_heuristic=not has_zero, min_via_cost=min(via) if via else 0)
    allc = cpu_threads()
    modes = {}
    checked, exact = 0, True

    def targeted(threads):
        nonlocal checked, exact

        def run(idx):
            nonlocal checked, exact
            on = [mrcpu.Net(nets[i].src, nets[i].dst) for i in idx]
            paths, costs = mrcpu.alone_batch(tspec, on, [wins[i] for i in idx], retry_full=True, threads=threads)
            for i, p, c in zip(idx, paths, costs):
                gp = got.path(i)
                exact = exact and ((p is None and gp is None) or (gp is not None and (p, c) == gp))
                checked += 1
        return timed_chunks(run, list(range(len(nets))), max(threads, 8), CPU_SECONDS / 4)

    def fields(threads):
        def run(idx):
            on = [mrcpu.Net(nets[i].src, nets[i].dst) for i in idx]
            mrcpu.fields_batch(spec, on, [[n.src] for n in on], [wins[i] for i in idx], threads=threads,
                               want_hops=True)
        return timed_chunks(run, list(range(len(nets))), threads, CPU_SECONDS / 4)

    for name, fn, unit in (("targeted_astar", targeted, "nets/s"), ("complete_fields", fields, "fields/s")):
        for t in (1, allc):
            done, dt = fn(t)
            modes[f"{name}_{t}t"] = {"value": round(done / dt, 3), "unit": unit, "threads": t,
                                     "items": done, "seconds": round(dt, 2)}
    # complete fields of the first nets against the GPU's
    k = min(4, len(nets))
    board.solve_fields([[nets[i].src] for i in range(k)], dsts=[nets[i].dst for i in range(k)],
                       windows=[wins[i] for i in range(k)])
    on = [mrcpu.Net(nets[i].src, nets[i].dst) for i in range(k)]
    rd, rh = mrcpu.fields_batch(spec, on, [[n.src] for n in on], [wins[i] for i in range(k)], threads=allc)
    fexact = True
    for i in range(k):
        gd, gh = board.read_field(i)
        fexact = fexact and np.array_equal(gd, rd[i]) and np.array_equal(gh, rh[i])
    head = modes[f"targeted_astar_{allc}t"]
    return {
        "value": head["value"],
        "unit": "nets/s",
        "cores": allc,
        "kind": "port",
        "sample": (f"{head['items']} nets of the same workload (step's nets in order, wrapping): oracle alone "
                   f"search, A* with the Manhattan heuristic {'on' if not has_zero else 'off (zero-cost cells)'}, "
                   f"window then full-board retry, canonical path, {allc} threads, {head['seconds']} s; "
                   "modes below also at 1 thread and as complete fields"),
        "host": host_info(),
        "modes": modes,
        "sample_paths_checked": checked,
        "sample_paths_bit_exact_vs_gpu": bool(exact),
        "complete_fields_bit_exact_vs_gpu": bool(fexact),
        "complete_fields_checked": k,
    }


def cpu_baseline_fields(s, board):
    """The oracle's complete multi-source fields for the step's nets, on one
    thread and on every core; the first fields' dist and hops planes are
    compared with the GPU's."""
    import numpy as np

    mrcpu, spec = oracle_spec(s)
    allc = cpu_threads()
    pins = s.pins
    modes = {}
    for t in (1, allc):
        def run(idx):
            nets = [mrcpu.Net(0, 0xFFFFFFFF, own_id=i) for i in idx]
            mrcpu.fields_batch(spec, nets, [pins[i] for i in idx], threads=t)
        done, dt = timed_chunks(run, list(range(len(pins))), t, CPU_SECONDS / 2)
        modes[f"complete_fields_{t}t"] = {"value": round(done / dt, 3), "unit": "fields/s", "threads": t,
                                          "items": done, "seconds": round(dt, 2)}
    k = min(2, len(pins))
    nets = [mrcpu.Net(0, 0xFFFFFFFF, own_id=i) for i in range(k)]
    dist, hops = mrcpu.fields_batch(spec, nets, [pins[i] for i in range(k)], threads=allc)
    exact = True
    for i in range(k):
        gd, gh = board.read_field(i)
        exact = exact and np.array_equal(gd, dist[i]) and np.array_equal(gh, hops[i])
    head = modes[f"complete_fields_{allc}t"]
    return {
        "value": head["value"],
        "unit": "fields/s",
        "cores": allc,
        "kind": "port",
        "sample": f"{head['items']} multi-source fields of the same board, {allc} threads, {head['seconds']} s",
        "host": host_info(),
        "modes": modes,
        "fields_bit_exact_vs_gpu": bool(exact),
        "fields_checked": k,
    }


def seam(s, steps=7, warmup=3):
    """The provider seam as NegotiatedRouter calls it: mrf_route_isolated_batch
    on a fresh context (its default form), the board's grid, masks and edge
    prices passed per call (alone-path prices: the request carries no history
    plane), full-board windows.  Per call the library uploads the caller's
    planes, builds the directed-edge and via masks on the device, solves, and
    releases the board again; the paths come back as the ABI's flat arrays."""
    import numpy as np

    import mrfield

    eng = mrfield.Engine(0)
    g = s.grid
    nets = s.nets()
    wins = np.asarray(s.windows or [(0, 0, g.w - 1, g.h - 1)] * len(nets), dtype=np.uint32).reshape(-1, 4)
    for _ in range(warmup):  # a fresh context's first calls size its buffers
        eng.route_isolated_batch_arrays(g, nets, wins, s.edges)
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        eng.route_isolated_batch_arrays(g, nets, wins, s.edges)
        ts.append(time.perf_counter() - t0)
    st = eng.stats()
    eng.close()
    ms = sorted(ts)[len(ts) // 2] * 1e3
    return {"entry": "mrf_route_isolated_batch", "value": round(len(nets) / (ms * 1e-3), 3), "unit": "nets/s",
            "ms_per_call_p50": round(ms, 3), "calls": steps, "warmup_calls": warmup,
            "relax_kernel_ms": round(st["relax_kernel_ms"], 3),
            "note": "fresh context's default (plain) form; the caller's planes and prices uploaded and the "
                    "masks built per call (host-to-device copies included); alone-path pricing (no history "
                    "plane); paths returned as flat arrays"}


def config0_main(a):
    """The reference's own GPU benchmark workload (mr-metal bench_board,
    crates/mr-metal/src/lib.rs:2761-2906): MetalRouter::route on a 128x128
    board with vertical walls and 64 nets, warm median of 7 calls.  Here the
    Router entry, mrf_router_route (board uploaded per call like the
    reference's), bit-exact against the oracle's Lee search."""
    import numpy as np

    import mrfield
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mrcpu

    s = mrfield.synth_config(0)
    g = s.grid
    nets = s.nets()
    eng = mrfield.Engine(0)
    for _ in range(max(a.warmup, 1)):
        got = eng.router_route(g, nets)
    ts = []
    t_start = time.perf_counter()
    for _ in range(a.steps):
        t0 = time.perf_counter()
        got = eng.router_route(g, nets)
        ts.append((time.perf_counter() - t0) * 1e3)
    wall = time.perf_counter() - t_start
    st = eng.stats()
    eng.close()
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost)
    lee = mrcpu.Spec(og, mode=mrcpu.MODE_LEE)
    buf = mrcpu.SearchBuf(g.n)
    exact = True
    c0 = time.perf_counter()
    for n, r in zip(nets, got):
        ref = mrcpu.astar(lee, n.src, n.dst, buf)
        exact = exact and ((ref is None and r is None) or (r is not None and ref is not None and r[0] == ref[0]))
    cpu_s = time.perf_counter() - c0
    p50 = float(np.percentile(ts, 50))
    out = {
        "metric": "ms per MetalRouter::route call, 128x128 bench_board, 64 nets (warm median)",
        "value": round(p50, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "reference_m4_warm_median_ms": [3.42, 3.75],
        "speedup_vs_reference_m4": round(3.42 / p50, 2),
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "reference bench_board: 128x128x1, side/8 single-column walls leaving row 0 "
                               "open, unit cost, 64 nets from xorshift 0x9E3779B97F4A7C15 (config 0)"},
        "nets_per_s": round(len(nets) / (p50 * 1e-3), 1),
        "p99_ms": round(float(np.percentile(ts, 99)), 4),
        "routed": int(sum(r is not None for r in got)),
        "relax_kernel_ms": round(st["relax_kernel_ms"], 4),
        "relax_launches": st["relax_launches"],
        "paths_bit_exact_vs_oracle_lee": bool(exact),
        "cpu_baseline": {"value": round(len(nets) / cpu_s, 1), "unit": "nets/s", "cores": 1, "kind": "port",
                         "sample": f"all {len(nets)} nets: oracle Lee search (targeted, early stop), 1 thread, "
                                   f"{cpu_s:.2f} s", "host": host_info()},
    }
    print(json.dumps(out))


def srj_main(a, tiles):
    """Real boards end to end: every committed SimpleRouteJson fixture
    (tests/golden/srj: the reference's CLI payloads and corpus boards) is
    rasterised by the library's producer with the CLI's rule selection
    (mrf_srj_rasterize) and all its nets routed through the provider seam
    (mrf_route_isolated_batch, complete fields on the full board) with
    geometric planar prices round(16 x line gap) and 160 per via.  A step is
    one pass over every board; paths are checked against the oracle after the
    timed region."""
    import glob

    import numpy as np

    import mrfield
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mrcpu

    docs = [(os.path.basename(f), open(f, "rb").read())
            for f in sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "srj", "*.json")))]
    eng = mrfield.Engine(0)
    eng.set_tiles(*tiles)
    eng.set_mode(FORMS[a.mode])

    def prices(lines):
        gaps = np.diff(np.asarray(lines, dtype=np.float64))
        return [int(v) for v in np.floor(gaps * 16.0 + 0.5)]

    def one(data):
        t0 = time.perf_counter()
        p = mrfield.srj_rasterize(data, mode=mrfield.SRJ_ROUTE)
        t1 = time.perf_counter()
        ed = mrfield.Edges(prices(p.x_lines), prices(p.y_lines), [160] * (p.layers - 1))
        got = eng.route_isolated_batch(p.grid(), [n.net() for n in p.nets], [(0, 0, p.w - 1, p.h - 1)] * len(p.nets),
                                       ed)
        t2 = time.perf_counter()
        return p, ed, got, t1 - t0, t2 - t1

    for _ in range(max(a.warmup, 1)):
        for _, d in docs:
            one(d)
    r_ms, s_ms, last = [], [], {}
    t_start = time.perf_counter()
    for _ in range(a.steps):
        for name, d in docs:
            p, ed, got, tr, ts = one(d)
            r_ms.append(tr * 1e3)
            s_ms.append(ts * 1e3)
            last[name] = (p, ed, got)
    wall = time.perf_counter() - t_start
    eng.close()
    n_nets = sum(len(v[0].nets) for v in last.values())
    exact, c_s = True, 0.0
    for name, (p, ed, got) in last.items():
        og = mrcpu.Grid(p.w, p.h, p.layers, p.cost, p.via_forbidden, p.board)
        c0 = time.perf_counter()
        ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, mrcpu.Edges(ed.x, ed.y, ed.via)),
                                         [mrcpu.Net(n.src, n.dst, n.pads, n.via_pads) for n in p.nets], None)
        c_s += time.perf_counter() - c0
        for i, r in enumerate(got):
            exact = exact and ((r is None) == (ref_p[i] is None)) and (r is None or r == (ref_p[i], ref_c[i]))
    out = {
        "metric": "boards/s (SimpleRouteJson rasterised + every net routed, bit-exact)",
        "value": round(len(docs) * a.steps / wall, 3),
        "unit": "boards/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "committed SimpleRouteJson fixtures (the reference's CLI payloads and corpus boards)",
        "config": {"workload": f"{len(docs)} boards, {n_nets} two-pin nets, complete fields on the full board",
                   "form": a.mode, "tiles": ",".join(str(t) for t in tiles)},
        "nets_per_s": round(n_nets * a.steps / wall, 1),
        "rasterise_ms_per_board_p50": round(float(np.percentile(r_ms, 50)), 3),
        "rasterise_ms_per_board_max": round(float(np.max(r_ms)), 3),
        "route_ms_per_board_p50": round(float(np.percentile(s_ms, 50)), 3),
        "route_ms_per_board_max": round(float(np.max(s_ms)), 3),
        "boards": {k: [v[0].w, v[0].h, v[0].layers, len(v[0].nets), sum(r is not None for r in v[2])]
                   for k, v in last.items()},
        "paths_bit_exact_vs_oracle": bool(exact),
        "cpu_baseline": {"value": round(n_nets / c_s, 1) if c_s > 0 else None, "unit": "nets/s", "cores": 1,
                         "kind": "port", "sample": f"all {n_nets} nets of one pass, oracle alone search "
                                                   f"(complete fields, full board), 1 thread, {c_s:.2f} s",
                         "host": host_info()},
    }
    print(json.dumps(out))


def serve_main(a, tiles):
    import serve_bench

    t0 = time.perf_counter()
    sides = tuple(int(v) for v in a.serve_sides.split(","))
    r = serve_bench.run(seed=a.seed, n_req=a.serve_requests, tiles=tiles, sides=sides,
                        admit_small_tiles=a.admit_small)
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
        "config": {"workload": f"{r['requests']} requests over {len(r['boards'])} resident boards (sides "
                               f"{a.serve_sides}), {r['nets']} nets, Poisson arrivals at 0.5x-4x the "
                               f"one-request-at-a-time capacity (seed {a.seed})", "boards": r["boards"],
                   "tiles": a.tiles, "admit_small_tiles": a.admit_small or "default"},
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
    eng.set_mode(FORMS[a.mode])
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
    if a.config == 0 and a.workload == "fields":
        config0_main(a)
        return
    if a.workload == "serve":
        serve_main(a, (tw, th, threads, ml))
        return
    if a.workload == "negotiate":
        negotiate_main(a, (tw, th, threads, ml))
        return
    if a.workload == "srj":
        srj_main(a, (tw, th, threads, ml))
        return
    eng = mrfield.Engine(0)
    eng.set_tiles(tw, th, threads, ml)
    eng.set_mode(FORMS[a.mode])
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
    if (a.autotune or (a.config == 3 and a.mode != "queue")) and not multi:
        tuning = board.autotune(nets, windows=wins, sample=64)
        form = {v: k for k, v in FORMS.items()}[tuning["mode"]]
        tw, th = tuning["tile_w"], tuning["tile_h"]
        tiles = f"{tw},{th},{tuning['threads']},{tuning['max_local']}"

    def step():
        if multi:
            board.solve_fields(s.pins, net_ids=list(range(len(s.pins))))
            return None
        return board.solve_arrays(nets, windows=wins, retry_full=wins is not None)

    lat, relax_ms, path_ms, visits, cells_visited, rounds_l, host_rounds = [], [], [], [], [], [], []
    kern_ms, launches, changed, valid_ms, call_ms, fcells, wave, lbytes = [], [], [], [], [], [], [], []
    timed_l = []
    got = None
    for i in range(a.warmup + a.steps):
        if i == a.warmup:
            # per-round events (about 3 us of host turnaround a round) in the
            # first of every --time-every steps of the timed region
            eng.set_round_timing(max(1, a.time_every))
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
            timed_l.append(st["timed_launches"])
            wave.append(st["wave_launches"])
            lbytes.append(st["label_bytes"])
            fcells.append(st["field_cells"])
    wall = time.perf_counter() - t_start
    n_nets = len(s.pins) if multi else len(nets)
    value = n_nets * a.steps / wall
    routed = int(np.count_nonzero(got.routed)) if got is not None else None
    bpc = BYTES_PER_CELL_VISIT[a.config]
    # per launch of the relaxation kernel: algorithmic bytes / average duration,
    # over the steps whose every launch carried its HIP event pair (plain
    # form; the other forms time their launch or replays in every step)
    tsel = [k for k in range(len(kern_ms)) if launches[k] and timed_l[k] == launches[k]]
    if form != "plain" or not tsel:
        tsel = list(range(len(kern_ms)))
    kern_ms = [kern_ms[k] for k in tsel]
    n_launch = float(np.mean([launches[k] for k in tsel]))
    cells_per_launch = float(np.mean([cells_visited[k] for k in tsel])) / max(n_launch, 1.0)
    avg_launch_ms = float(np.mean(kern_ms)) / max(n_launch, 1.0)
    achieved = cells_per_launch * bpc / (avg_launch_ms * 1e-3) / 1e9
    workload = f"c{a.config}-{g.w}x{g.h}x{g.layers}-{n_nets}nets-t{tw}x{th}"
    kernel = {"graph": "relax_list", "plain": "relax_round", "queue": "relax_queue"}[form]
    if sum(wave) == sum(launches) > 0:
        # one wavefront per tile visit (the queue form's persistent variant)
        kernel = {"plain": "relax_wave", "queue": "relax_wave_queue"}.get(form, kernel)
    traffic = load_traffic(kernel, workload)
    # compulsory floor of a step: every field cell's dist written once (4 B)
    # and the board planes read once per field (bpc - 16 B per cell)
    floor_bytes = float(np.mean(fcells)) * (4 + bpc - 16)
    kern_s = float(np.mean(kern_ms)) * 1e-3
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
        "kernel_timing": (f"HIP events around every launch in {len(tsel)} of {a.steps} timed steps"
                          + (f" (every {max(1, a.time_every)}th)" if form == "plain" else "")),
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
            "traffic": traffic,
            "kernel": kernel,
            "bytes_per_cell_visit": bpc,
            # the kernel's own label width: 8 B labels move bpc bytes per cell
            # visit; packed 4 B labels move 8 fewer (the survey's unit above
            # counts u64 labels either way)
            "label_bytes": int(max(lbytes)) if lbytes else 8,
            # the wave form's lean visits load 12 B of step price pairs per
            # cell from the board's step planes instead of its 1-7 B of planes
            "bytes_moved_per_cell_visit": (2 * (int(max(lbytes)) if lbytes else 8)
                                           + (12 if kernel.startswith("relax_wave") and a.config != 4
                                              else bpc - 16)),
            "cell_visits_per_launch": int(cells_per_launch),
            "achieved_from_traffic": (round(traffic / (avg_launch_ms * 1e-3) / 1e9, 2) if traffic else None),
            "traffic_over_algorithmic": (round(traffic / (cells_per_launch * bpc), 3) if traffic else None),
            "compulsory_bytes_per_step": int(floor_bytes),
            "compulsory_GBs_over_kernel_time": round(floor_bytes / kern_s / 1e9, 2) if kern_s > 0 else None,
        },
        "cpu_baseline": None,
    }
    if tuning is not None:
        out["autotune"] = tuning
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_fields(s, board) if multi else cpu_baseline_paths(s, got, board)
    board.close()
    eng.close()
    if not multi and not a.no_seam:
        out["seam"] = seam(s)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
