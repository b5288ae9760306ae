"""Python host binding of libmrfield.so (the MI355X field engine's C ABI).

Mirrors the reference's provider seam for tests and benchmarks:
``Engine.route_isolated_batch`` is IsolatedRouteProvider::route_isolated_batch
(crates/mr-cpu/src/negotiated.rs:554-560) through MetalIsolatedProvider
(crates/mr-cli/src/lib.rs:612-625); ``route_isolated_pass`` is
metal_route_isolated_batch (crates/mr-metal/src/lib.rs:414); ``sweep_fields``
is metal_sweep_fields (:172); ``router_route`` is MetalRouter::route (:1143).

There is no CPU fallback: every call runs the HIP kernels, and loading fails
loudly when the library or a HIP device is missing.  Errors raise
``MrfError`` carrying the RouterError code (crates/mr-core/src/lib.rs:699-708).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrfield.so")

MAX = 0xFFFFFFFF
OK, BACKEND_UNAVAILABLE, INVALID_ENDPOINT, MALFORMED_GRID = 0, 1, 2, 3
MODE_PLAIN, MODE_PERSISTENT, MODE_GRAPH, MODE_QUEUE = 0, 1, 2, 3

U32P = C.POINTER(C.c_uint32)
U8P = C.POINTER(C.c_uint8)
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)


class Dims(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("layers", C.c_uint32)]


class GridS(C.Structure):
    _fields_ = [("cost", U32P), ("n_cost", C.c_size_t), ("via_forbidden", U8P), ("n_via_forbidden", C.c_size_t),
                ("board_constraint", U8P), ("n_board_constraint", C.c_size_t)]


class NetS(C.Structure):
    _fields_ = [("src", C.c_uint32), ("dst", C.c_uint32), ("pads", U32P), ("n_pads", C.c_size_t),
                ("via_pads", U32P), ("n_via_pads", C.c_size_t)]


class WindowS(C.Structure):
    _fields_ = [("x0", C.c_uint32), ("y0", C.c_uint32), ("x1", C.c_uint32), ("y1", C.c_uint32)]


class EdgesS(C.Structure):
    _fields_ = [("x", U32P), ("nx", C.c_size_t), ("y", U32P), ("ny", C.c_size_t), ("via", U32P),
                ("via_allowed", U8P), ("nvia", C.c_size_t)]


class PricingS(C.Structure):
    _fields_ = [("additive", U32P), ("n_additive", C.c_size_t), ("halo_owner", I32P), ("n_halo_owner", C.c_size_t),
                ("halo_penalty", C.c_uint32), ("net_ids", I32P)]


class ErrorS(C.Structure):
    _fields_ = [("code", C.c_int), ("net", C.c_uint32), ("message", C.c_char * 256)]


class PathsS(C.Structure):
    _fields_ = [("n", C.c_size_t), ("routed", U8P), ("search_cost", U32P), ("offsets", U64P), ("cells", U32P)]


class StatsS(C.Structure):
    _fields_ = [("relax_ms", C.c_double), ("path_ms", C.c_double), ("total_ms", C.c_double),
                ("tile_visits", C.c_uint64), ("local_rounds", C.c_uint64), ("cells_visited", C.c_uint64),
                ("field_cells", C.c_uint64), ("host_rounds", C.c_uint32), ("persistent_ms", C.c_double),
                ("relax_kernel_ms", C.c_double), ("relax_launches", C.c_uint32), ("changed_visits", C.c_uint64),
                ("validate_ms", C.c_double), ("wave_launches", C.c_uint32), ("label_bytes", C.c_uint32),
                ("timed_launches", C.c_uint32)]


class TuningS(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("threads", C.c_uint32), ("max_local", C.c_uint32),
                ("mode", C.c_int), ("graph_rounds", C.c_uint32), ("delta", C.c_uint32), ("sample_ms", C.c_double),
                ("default_ms", C.c_double), ("tried", C.c_uint32)]


class ServerConfigS(C.Structure):
    _fields_ = [("label_bytes", C.c_size_t), ("max_boards", C.c_uint32), ("max_layers", C.c_uint32),
                ("max_inflight", C.c_uint32), ("tile_w", C.c_uint32), ("tile_h", C.c_uint32),
                ("threads", C.c_uint32), ("admit_below_tiles", C.c_uint32), ("admit_small_tiles", C.c_uint32)]


class RequestStatsS(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("solve_ms", C.c_double), ("total_ms", C.c_double),
                ("rounds", C.c_uint32), ("peers_max", C.c_uint32)]


class ServerStatsS(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("rounds", C.c_uint64), ("tile_visits", C.c_uint64), ("kernel_ms", C.c_double),
                ("label_bytes", C.c_size_t), ("label_bytes_peak", C.c_size_t), ("inflight_peak", C.c_uint32),
                ("admitted_ahead", C.c_uint64)]


class SynthS(C.Structure):
    _fields_ = [("dims", Dims), ("cost", U32P), ("board_constraint", U8P), ("additive", U32P),
                ("halo_owner", I32P), ("halo_penalty", C.c_uint32), ("x_edges", U32P), ("y_edges", U32P),
                ("via_edges", U32P), ("via_allowed", U8P), ("n_nets", C.c_size_t), ("src", U32P), ("dst", U32P),
                ("pin_off", U64P), ("pins", U32P), ("windows", C.POINTER(WindowS))]


class NegotiationS(C.Structure):
    _fields_ = [
        ("history", U32P), ("present", U32P), ("present_halo", U32P), ("n_cells", C.c_size_t),
        ("pfac", C.c_uint32), ("fused", C.c_int),
        ("own_path_off", C.POINTER(C.c_uint64)), ("own_path", U32P),
        ("own_halo_off", C.POINTER(C.c_uint64)), ("own_halo", U32P), ("own_halo_count", U32P),
    ]


class SrjParamsS(C.Structure):
    _fields_ = [("mode", C.c_int), ("resolution", C.c_double), ("layers", C.c_uint32),
                ("clearance_cells", C.c_uint32), ("min_clearance_mm", C.c_double), ("via_pad_mm", C.c_double),
                ("cell_budget", C.c_uint64), ("model", C.c_int), ("rules", C.c_void_p)]


class SrjNetS(C.Structure):
    _fields_ = [("name", C.c_char_p), ("src", C.c_uint32), ("dst", C.c_uint32), ("pads", U32P),
                ("n_pads", C.c_size_t), ("via_pads", U32P), ("n_via_pads", C.c_size_t)]


class SrjRulesS(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("trace_width_mm", "obstacle_margin_mm", "trace_to_pad_clearance_mm",
                                          "via_to_pad_clearance_mm", "pad_to_pad_clearance_mm",
                                          "via_pad_diameter_mm", "via_hole_diameter_mm",
                                          "via_hole_to_hole_clearance_mm")]


class SrjProblemS(C.Structure):
    _fields_ = [("dims", Dims), ("x_lines", C.POINTER(C.c_double)), ("y_lines", C.POINTER(C.c_double)),
                ("cost", U32P), ("via_forbidden", U8P), ("board_constraint", U8P), ("n_nets", C.c_size_t),
                ("nets", C.POINTER(SrjNetS)), ("n_pins", C.c_size_t), ("pin_cells", U32P),
                ("pin_xy", C.POINTER(C.c_double)), ("resolution", C.c_double), ("trace_width_mm", C.c_double),
                ("min_clearance_mm", C.c_double), ("via_pad_mm", C.c_double), ("clearance_cells", C.c_uint32),
                ("exact_model", C.c_int), ("physical", C.c_int), ("rules", SrjRulesS),
                ("fill_dropped", C.c_uint32)]


class MrfError(RuntimeError):
    def __init__(self, code: int, message: str, net: int = 0):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.net = net


def build() -> None:
    """Compile libmrfield.so for gfx950 (hipcc, in-tree)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MrfError(BACKEND_UNAVAILABLE, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    E = C.POINTER(ErrorS)
    sig = {
        "mrf_abi_version": (C.c_int, []),
        "mrf_open": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), E]),
        "mrf_close": (None, [C.c_void_p]),
        "mrf_set_tiles": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, E]),
        "mrf_set_mode": (C.c_int, [C.c_void_p, C.c_int]),
        "mrf_set_delta": (C.c_int, [C.c_void_p, C.c_uint32]),
        "mrf_set_round_timing": (C.c_int, [C.c_void_p, C.c_uint32]),
        "mrf_last_stats": (C.c_int, [C.c_void_p, C.POINTER(StatsS)]),
        "mrf_route_isolated_batch": (C.c_int, [C.c_void_p, Dims, C.POINTER(GridS), C.POINTER(NetS), C.c_size_t,
                                               C.POINTER(WindowS), C.POINTER(EdgesS), C.POINTER(PathsS), E]),
        "mrf_route_isolated_pass": (C.c_int, [C.c_void_p, Dims, C.POINTER(GridS), C.POINTER(NetS), C.c_size_t,
                                              C.POINTER(WindowS), C.POINTER(EdgesS), C.POINTER(PathsS), E]),
        "mrf_sweep_fields": (C.c_int, [C.c_void_p, Dims, C.POINTER(GridS), U32P, C.c_size_t, U32P, E]),
        "mrf_router_route": (C.c_int, [C.c_void_p, Dims, C.POINTER(GridS), C.POINTER(NetS), C.c_size_t,
                                       C.POINTER(PathsS), E]),
        "mrf_fields": (C.c_int, [C.c_void_p, Dims, C.POINTER(GridS), C.POINTER(EdgesS), C.POINTER(PricingS),
                                 C.POINTER(U32P), C.POINTER(C.c_size_t), U32P, C.c_size_t, C.POINTER(WindowS), U32P,
                                 U32P, E]),
        "mrf_free_paths": (None, [C.POINTER(PathsS)]),
        "mrf_set_label_budget": (C.c_int, [C.c_void_p, C.c_size_t]),
        "mrf_get_label_budget": (C.c_size_t, [C.c_void_p]),
        "mrf_congestion_from": (C.c_int, [Dims, C.POINTER(PathsS), U32P, E]),
        "mrf_negotiation_settle": (C.c_int, [C.c_size_t, C.c_size_t, U64P, U32P, C.POINTER(C.c_int64), C.c_uint32,
                                             U32P, U32P, U8P, U32P, C.POINTER(C.c_size_t), E]),
        "mrf_board_create": (C.c_int, [C.c_void_p, Dims, C.POINTER(GridS), C.POINTER(EdgesS), C.POINTER(PricingS),
                                       C.POINTER(C.c_void_p), E]),
        "mrf_board_destroy": (None, [C.c_void_p]),
        "mrf_board_bytes": (C.c_size_t, [C.c_void_p]),
        "mrf_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(NetS), C.c_size_t, C.POINTER(WindowS), C.c_int,
                                C.POINTER(PathsS), E]),
        "mrf_solve_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(U32P), C.POINTER(C.c_size_t), U32P, I32P,
                                       C.c_size_t, C.POINTER(WindowS), E]),
        "mrf_read_field": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, U32P, U32P, E]),
        "mrf_board_set_additive": (C.c_int, [C.c_void_p, U32P, C.c_size_t, E]),
        "mrf_route_legal_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(NetS), C.c_size_t, C.POINTER(WindowS),
                                            C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_size_t,
                                            C.POINTER(C.c_int64), C.POINTER(PathsS), E]),
        "mrf_route_negotiated_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(NetS), C.c_size_t,
                                                 C.POINTER(WindowS), C.c_int, C.POINTER(NegotiationS),
                                                 C.POINTER(PathsS), E]),
        "mrf_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(NetS), C.c_size_t, C.POINTER(WindowS),
                                   C.c_size_t, C.POINTER(TuningS), E]),
        "mrf_server_create": (C.c_int, [C.c_int, C.POINTER(ServerConfigS), C.POINTER(C.c_void_p), E]),
        "mrf_server_destroy": (None, [C.c_void_p]),
        "mrf_server_add_board": (C.c_int, [C.c_void_p, Dims, C.POINTER(GridS), C.POINTER(EdgesS),
                                           C.POINTER(PricingS), C.POINTER(C.c_uint32), E]),
        "mrf_server_remove_board": (C.c_int, [C.c_void_p, C.c_uint32, E]),
        "mrf_server_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(NetS), C.c_size_t, C.POINTER(WindowS),
                                        C.c_int, I32P, C.POINTER(C.c_uint64), E]),
        "mrf_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, C.POINTER(PathsS),
                                      C.POINTER(RequestStatsS), E]),
        "mrf_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStatsS)]),
        "mrf_synth_config": (C.c_int, [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(SynthS), E]),
        "mrf_synth_board": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(SynthS)]),
        "mrf_free_synth": (None, [C.POINTER(SynthS)]),
        "mrf_solve_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(NetS), C.c_size_t, C.POINTER(WindowS),
                                        C.c_int, U8P, U32P, U64P, U32P, C.c_size_t, C.POINTER(C.c_size_t), E]),
        "mrf_noalloc_scratch_bytes": (C.c_size_t, []),
        "mrf_srj_rasterize": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(SrjParamsS),
                                        C.POINTER(C.POINTER(SrjProblemS)), E]),
        "mrf_srj_free": (None, [C.POINTER(SrjProblemS)]),
        "mrf_srj_uniform_rules": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(SrjRulesS), C.POINTER(C.c_int), E]),
        "mrf_srj_outline_plane": (C.c_int, [C.c_char_p, C.c_size_t, C.c_double, C.c_double, C.POINTER(C.c_double),
                                            C.c_size_t, C.POINTER(C.c_double), C.c_size_t, U8P, C.POINTER(C.c_int),
                                            E]),
        "mrf_srj_enforce_budget": (C.c_int, [C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_double), C.c_size_t,
                                             C.POINTER(C.c_double), C.POINTER(C.c_size_t), C.POINTER(C.c_double),
                                             C.POINTER(C.c_size_t), C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols() -> List[str]:
    return [n for n in dir(lib()) if n.startswith("mrf_")]


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _p32(a):
    return a.ctypes.data_as(U32P) if a is not None else None


def _p8(a):
    return a.ctypes.data_as(U8P) if a is not None else None


def _check(rc: int, err: ErrorS):
    if rc != OK:
        raise MrfError(rc, err.message.decode(errors="replace"), err.net)


@dataclass
class Grid:
    """mr-core Grid (crates/mr-core/src/lib.rs:394-408)."""

    w: int
    h: int
    layers: int
    cost: np.ndarray
    via_forbidden: Optional[np.ndarray] = None
    board: Optional[np.ndarray] = None

    def __post_init__(self):
        self.cost = _u32(self.cost)
        if self.via_forbidden is not None:
            self.via_forbidden = np.ascontiguousarray(np.asarray(self.via_forbidden, dtype=np.uint8))
        if self.board is not None:
            self.board = np.ascontiguousarray(np.asarray(self.board, dtype=np.uint8))

    @property
    def n(self):
        return self.w * self.h * self.layers

    def dims(self):
        return Dims(self.w, self.h, self.layers)

    def cstruct(self):
        return GridS(_p32(self.cost), len(self.cost),
                     _p8(self.via_forbidden), 0 if self.via_forbidden is None else len(self.via_forbidden),
                     _p8(self.board), 0 if self.board is None else len(self.board))


@dataclass
class Edges:
    """Rounded request prices (negotiated.rs:530-537); None via = forbidden step."""

    x: Sequence[int]
    y: Sequence[int]
    via: Sequence[Optional[int]]

    def cstruct(self):
        self._x = _u32(list(self.x) + [0])
        self._y = _u32(list(self.y) + [0])
        self._v = _u32([0 if v is None else v for v in self.via] + [0])
        self._a = np.ascontiguousarray(np.array([0 if v is None else 1 for v in self.via] + [0], dtype=np.uint8))
        return EdgesS(_p32(self._x), len(self.x), _p32(self._y), len(self.y), _p32(self._v), _p8(self._a),
                      len(self.via))


@dataclass
class Net:
    src: int
    dst: int
    pads: Sequence[int] = ()
    via_pads: Sequence[int] = ()


@dataclass
class Pricing:
    additive: Optional[np.ndarray] = None
    halo_owner: Optional[np.ndarray] = None
    halo_penalty: int = 0
    net_ids: Optional[Sequence[int]] = None

    def cstruct(self):
        self._a = _u32(self.additive) if self.additive is not None else None
        self._h = (np.ascontiguousarray(np.asarray(self.halo_owner, dtype=np.int32))
                   if self.halo_owner is not None else None)
        self._i = (np.ascontiguousarray(np.asarray(self.net_ids, dtype=np.int32))
                   if self.net_ids is not None else None)
        return PricingS(_p32(self._a), 0 if self._a is None else len(self._a),
                        self._h.ctypes.data_as(I32P) if self._h is not None else None,
                        0 if self._h is None else len(self._h), self.halo_penalty,
                        self._i.ctypes.data_as(I32P) if self._i is not None else None)


_NO_PADS = np.zeros(1, dtype=np.uint32)


_NET_DTYPE = np.dtype([("src", np.uint32), ("dst", np.uint32), ("pads", np.uintp), ("n_pads", np.uintp),
                       ("via_pads", np.uintp), ("n_via_pads", np.uintp)], align=True)


def _nets(nets: Sequence[Net]):
    if all(not n.pads and not n.via_pads for n in nets):
        # vectorised: one numpy record per net, laid out as mrf_net
        assert _NET_DTYPE.itemsize == C.sizeof(NetS)
        rec = np.zeros(max(len(nets), 1), dtype=_NET_DTYPE)
        if nets:
            rec["src"] = [n.src for n in nets]
            rec["dst"] = [n.dst for n in nets]
        rec["pads"] = _NO_PADS.ctypes.data
        rec["via_pads"] = _NO_PADS.ctypes.data
        return C.cast(rec.ctypes.data, C.POINTER(NetS)), [rec, _NO_PADS]
    keep = [_NO_PADS]
    arr = (NetS * max(len(nets), 1))()
    empty = _p32(_NO_PADS)
    for i, n in enumerate(nets):
        if n.pads:
            p = _u32(list(n.pads) + [0])
            keep.append(p)
            pp = _p32(p)
        else:
            pp = empty
        if n.via_pads:
            v = _u32(list(n.via_pads) + [0])
            keep.append(v)
            vp = _p32(v)
        else:
            vp = empty
        arr[i] = NetS(n.src, n.dst, pp, len(n.pads), vp, len(n.via_pads))
    return arr, keep


def _csr(per_net, n: int):
    """Per-net cell lists as (u64 offsets[n + 1], u32 cells): from a flat
    (offsets, cells) pair, or from a sequence of per-net cell sequences."""
    if isinstance(per_net, tuple):
        off = np.ascontiguousarray(per_net[0], dtype=np.uint64)
        cells = np.ascontiguousarray(per_net[1], dtype=np.uint32)
        if off.size != n + 1 or (n and int(off[-1]) > cells.size):
            raise MrfError(BACKEND_UNAVAILABLE, "own-footprint offsets do not match the nets")
    else:
        arrs = [np.asarray(p).ravel() for p in per_net]
        off = np.zeros(n + 1, dtype=np.uint64)
        np.cumsum([a.size for a in arrs], out=off[1:])
        cells = np.concatenate(arrs).astype(np.uint32) if off[-1] else np.zeros(0, dtype=np.uint32)
    if not cells.size:
        cells = np.zeros(1, dtype=np.uint32)
    return off, cells


def _windows(wins, n):
    """Windows as mrf_window records: from (x0, y0, x1, y1) tuples, or from an
    (n, 4) integer array (passed without per-window conversion)."""
    if wins is None:
        return None
    if isinstance(wins, np.ndarray):
        a = np.ascontiguousarray(wins, dtype=np.uint32).reshape(-1, 4)
        if a.shape[0] != n:
            raise MrfError(BACKEND_UNAVAILABLE, "one window per net")
        if n == 0:
            a = np.zeros((1, 4), dtype=np.uint32)
        return a.ctypes.data_as(C.POINTER(WindowS))
    return (WindowS * max(n, 1))(*[WindowS(*w) for w in wins])


class PathBatch:
    """Paths of one call as flat arrays: path i is cells[offsets[i]:offsets[i+1]]
    when routed[i], with search cost cost[i]."""

    def __init__(self, routed, cost, offsets, cells):
        self.routed, self.cost, self.offsets, self.cells = routed, cost, offsets, cells

    def __len__(self):
        return len(self.routed)

    def path(self, i: int) -> Optional[Tuple[List[int], int]]:
        if not self.routed[i]:
            return None
        return self.cells[self.offsets[i]:self.offsets[i + 1]].tolist(), int(self.cost[i])


def _take_arrays(p: PathsS) -> PathBatch:
    try:
        n = p.n
        off = np.ctypeslib.as_array(p.offsets, shape=(n + 1,)).copy()
        total = int(off[-1])
        cells = np.ctypeslib.as_array(p.cells, shape=(max(total, 1),))[:total].copy()
        routed = np.ctypeslib.as_array(p.routed, shape=(max(n, 1),))[:n].copy()
        cost = np.ctypeslib.as_array(p.search_cost, shape=(max(n, 1),))[:n].copy()
    finally:
        lib().mrf_free_paths(C.byref(p))
    return PathBatch(routed, cost, off, cells)


def _take_paths(p: PathsS) -> List[Optional[Tuple[List[int], int]]]:
    out = []
    try:
        n = p.n
        if n:
            off = np.ctypeslib.as_array(p.offsets, shape=(n + 1,)).copy()
            total = int(off[-1])
            cells = np.ctypeslib.as_array(p.cells, shape=(max(total, 1),))[:total].copy()
            routed = np.ctypeslib.as_array(p.routed, shape=(n,)).copy()
            cost = np.ctypeslib.as_array(p.search_cost, shape=(n,)).copy()
            for i in range(n):
                if routed[i]:
                    out.append((cells[off[i]:off[i + 1]].tolist(), int(cost[i])))
                else:
                    out.append(None)
    finally:
        lib().mrf_free_paths(C.byref(p))
    return out


class Engine:
    """One context (HIP stream + workspace) on one device."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        err = ErrorS()
        _check(lib().mrf_open(device, C.byref(self._h), C.byref(err)), err)

    def close(self):
        if self._h:
            lib().mrf_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tiles(self, tw=16, th=16, threads=64, max_local=32):
        err = ErrorS()
        _check(lib().mrf_set_tiles(self._h, tw, th, threads, max_local, C.byref(err)), err)

    def set_mode(self, mode: int):
        if lib().mrf_set_mode(self._h, mode) != OK:
            raise MrfError(BACKEND_UNAVAILABLE, "bad mode")

    def set_delta(self, delta: int):
        """Delta-stepping window of the plain form (MAX = off)."""
        if lib().mrf_set_delta(self._h, delta) != OK:
            raise MrfError(BACKEND_UNAVAILABLE, "bad delta")

    def set_round_timing(self, every: int):
        """Bracket each plain-form round with HIP events in every `every`-th
        solving call (1 = every call; 0 = none, the library default)."""
        if lib().mrf_set_round_timing(self._h, int(every)) != OK:
            raise MrfError(BACKEND_UNAVAILABLE, "bad round timing")

    def set_label_budget(self, nbytes: int):
        """Label bytes one device batch may hold (0 = the library default)."""
        if lib().mrf_set_label_budget(self._h, int(nbytes)) != OK:
            raise MrfError(BACKEND_UNAVAILABLE, "bad label budget")

    def label_budget(self) -> int:
        return int(lib().mrf_get_label_budget(self._h))

    def stats(self) -> dict:
        s = StatsS()
        lib().mrf_last_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in StatsS._fields_}

    # -- provider seam -------------------------------------------------------
    def _isolated(self, fn, grid: Grid, nets, windows, edges: Edges, take=None):
        arr, keep = _nets(nets)
        g = grid.cstruct()
        e = edges.cstruct()
        w = _windows(windows, len(nets))
        out = PathsS()
        err = ErrorS()
        rc = fn(self._h, grid.dims(), C.byref(g), arr, len(nets), w, C.byref(e), C.byref(out), C.byref(err))
        _check(rc, err)
        return (take or _take_paths)(out)

    def route_isolated_batch(self, grid, nets, windows, edges):
        return self._isolated(lib().mrf_route_isolated_batch, grid, nets, windows, edges)

    def route_isolated_batch_arrays(self, grid, nets, windows, edges) -> PathBatch:
        """Like route_isolated_batch(), with the paths left as flat numpy arrays."""
        return self._isolated(lib().mrf_route_isolated_batch, grid, nets, windows, edges, _take_arrays)

    def route_isolated_pass(self, grid, nets, windows, edges):
        return self._isolated(lib().mrf_route_isolated_pass, grid, nets, windows, edges)

    def router_route(self, grid: Grid, nets):
        arr, keep = _nets(nets)
        g = grid.cstruct()
        out = PathsS()
        err = ErrorS()
        _check(lib().mrf_router_route(self._h, grid.dims(), C.byref(g), arr, len(nets), C.byref(out),
                                      C.byref(err)), err)
        return _take_paths(out)

    def sweep_fields(self, grid: Grid, sources: Sequence[int]) -> np.ndarray:
        s = _u32(list(sources) + [0])
        out = np.zeros((max(len(sources), 1), grid.n), dtype=np.uint32)
        g = grid.cstruct()
        err = ErrorS()
        _check(lib().mrf_sweep_fields(self._h, grid.dims(), C.byref(g), _p32(s), len(sources), _p32(out),
                                      C.byref(err)), err)
        return out[: len(sources)]

    def fields(self, grid: Grid, edges: Edges, sources: Sequence[Sequence[int]], dsts=None, windows=None,
               pricing: Optional[Pricing] = None, want_hops=True):
        n = len(sources)
        srcs = [_u32(list(s) + [0]) for s in sources]
        sp = (U32P * max(n, 1))(*[_p32(s) for s in srcs])
        ns = (C.c_size_t * max(n, 1))(*[len(s) for s in sources])
        d = _u32(list(dsts) + [0]) if dsts is not None else None
        dist = np.zeros((max(n, 1), grid.n), dtype=np.uint32)
        hops = np.zeros((max(n, 1), grid.n), dtype=np.uint32) if want_hops else None
        g = grid.cstruct()
        e = edges.cstruct()
        pr = pricing.cstruct() if pricing is not None else None
        err = ErrorS()
        _check(lib().mrf_fields(self._h, grid.dims(), C.byref(g), C.byref(e), C.byref(pr) if pr else None, sp, ns,
                                _p32(d), n, _windows(windows, n), _p32(dist), _p32(hops), C.byref(err)), err)
        return dist[:n], (hops[:n] if hops is not None else None)


def congestion_from(dims: Tuple[int, int, int], paths) -> np.ndarray:
    """mrf_congestion_from over a list of paths (each a list of cells, a
    (cells, cost) pair, or None for an unrouted net)."""
    w, h, L = dims
    cells_l = [(list(p[0]) if isinstance(p, tuple) else list(p)) if p is not None else [] for p in paths]
    n = len(cells_l)
    routed = np.array([1 if p is not None else 0 for p in paths] + [0], dtype=np.uint8)
    off = np.zeros(n + 1, dtype=np.uint64)
    for i, c in enumerate(cells_l):
        off[i + 1] = off[i] + len(c)
    flat = _u32([v for c in cells_l for v in c] + [0])
    cost = np.zeros(n + 1, dtype=np.uint32)
    ps = PathsS(n, routed.ctypes.data_as(U8P), _p32(cost), off.ctypes.data_as(U64P), _p32(flat))
    out = np.zeros(max(w * h * L, 1), dtype=np.uint32)
    err = ErrorS()
    _check(lib().mrf_congestion_from(Dims(w, h, L), C.byref(ps), _p32(out), C.byref(err)), err)
    return out[: w * h * L]


class Board:
    """A board whose planes stay resident in HBM across solves."""

    def __init__(self, eng: Engine, grid: Grid, edges: Edges, pricing: Optional[Pricing] = None):
        self.eng = eng
        self.grid = grid
        self._h = C.c_void_p()
        g = grid.cstruct()
        e = edges.cstruct()
        self._pr = pricing.cstruct() if pricing is not None else None
        err = ErrorS()
        _check(lib().mrf_board_create(eng._h, grid.dims(), C.byref(g), C.byref(e),
                                      C.byref(self._pr) if self._pr else None, C.byref(self._h), C.byref(err)), err)

    def close(self):
        if self._h:
            lib().mrf_board_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bytes(self) -> int:
        return int(lib().mrf_board_bytes(self._h))

    def _solve(self, nets, windows, retry_full):
        arr, keep = _nets(nets)
        out = PathsS()
        err = ErrorS()
        _check(lib().mrf_solve(self.eng._h, self._h, arr, len(nets), _windows(windows, len(nets)),
                               1 if retry_full else 0, C.byref(out), C.byref(err)), err)
        return out

    def solve(self, nets, windows=None, retry_full=True):
        return _take_paths(self._solve(nets, windows, retry_full))

    def autotune(self, nets, windows=None, sample=64) -> dict:
        """Measure tile shapes / workgroup sizes / round caps / forms on a
        sample of `nets`; the fastest stays set on the engine."""
        arr, keep = _nets(nets)
        t = TuningS()
        err = ErrorS()
        _check(lib().mrf_autotune(self.eng._h, self._h, arr, len(nets), _windows(windows, len(nets)), sample,
                                  C.byref(t), C.byref(err)), err)
        return {k: getattr(t, k) for k, _ in TuningS._fields_}

    def solve_noalloc(self, nets, windows=None, retry_full=True, cells_cap=1 << 20) -> PathBatch:
        """mrf_solve_noalloc: the same solve with every device buffer carved
        from the library's scratch array; paths into arrays owned here."""
        arr, keep = _nets(nets)
        n = len(nets)
        routed = np.zeros(max(n, 1), dtype=np.uint8)
        cost = np.zeros(max(n, 1), dtype=np.uint32)
        off = np.zeros(n + 1, dtype=np.uint64)
        cells = np.zeros(max(cells_cap, 1), dtype=np.uint32)
        n_cells = C.c_size_t(0)
        err = ErrorS()
        _check(lib().mrf_solve_noalloc(self.eng._h, self._h, arr, n, _windows(windows, n), 1 if retry_full else 0,
                                       routed.ctypes.data_as(U8P), _p32(cost), off.ctypes.data_as(U64P), _p32(cells),
                                       cells_cap, C.byref(n_cells), C.byref(err)), err)
        return PathBatch(routed[:n], cost[:n], off, cells[:n_cells.value])

    def solve_arrays(self, nets, windows=None, retry_full=True) -> PathBatch:
        """Like solve(), with the paths left as flat numpy arrays."""
        return _take_arrays(self._solve(nets, windows, retry_full))

    def set_additive(self, additive) -> None:
        """Replace the board's additive price plane (w*h*layers u32); None drops it."""
        err = ErrorS()
        if additive is None:
            _check(lib().mrf_board_set_additive(self._h, None, 0, C.byref(err)), err)
            return
        a = _u32(additive)
        _check(lib().mrf_board_set_additive(self._h, _p32(a), a.size, C.byref(err)), err)

    def route_legal(self, nets, owner, halo, own_groups, windows=None, retry_full=True, arrays=False):
        """Legalization searches (route_legal_once, negotiated.rs:2885-3052)
        against committed copper owners and clearance halo (w*h*layers group
        ids, -1 free, halo -2 mixed; either may be None), one group per net."""
        arr, keep = _nets(nets)
        P64 = C.POINTER(C.c_int64)
        o = np.ascontiguousarray(np.asarray(owner, dtype=np.int64)) if owner is not None else None
        h = np.ascontiguousarray(np.asarray(halo, dtype=np.int64)) if halo is not None else None
        n = o.size if o is not None else (h.size if h is not None else 0)
        if o is not None and h is not None and o.size != h.size:
            raise MrfError(MALFORMED_GRID, "owner and halo planes differ in length")
        gid = np.ascontiguousarray(np.asarray(list(own_groups), dtype=np.int64))
        if gid.size != len(nets):
            raise MrfError(BACKEND_UNAVAILABLE, "one group per net")
        out = PathsS()
        err = ErrorS()
        _check(lib().mrf_route_legal_batch(self.eng._h, self._h, arr, len(nets), _windows(windows, len(nets)),
                                           1 if retry_full else 0, o.ctypes.data_as(P64) if o is not None else None,
                                           h.ctypes.data_as(P64) if h is not None else None, n,
                                           gid.ctypes.data_as(P64) if gid.size else None, C.byref(out),
                                           C.byref(err)), err)
        return _take_arrays(out) if arrays else _take_paths(out)

    def route_negotiated(self, nets, history, present, pfac, *, present_halo=None, fused=False, own_paths=None,
                         own_halos=None, windows=None, retry_full=True, arrays=False):
        """One negotiation iteration's searches (route_negotiated per net,
        negotiated.rs:2746-2856) in one device batch.

        own_paths: per net its own copper (sequence of cells), or None for the
        batch passing no own path.  own_halos: per net its own halo, a
        {cell: count} dict or cells repeated by multiplicity, or None.  Either
        may also be given flat as a pair (offsets, cells): net i's cells are
        cells[offsets[i]:offsets[i + 1]] (halo cells repeated by multiplicity).
        """
        arr, keep = _nets(nets)
        hist, pres = _u32(history), _u32(present)
        halo = _u32(present_halo) if present_halo is not None else None
        if pres.size != hist.size or (halo is not None and halo.size != hist.size):
            raise MrfError(MALFORMED_GRID, "negotiation planes differ in length")
        ng = NegotiationS(_p32(hist), _p32(pres), _p32(halo), hist.size, int(pfac), 1 if fused else 0)
        if own_paths is not None:
            poff, pc = _csr(own_paths, len(nets))
            keep += [poff, pc]
            ng.own_path_off = poff.ctypes.data_as(C.POINTER(C.c_uint64))
            ng.own_path = _p32(pc)
        if isinstance(own_halos, tuple):
            hoff, hc = _csr(own_halos, len(nets))
            keep += [hoff, hc]
            ng.own_halo_off = hoff.ctypes.data_as(C.POINTER(C.c_uint64))
            ng.own_halo = _p32(hc)
        elif own_halos is not None:
            cells, counts = [], []
            weighted = any(isinstance(h, dict) for h in own_halos)
            for hset in own_halos:
                if isinstance(hset, dict):
                    cells.append(np.fromiter(hset.keys(), dtype=np.uint32, count=len(hset)))
                    counts.append(np.fromiter(hset.values(), dtype=np.uint32, count=len(hset)))
                else:
                    a = np.asarray(hset, dtype=np.uint32).ravel()
                    cells.append(a)
                    if weighted:
                        counts.append(np.ones(a.size, dtype=np.uint32))
            hoff = np.zeros(len(nets) + 1, dtype=np.uint64)
            np.cumsum([a.size for a in cells], out=hoff[1:])
            hc = np.concatenate(cells) if hoff[-1] else np.zeros(1, dtype=np.uint32)
            keep += [hoff, hc]
            ng.own_halo_off = hoff.ctypes.data_as(C.POINTER(C.c_uint64))
            ng.own_halo = _p32(hc)
            if weighted:
                hn = np.concatenate(counts) if hoff[-1] else np.zeros(1, dtype=np.uint32)
                keep.append(hn)
                ng.own_halo_count = _p32(hn)
        out = PathsS()
        err = ErrorS()
        _check(lib().mrf_route_negotiated_batch(self.eng._h, self._h, arr, len(nets), _windows(windows, len(nets)),
                                                1 if retry_full else 0, C.byref(ng), C.byref(out), C.byref(err)), err)
        return _take_arrays(out) if arrays else _take_paths(out)

    def solve_fields(self, sources, dsts=None, net_ids=None, windows=None):
        n = len(sources)
        srcs = [_u32(list(s) + [0]) for s in sources]
        sp = (U32P * max(n, 1))(*[_p32(s) for s in srcs])
        ns = (C.c_size_t * max(n, 1))(*[len(s) for s in sources])
        d = _u32(list(dsts) + [0]) if dsts is not None else None
        ids = np.ascontiguousarray(np.asarray(list(net_ids) + [0], dtype=np.int32)) if net_ids is not None else None
        err = ErrorS()
        _check(lib().mrf_solve_fields(self.eng._h, self._h, sp, ns, _p32(d),
                                      ids.ctypes.data_as(I32P) if ids is not None else None, n,
                                      _windows(windows, n), C.byref(err)), err)

    def read_field(self, i: int, want_hops=True):
        dist = np.zeros(self.grid.n, dtype=np.uint32)
        hops = np.zeros(self.grid.n, dtype=np.uint32) if want_hops else None
        err = ErrorS()
        _check(lib().mrf_read_field(self.eng._h, self._h, i, _p32(dist), _p32(hops), C.byref(err)), err)
        return dist, hops


def negotiation_settle(path_off, path_cells, groups, scale: int, present: np.ndarray, history: np.ndarray):
    """mrf_negotiation_settle: recount `present` (uint32, in place) from the
    settled paths (path i = path_cells[path_off[i]:path_off[i + 1]]), add
    `scale` to `history` (uint32, in place, saturating) on every cell held by
    two or more groups, and return (over-used cells ascending, per net whether
    its path holds one)."""
    if present.dtype != np.uint32 or history.dtype != np.uint32 or not present.flags.c_contiguous \
            or not history.flags.c_contiguous or present.size != history.size:
        raise MrfError(BACKEND_UNAVAILABLE, "present and history must be contiguous uint32 planes of one size")
    off = np.ascontiguousarray(path_off, dtype=np.uint64)
    cells = np.ascontiguousarray(path_cells, dtype=np.uint32)
    grp = np.ascontiguousarray(groups, dtype=np.int64)
    n = off.size - 1
    if grp.size != n:
        raise MrfError(BACKEND_UNAVAILABLE, "one group per path")
    touches = np.zeros(max(n, 1), dtype=np.uint8)
    over = np.zeros(max(present.size, 1), dtype=np.uint32)
    n_over = C.c_size_t(0)
    err = ErrorS()
    _check(lib().mrf_negotiation_settle(present.size, n, off.ctypes.data_as(U64P), _p32(cells if cells.size else
                                        np.zeros(1, dtype=np.uint32)), grp.ctypes.data_as(C.POINTER(C.c_int64)),
                                        int(scale), _p32(present), _p32(history),
                                        touches.ctypes.data_as(U8P), _p32(over), C.byref(n_over), C.byref(err)), err)
    return over[:n_over.value].astype(np.int64), touches[:n].astype(bool)


NOT_READY = 4


class Server:
    """Serving loop: resident boards, HBM-budget admission, continuous batching."""

    def __init__(self, device: int = 0, label_bytes: int = 0, max_boards: int = 0, max_layers: int = 0,
                 max_inflight: int = 0, tile_w: int = 0, tile_h: int = 0, threads: int = 0,
                 admit_below_tiles: int = 0, admit_small_tiles: int = 0):
        self._h = C.c_void_p()
        cfg = ServerConfigS(label_bytes, max_boards, max_layers, max_inflight, tile_w, tile_h, threads,
                            admit_below_tiles, admit_small_tiles)
        err = ErrorS()
        _check(lib().mrf_server_create(device, C.byref(cfg), C.byref(self._h), C.byref(err)), err)
        self._boards = {}

    def close(self):
        if self._h:
            lib().mrf_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_board(self, grid: Grid, edges: Edges, pricing: Optional[Pricing] = None) -> int:
        g = grid.cstruct()
        e = edges.cstruct()
        pr = pricing.cstruct() if pricing is not None else None
        bid = C.c_uint32()
        err = ErrorS()
        _check(lib().mrf_server_add_board(self._h, grid.dims(), C.byref(g), C.byref(e),
                                          C.byref(pr) if pr else None, C.byref(bid), C.byref(err)), err)
        return int(bid.value)

    def remove_board(self, board_id: int) -> None:
        err = ErrorS()
        _check(lib().mrf_server_remove_board(self._h, board_id, C.byref(err)), err)

    def submit(self, board_id: int, nets, windows=None, retry_full=True, net_ids=None) -> int:
        arr, keep = _nets(nets)
        ids = np.ascontiguousarray(np.asarray(list(net_ids), dtype=np.int32)) if net_ids is not None else None
        rid = C.c_uint64()
        err = ErrorS()
        _check(lib().mrf_server_submit(self._h, board_id, arr, len(nets), _windows(windows, len(nets)),
                                       1 if retry_full else 0, ids.ctypes.data_as(I32P) if ids is not None else None,
                                       C.byref(rid), C.byref(err)), err)
        return int(rid.value)

    def wait(self, request_id: int, timeout_ms: int = -1, arrays: bool = False):
        """(paths, stats); paths as solve() returns them (or a PathBatch when
        arrays is true).  None when the request has not finished in time."""
        out = PathsS()
        st = RequestStatsS()
        err = ErrorS()
        rc = lib().mrf_server_wait(self._h, request_id, timeout_ms, C.byref(out), C.byref(st), C.byref(err))
        if rc == NOT_READY:
            return None
        _check(rc, err)
        paths = _take_arrays(out) if arrays else _take_paths(out)
        return paths, {k: getattr(st, k) for k, _ in RequestStatsS._fields_}

    def stats(self) -> dict:
        s = ServerStatsS()
        lib().mrf_server_get_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStatsS._fields_}


@dataclass
class Synth:
    grid: Grid
    edges: Edges
    additive: Optional[np.ndarray]
    halo_owner: Optional[np.ndarray]
    halo_penalty: int
    src: np.ndarray
    dst: np.ndarray
    pins: Optional[List[List[int]]]
    windows: Optional[List[Tuple[int, int, int, int]]]

    def nets(self) -> List[Net]:
        return [Net(int(s), int(d)) for s, d in zip(self.src, self.dst)]

    def pricing(self, net_ids=None) -> Optional[Pricing]:
        if self.additive is None and self.halo_owner is None:
            return None
        return Pricing(self.additive, self.halo_owner, self.halo_penalty, net_ids)


def _from_synth(s: SynthS) -> Synth:
    try:
        d = s.dims
        n = d.w * d.h * d.layers
        arr = lambda p, k, dt=None: np.ctypeslib.as_array(p, shape=(k,)).copy()
        cost = arr(s.cost, n)
        board = arr(s.board_constraint, n) if s.board_constraint else None
        add = arr(s.additive, n) if s.additive else None
        halo = arr(s.halo_owner, n) if s.halo_owner else None
        ex = [int(v) for v in arr(s.x_edges, max(d.w, 1))[: d.w - 1]]
        ey = [int(v) for v in arr(s.y_edges, max(d.h, 1))[: d.h - 1]]
        ve = arr(s.via_edges, max(d.layers, 1))[: d.layers - 1]
        va = arr(s.via_allowed, max(d.layers, 1))[: d.layers - 1]
        via = [int(v) if a else None for v, a in zip(ve, va)]
        nn = s.n_nets
        src = arr(s.src, max(nn, 1))[:nn]
        dst = arr(s.dst, max(nn, 1))[:nn]
        pins = None
        if s.pin_off:
            off = arr(s.pin_off, nn + 1)
            allp = arr(s.pins, max(int(off[-1]), 1))
            pins = [[int(c) for c in allp[off[i]:off[i + 1]]] for i in range(nn)]
        wins = None
        if s.windows:
            wins = [(s.windows[i].x0, s.windows[i].y0, s.windows[i].x1, s.windows[i].y1) for i in range(nn)]
        return Synth(Grid(d.w, d.h, d.layers, cost, None, board), Edges(ex, ey, via), add, halo,
                     int(s.halo_penalty), src, dst, pins, wins)
    finally:
        lib().mrf_free_synth(C.byref(s))


def synth_config(config: int, seed: int = 1, scale_nets: int = 0) -> Synth:
    s = SynthS()
    err = ErrorS()
    _check(lib().mrf_synth_config(config, seed, scale_nets, C.byref(s), C.byref(err)), err)
    return _from_synth(s)


def synth_board(side: int, layers: int, nets: int, seed: int) -> Synth:
    s = SynthS()
    rc = lib().mrf_synth_board(side, layers, nets, seed, C.byref(s))
    if rc != OK:
        raise MrfError(rc, "synth_board failed")
    return _from_synth(s)


# ---- SimpleRouteJson boards -------------------------------------------------

SRJ_ROUTE, SRJ_LAYERS, SRJ_PHYSICAL = 0, 1, 2
BOARD_TRACE_NODE, BOARD_VIA_NODE = 1, 2
BOARD_EDGE_NEG_Y, BOARD_EDGE_NEG_X, BOARD_EDGE_POS_X, BOARD_EDGE_POS_Y = 4, 8, 16, 32


@dataclass
class SrjNet:
    """NetEndpoints with its name (crates/mr-core/src/lib.rs:520-537)."""

    name: str
    src: int
    dst: int
    pads: List[int]
    via_pads: List[int]

    def net(self) -> Net:
        return Net(self.src, self.dst, self.pads, self.via_pads)


@dataclass
class SrjProblem:
    """RasterizedProblem (crates/mr-srj/src/lib.rs:1303-1320) plus the rules it
    was built with."""

    w: int
    h: int
    layers: int
    x_lines: np.ndarray
    y_lines: np.ndarray
    cost: np.ndarray
    via_forbidden: Optional[np.ndarray]
    board: Optional[np.ndarray]
    nets: List[SrjNet]
    pins: dict
    resolution: float
    trace_width_mm: float
    min_clearance_mm: float
    via_pad_mm: float
    clearance_cells: int
    exact_model: bool
    physical: bool
    rules: Optional[dict]
    fill_dropped: int

    def grid(self) -> Grid:
        return Grid(self.w, self.h, self.layers, self.cost, self.via_forbidden, self.board)

    def idx(self, x: int, y: int, layer: int = 0) -> int:
        return (layer * self.h + y) * self.w + x

    def is_obstacle(self, cell: int) -> bool:
        return int(self.cost[cell]) == MAX


def _rules_dict(r: SrjRulesS) -> dict:
    d = {k: getattr(r, k) for k, _ in SrjRulesS._fields_}
    for k in ("pad_to_pad_clearance_mm", "via_hole_to_hole_clearance_mm"):
        if d[k] < 0:
            d[k] = None
    return d


def srj_rasterize(doc, mode: int = SRJ_ROUTE, resolution: float = 0.0, layers: int = 0, clearance_cells: int = 0,
                  min_clearance_mm: float = 0.0, via_pad_mm: float = 0.0, cell_budget: int = 0, model: int = 0,
                  rules: Optional[dict] = None) -> SrjProblem:
    """Rasterise a SimpleRouteJson document (str, bytes or a JSON-able object).
    ``mode``: SRJ_ROUTE (the CLI's rule selection), SRJ_LAYERS
    (rasterize_with_layers; all-zero rules = rasterize(); ``model`` 1 / 2
    forces the legacy / exact clearance model), SRJ_PHYSICAL (``rules`` as
    srj_uniform_rules returns them, or None to resolve the problem's own)."""
    import json as _json

    if isinstance(doc, (dict, list)):
        doc = _json.dumps(doc)
    data = doc.encode() if isinstance(doc, str) else bytes(doc)
    rs = None
    if rules is not None:
        rs = SrjRulesS(*[(-1.0 if rules[k] is None else float(rules[k])) for k, _ in SrjRulesS._fields_])
    prm = SrjParamsS(mode, resolution, layers, clearance_cells, min_clearance_mm, via_pad_mm, cell_budget, model,
                     C.cast(C.pointer(rs), C.c_void_p) if rs is not None else None)
    out = C.POINTER(SrjProblemS)()
    err = ErrorS()
    _check(lib().mrf_srj_rasterize(data, len(data), C.byref(prm), C.byref(out), C.byref(err)), err)
    try:
        p = out.contents
        n = p.dims.w * p.dims.h * p.dims.layers
        nets = []
        for i in range(p.n_nets):
            q = p.nets[i]
            nets.append(SrjNet(q.name.decode(), int(q.src), int(q.dst),
                               [int(v) for v in np.ctypeslib.as_array(q.pads, (q.n_pads,))] if q.n_pads else [],
                               [int(v) for v in np.ctypeslib.as_array(q.via_pads, (q.n_via_pads,))]
                               if q.n_via_pads else []))
        pins = {}
        if p.n_pins:
            cells = np.ctypeslib.as_array(p.pin_cells, (p.n_pins,))
            xy = np.ctypeslib.as_array(p.pin_xy, (2 * p.n_pins,))
            pins = {int(c): (float(xy[2 * k]), float(xy[2 * k + 1])) for k, c in enumerate(cells)}
        return SrjProblem(
            int(p.dims.w), int(p.dims.h), int(p.dims.layers),
            np.ctypeslib.as_array(p.x_lines, (p.dims.w,)).copy(), np.ctypeslib.as_array(p.y_lines, (p.dims.h,)).copy(),
            np.ctypeslib.as_array(p.cost, (n,)).copy(),
            np.ctypeslib.as_array(p.via_forbidden, (n,)).copy() if p.via_forbidden else None,
            np.ctypeslib.as_array(p.board_constraint, (n,)).copy() if p.board_constraint else None,
            nets, pins, p.resolution, p.trace_width_mm, p.min_clearance_mm, p.via_pad_mm, int(p.clearance_cells),
            bool(p.exact_model), bool(p.physical), _rules_dict(p.rules) if p.physical else None,
            int(p.fill_dropped))
    finally:
        lib().mrf_srj_free(out)


def srj_uniform_rules(doc) -> Optional[dict]:
    """SimpleRouteJson::uniform_physical_rules (crates/mr-srj/src/lib.rs:286-528)."""
    import json as _json

    if isinstance(doc, (dict, list)):
        doc = _json.dumps(doc)
    data = doc.encode() if isinstance(doc, str) else bytes(doc)
    r = SrjRulesS()
    ok = C.c_int()
    err = ErrorS()
    _check(lib().mrf_srj_uniform_rules(data, len(data), C.byref(r), C.byref(ok), C.byref(err)), err)
    return _rules_dict(r) if ok.value else None


def srj_enforce_budget(x_features, y_features, x_fill, y_fill, budget: int):
    """enforce_budget (crates/mr-srj/src/lib.rs:1047-1128): the kept fill lanes."""
    D = C.POINTER(C.c_double)
    xf = np.ascontiguousarray(x_features, dtype=np.float64)
    yf = np.ascontiguousarray(y_features, dtype=np.float64)
    xs = np.ascontiguousarray(np.array(x_fill, dtype=np.float64))
    ys = np.ascontiguousarray(np.array(y_fill, dtype=np.float64))
    nx, ny = C.c_size_t(len(xs)), C.c_size_t(len(ys))
    rc = lib().mrf_srj_enforce_budget(xf.ctypes.data_as(D), len(xf), yf.ctypes.data_as(D), len(yf),
                                      xs.ctypes.data_as(D), C.byref(nx), ys.ctypes.data_as(D), C.byref(ny),
                                      C.c_uint64(budget))
    if rc != OK:
        raise MrfError(rc, "enforce_budget failed")
    return [float(v) for v in xs[:nx.value]], [float(v) for v in ys[:ny.value]]


def srj_outline_plane(doc, trace_width_mm: float, via_pad_mm: float, x_lines, y_lines):
    """The board-outline mask plane over given lines (board_outline.rs:72-108,
    :223-229): None when the problem declares no outline contract."""
    import json as _json

    if isinstance(doc, (dict, list)):
        doc = _json.dumps(doc)
    data = doc.encode() if isinstance(doc, str) else bytes(doc)
    D = C.POINTER(C.c_double)
    xl = np.ascontiguousarray(x_lines, dtype=np.float64)
    yl = np.ascontiguousarray(y_lines, dtype=np.float64)
    out = np.zeros(len(xl) * len(yl), dtype=np.uint8)
    state = C.c_int()
    err = ErrorS()
    _check(lib().mrf_srj_outline_plane(data, len(data), trace_width_mm, via_pad_mm, xl.ctypes.data_as(D), len(xl),
                                       yl.ctypes.data_as(D), len(yl), _p8(out), C.byref(state), C.byref(err)), err)
    return None if state.value == 0 else out
