/*
 * mrfield.h -- C ABI of the MI355X batched shortest-path field engine
 * (libmrfield.so, built from metal-route_amd/csrc).
 *
 * The ABI is the drop-in boundary for metalroute's isolated-net provider seam
 * and the two secondary seams of its Metal engine.  Paths below are relative to
 * the reference repository root.
 *
 *   mrf_route_isolated_batch  replaces IsolatedRouteProvider::route_isolated_batch
 *                             (crates/mr-cpu/src/negotiated.rs:554-560) as
 *                             implemented by MetalIsolatedProvider
 *                             (crates/mr-cli/src/lib.rs:612-625) over
 *                             metal_isolated_paths_with (:549-609): window pass,
 *                             then full-board retry of unreachable nets.
 *   mrf_route_isolated_pass   replaces mr_metal::metal_route_isolated_batch /
 *                             _ragged (crates/mr-metal/src/lib.rs:414, :695):
 *                             one pass, no retry.
 *   mrf_sweep_fields          replaces mr_metal::metal_sweep_fields
 *                             (crates/mr-metal/src/lib.rs:172-192).
 *   mrf_router_route          replaces MetalRouter::route (Router trait,
 *                             crates/mr-metal/src/lib.rs:1143-1235).
 *   mrf_fields                complete multi-source canonical fields with
 *                             edge prices and negotiation pricing (no single
 *                             reference function: the field form of
 *                             route_legal_once / route_negotiated,
 *                             crates/mr-cpu/src/negotiated.rs:2746-3052).
 *   mrf_board_* / mrf_solve   device-resident boards for repeated solves and
 *                             the serving loop (no reference counterpart; the
 *                             reference re-uploads per call, mr-metal gpu.rs:880-888).
 *
 * Conventions
 *   * Cells use mr-core's mapping idx3(x, y, l) = (l*h + y)*w + x
 *     (crates/mr-core/src/lib.rs:90-116); UINT32_MAX is OBSTACLE / unreachable.
 *   * Every operation is all-or-error: on a non-zero return nothing in an
 *     output is valid (crates/mr-cpu/src/negotiated.rs:539-553).
 *   * Error codes mirror RouterError (crates/mr-core/src/lib.rs:699-708).
 *   * Board masks and via masks are supported (the reference's Metal engine
 *     rejected them, crates/mr-metal/src/lib.rs:51-59, :428-432).
 *   * A context serialises its GPU lane; a call that finds the lane busy returns
 *     MRF_BACKEND_UNAVAILABLE at once (crates/mr-cli/src/lib.rs:499, :620-622).
 */
#ifndef MRFIELD_H
#define MRFIELD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRF_ABI_VERSION 3

/* RouterError (crates/mr-core/src/lib.rs:699-708) */
enum {
    MRF_OK = 0,
    MRF_BACKEND_UNAVAILABLE = 1,
    MRF_INVALID_ENDPOINT = 2,
    MRF_MALFORMED_GRID = 3,
    MRF_NOT_READY = 4 /* serving loop only: the request has not finished yet */
};

/* mr-core Grid::BOARD_* bits (crates/mr-core/src/lib.rs:411-416) */
enum {
    MRF_BOARD_TRACE_NODE = 1 << 0,
    MRF_BOARD_VIA_NODE = 1 << 1,
    MRF_BOARD_EDGE_NEG_Y = 1 << 2,
    MRF_BOARD_EDGE_NEG_X = 1 << 3,
    MRF_BOARD_EDGE_POS_X = 1 << 4,
    MRF_BOARD_EDGE_POS_Y = 1 << 5
};

typedef struct mrf_ctx mrf_ctx;
typedef struct mrf_board mrf_board;

/* Dims (crates/mr-core/src/lib.rs:40-60) */
typedef struct {
    uint32_t w, h, layers;
} mrf_dims;

/* Grid (crates/mr-core/src/lib.rs:394-408).  A zero length means the empty
 * Vec (no mask).  Lengths are checked like Grid::is_well_formed (:508-512). */
typedef struct {
    const uint32_t *cost;
    size_t n_cost;
    const uint8_t *via_forbidden;
    size_t n_via_forbidden;
    const uint8_t *board_constraint;
    size_t n_board_constraint;
} mrf_grid;

/* NetEndpoints (crates/mr-core/src/lib.rs:520-537) without the name. */
typedef struct {
    uint32_t src, dst;
    const uint32_t *pads;
    size_t n_pads;
    const uint32_t *via_pads;
    size_t n_via_pads;
} mrf_net;

/* IsolatedRouteWindow / MetalWindow, inclusive (negotiated.rs:495-519). */
typedef struct {
    uint32_t x0, y0, x1, y1;
} mrf_window;

/* Edge prices of IsolatedRouteRequest (negotiated.rs:530-537): x[w-1], y[h-1],
 * via[layers-1] with via_allowed[k] == 0 meaning None. */
typedef struct {
    const uint32_t *x;
    size_t nx;
    const uint32_t *y;
    size_t ny;
    const uint32_t *via;
    const uint8_t *via_allowed;
    size_t nvia;
} mrf_edges;

/* Negotiation pricing (negotiated.rs:2777-2797): entering v adds
 * additive[v] + halo_penalty * [halo_owner[v] is neither -1 nor the net's id].
 * All fields may be empty for the alone-path price. */
typedef struct {
    const uint32_t *additive;
    size_t n_additive;
    const int32_t *halo_owner;
    size_t n_halo_owner;
    uint32_t halo_penalty;
    const int32_t *net_ids; /* per net own id, or NULL (all -1) */
} mrf_pricing;

typedef struct {
    int code;
    uint32_t net; /* offending net index for MRF_INVALID_ENDPOINT */
    char message[256];
} mrf_error;

/* Library-owned path batch; release with mrf_free_paths. */
typedef struct {
    size_t n;
    uint8_t *routed;       /* n: 0 = None (unreachable) */
    uint32_t *search_cost; /* n: fixed-point search cost (MetalIsolatedRoute::search_cost) */
    uint64_t *offsets;     /* n+1: path i is cells[offsets[i] .. offsets[i+1]) */
    uint32_t *cells;       /* src ..= dst */
} mrf_paths;

/* Solve statistics of the last call on a context. */
typedef struct {
    double relax_ms;       /* device time of the relaxation (HIP events) */
    double path_ms;        /* device time of the path reconstruction */
    double total_ms;       /* host wall time of the call */
    uint64_t tile_visits;  /* (field, tile) loads from HBM */
    uint64_t local_rounds; /* in-LDS relaxation rounds */
    uint64_t cells_visited;
    uint64_t field_cells;
    uint32_t host_rounds;  /* launches of the round kernel (+1 for the persistent kernel) */
    double persistent_ms;  /* persistent form: the work-queue kernel alone */
  double relax_kernel_ms;  /* relaxation kernels alone: sum of per-launch HIP event times (plain
                              form in calls timed by mrf_set_round_timing, else the relaxation
                              phase; the graph form reports its replay time) */
  uint32_t relax_launches; /* relaxation kernel launches */
  uint64_t changed_visits; /* tile visits that improved at least one label */
  double validate_ms;      /* host re-validation of the device paths */
  uint32_t wave_launches;  /* relaxation launches that ran one wavefront per tile (tile width 16,
                              64 threads; lean unit-weight narrow batches in the plain form) */
  uint32_t label_bytes;    /* label width of the last solve: 8, or 4 for packed labels (wave form on
                              boards whose every open step costs at least 1, certified per batch) */
  uint32_t timed_launches; /* plain-form launches whose event times relax_kernel_ms sums (all of
                              them in a timed call, none otherwise: see mrf_set_round_timing) */
} mrf_stats;

/* Solve form of a context.
 *   PLAIN       one launch per relaxation round; the host reads the number of
 *               activated tiles after each round.
 *   PERSISTENT  resident grid with a device barrier between rounds: disabled in
 *               this library, mrf_solve* return MRF_BACKEND_UNAVAILABLE for it.
 *   GRAPH       grid-stride launches over a device work list, rounds captured
 *               in a hipGraph and replayed; one host read per replay.
 *   QUEUE       one persistent launch: workgroups pop tiles from device work
 *               queues (one per XCD, fields balanced over them by tiles) and
 *               push the neighbours their visits improve; no rounds, no
 *               barrier between workgroups.  Every wait inside the launch is
 *               bounded (MRF_QUEUE_WAIT_MS, read when a context opens,
 *               default 2000); a launch that exceeds a bound fails the call.
 *               The persistent work-queue batching form of
 *               crates/mr-metal/src/gpu.rs:1630-2039 (ragged batches), on
 *               the device instead of a host round loop.
 * A new context starts in PLAIN.  Results are the same in every form. */
enum { MRF_MODE_PLAIN = 0, MRF_MODE_PERSISTENT = 1, MRF_MODE_GRAPH = 2, MRF_MODE_QUEUE = 3 };

int mrf_abi_version(void);

/* Context: one HIP stream on one device, reusable device workspace. */
int mrf_open(int device, mrf_ctx **out, mrf_error *err);
void mrf_close(mrf_ctx *ctx);
/* Tile shape (tw x th cells, all layers), workgroup size (64..256), in-LDS
 * round cap.  Instantiated shapes: 32x32, 32x16, 16x16, 16x8, 8x4, 4x1.
 * 16 x 16 at 64 threads (a new context's shape, round cap 32) is the wave
 * form: one wavefront per tile visit for plain-form launches of fields with
 * no halo-owner or legal planes, overrides, own pads or via-exempt pads, on
 * boards without a weight plane, up to 4 layers, with every step price below
 * 0x7FFF and no delta stepping; every other launch of that shape runs the
 * workgroup kernels at 256 threads.  Results are the same either way. */
int mrf_set_tiles(mrf_ctx *ctx, uint32_t tw, uint32_t th, uint32_t threads, uint32_t max_local,
                  mrf_error *err);
int mrf_set_mode(mrf_ctx *ctx, int mode);
/* Plain form: process, each round, only the tiles whose activating border
 * cost is within `delta` of the smallest pending one; the others wait for a
 * later round (delta stepping over tiles: fewer revisits, more rounds).
 * UINT32_MAX (the default) processes every active tile each round.  Results
 * do not depend on it. */
int mrf_set_delta(mrf_ctx *ctx, uint32_t delta);
int mrf_last_stats(mrf_ctx *ctx, mrf_stats *out);
/* Measurement control, no counterpart in the reference: which solving calls
 * bracket each plain-form relaxation launch with a pair of HIP events (about
 * 3 us of host turnaround per round).  0 (the default): none; 1: every call;
 * k: the first of every k calls from now on.  An untimed call reports its
 * relaxation phase as relax_kernel_ms and timed_launches 0.  Results do not
 * depend on it. */
int mrf_set_round_timing(mrf_ctx *ctx, uint32_t every);

/* Label memory (bytes) one device batch may hold; a call whose fields need
 * more is split into batches that fit (mr-metal's MAX_BATCH_CELLS /
 * fields_per_batch chunking, crates/mr-metal/src/lib.rs:251-296, :626-678).
 * 0 restores the default, min(48 GiB, half of device memory); larger values
 * are clamped to it.  A single field larger than the budget still runs
 * alone in its own batch. */
int mrf_set_label_budget(mrf_ctx *ctx, size_t bytes);
size_t mrf_get_label_budget(mrf_ctx *ctx);

/* BoardRoute::congestion_from (crates/mr-core/src/lib.rs:575-593), the
 * congestion plane MetalRouter::route returns beside its paths
 * (crates/mr-metal/src/lib.rs:1227): out[c] (dims.len() u32) = number of
 * routed paths holding cell c, each path counting a cell once however often
 * it visits it.  Host only. */
int mrf_congestion_from(mrf_dims dims, const mrf_paths *paths, uint32_t *out, mrf_error *err);

/* IsolatedRouteProvider::route_isolated_batch + MetalIsolatedProvider adapter
 * (negotiated.rs:554-560; mr-cli lib.rs:549-625): solve each net inside its
 * window, then only the unreachable nets on the full board. */
int mrf_route_isolated_batch(mrf_ctx *ctx, mrf_dims dims, const mrf_grid *grid, const mrf_net *nets,
                             size_t n_nets, const mrf_window *windows, const mrf_edges *edges,
                             mrf_paths *out, mrf_error *err);

/* metal_route_isolated_batch (mr-metal lib.rs:414-558): one windowed pass;
 * routed[i] == 0 means unreachable inside windows[i]. */
int mrf_route_isolated_pass(mrf_ctx *ctx, mrf_dims dims, const mrf_grid *grid, const mrf_net *nets,
                            size_t n_nets, const mrf_window *windows, const mrf_edges *edges,
                            mrf_paths *out, mrf_error *err);

/* metal_sweep_fields (mr-metal lib.rs:172-192): planar single-source fields
 * with additive enter cost (bfs_distance_field, mr-cpu sweep.rs:143-148);
 * dist_out holds n_sources * dims.len() u32, in source order. */
int mrf_sweep_fields(mrf_ctx *ctx, mrf_dims dims, const mrf_grid *grid, const uint32_t *sources,
                     size_t n_sources, uint32_t *dist_out, mrf_error *err);

/* MetalRouter::route / LeeRouter::route (mr-metal lib.rs:1143-1235,
 * mr-cpu lee.rs:68-110): per-net canonical paths with cost = sum of enter costs
 * excluding the source.  routed[i] == 0 lists the net in BoardRoute::unrouted. */
int mrf_router_route(mrf_ctx *ctx, mrf_dims dims, const mrf_grid *grid, const mrf_net *nets,
                     size_t n_nets, mrf_paths *out, mrf_error *err);

/* Complete canonical fields: field i is seeded by sources[i][0..n_sources[i])
 * at label (0, 0) inside windows[i] (NULL = full board).  Sources and dsts[i]
 * (NULL or UINT32_MAX = none) stay passable like the alone search's endpoints.
 * dist_out (and hops_out unless NULL) receive n_fields * dims.len() u32 each. */
int mrf_fields(mrf_ctx *ctx, mrf_dims dims, const mrf_grid *grid, const mrf_edges *edges,
               const mrf_pricing *pricing, const uint32_t *const *sources, const size_t *n_sources,
               const uint32_t *dsts, size_t n_fields, const mrf_window *windows, uint32_t *dist_out,
               uint32_t *hops_out, mrf_error *err);

void mrf_free_paths(mrf_paths *paths);

/* ---- device-resident boards (serving and repeated solves) ---------------- */
int mrf_board_create(mrf_ctx *ctx, mrf_dims dims, const mrf_grid *grid, const mrf_edges *edges,
                     const mrf_pricing *pricing, mrf_board **out, mrf_error *err);
void mrf_board_destroy(mrf_board *board);
size_t mrf_board_bytes(const mrf_board *board);
/* Solve nets on a resident board: one pass in `windows` (NULL = full board),
 * then, when retry_full != 0, the unreachable nets on the full board.
 * Multi-pin nets: pass pins as sources via mrf_solve_fields. */
int mrf_solve(mrf_ctx *ctx, mrf_board *board, const mrf_net *nets, size_t n_nets,
              const mrf_window *windows, int retry_full, mrf_paths *out, mrf_error *err);
/* Allocation-free form of mrf_solve.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory during a call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The call makes no hipMalloc, hipFree, hipHostMalloc or hipEventCreate:
 * every device buffer of the solve (labels, tile states, work lists,
 * descriptors, path buffers) is carved from the library's fixed scratch
 * array on the context's device, mrf_noalloc_scratch_bytes() bytes (64 MiB),
 * which each device holds until the process ends.  Nets whose labels exceed
 * 5/8 of the array run in several device batches; a call that still needs
 * more (one field too large, or too many path cells at once) fails with
 * MRF_BACKEND_UNAVAILABLE and a message naming the limit.  It runs the plain
 * form whatever the context's mode; results equal mrf_solve's.  Paths go to
 * caller arrays: routed[n_nets], search_cost[n_nets], offsets[n_nets + 1],
 * cells[cells_cap]; *n_cells receives the number of path cells (also when
 * cells_cap is too small, which fails the call).  Calls that overlap on one
 * device, from any context, share the array and race: serialise them. */
int mrf_solve_noalloc(mrf_ctx *ctx, mrf_board *board, const mrf_net *nets, size_t n_nets,
                      const mrf_window *windows, int retry_full, uint8_t *routed, uint32_t *search_cost,
                      uint64_t *offsets, uint32_t *cells, size_t cells_cap, size_t *n_cells, mrf_error *err);
size_t mrf_noalloc_scratch_bytes(void);
/* Fields on a resident board; labels stay on the device until the next call,
 * read them with mrf_read_field. */
int mrf_solve_fields(mrf_ctx *ctx, mrf_board *board, const uint32_t *const *sources,
                     const size_t *n_sources, const uint32_t *dsts, const int32_t *net_ids,
                     size_t n_fields, const mrf_window *windows, mrf_error *err);
int mrf_read_field(mrf_ctx *ctx, mrf_board *board, size_t field, uint32_t *dist_out,
                   uint32_t *hops_out, mrf_error *err);

/* Replace the resident board's additive price plane (n == w*h*layers), or
 * drop it (additive == NULL, n == 0).  The negotiation loop reprices the
 * board once per iteration (negotiated.rs:1610-1667);
 * mrf_route_negotiated_batch calls this itself. */
int mrf_board_set_additive(mrf_board *board, const uint32_t *additive, size_t n, mrf_error *err);

/* ---- negotiated-congestion batches --------------------------------------------
 * One negotiation iteration's searches in one batch: the rayon map over the
 * dirty nets of the Jacobi loop (negotiated.rs:1679-1778), each net routed by
 * route_negotiated (negotiated.rs:2746-2856) in its window with a full-board
 * retry.  Entering cell v over base step price E*w costs
 *   passable_search_cost(E*w + history[v] + foreign(v))
 * with pf = pfac * 16 (SCALE == CLEARANCE_NEG_WEIGHT, negotiated.rs:71, :99),
 * self_present(v) = [v on the net's own path] (own_path_off != NULL) and
 * self_halo(v) = multiplicity of v in the net's own halo list:
 *   fused (shared_congestion, negotiated.rs:1654-1667):
 *     foreign = (pf*present[v] + pf*present_halo[v]) - pf*self_present - pf*self_halo
 *   otherwise:
 *     foreign = pf*(present[v] - self_present) + pf*(present_halo[v] - self_halo)
 * every sum and difference saturating (negotiated.rs:2781-2795).  The library
 * writes history + the shared term into the board's additive plane and prices
 * each net's own footprint apart on the device; the board must have no halo-
 * owner plane.  Paths and costs as mrf_solve. */
typedef struct {
    const uint32_t *history;      /* n_cells */
    const uint32_t *present;      /* n_cells */
    const uint32_t *present_halo; /* n_cells, or NULL for all zero (no clearance) */
    size_t n_cells;
    uint32_t pfac;                /* 1 + iteration */
    int fused;                    /* shared congestion snapshot (Some) or separate terms */
    /* Own copper of each net: own_path[own_path_off[i] .. own_path_off[i+1]),
     * or own_path_off == NULL when the batch passes no own path (None). */
    const uint64_t *own_path_off;
    const uint32_t *own_path;
    /* Own halo of each net (CountedCellSet): entries
     * [own_halo_off[i] .. own_halo_off[i+1]) of own_halo, each adding
     * own_halo_count[k] (1 when own_halo_count == NULL) to its cell's
     * multiplicity; own_halo_off == NULL for none. */
    const uint64_t *own_halo_off;
    const uint32_t *own_halo;
    const uint32_t *own_halo_count;
} mrf_negotiation;

int mrf_route_negotiated_batch(mrf_ctx *ctx, mrf_board *board, const mrf_net *nets, size_t n_nets,
                               const mrf_window *windows, int retry_full, const mrf_negotiation *neg,
                               mrf_paths *out, mrf_error *err);

/* The settle step of one negotiation iteration over the settled paths
 * (negotiated.rs:1779-1823 merge, :1924-1950 over-use and history), on the
 * host: present[c] = number of paths holding cell c (each path holds a cell
 * once); a cell is over-used when paths of two or more distinct groups hold
 * it; history[c] += scale (saturating at UINT32_MAX) on every over-used cell;
 * touches[i] = 1 when net i's path holds an over-used cell.  Path i is
 * path_cells[path_off[i] .. path_off[i + 1]); groups[i] is net i's connection
 * group.  The over-used cells are written ascending to `overused` (capacity
 * n_cells) and counted in *n_overused.  No device is used.  It restates the
 * part of NegotiatedRouter's loop body that follows each batch; the reference
 * has no separate entry point for it. */
int mrf_negotiation_settle(size_t n_cells, size_t n_nets, const uint64_t *path_off, const uint32_t *path_cells,
                           const int64_t *groups, uint32_t scale, uint32_t *present, uint32_t *history,
                           uint8_t *touches, uint32_t *overused, size_t *n_overused, mrf_error *err);

/* ---- legalization searches ------------------------------------------------------
 * route_legal_once (negotiated.rs:2885-3052) for a batch of nets against one
 * committed snapshot, as the staged legalization's parallel phase runs them
 * (negotiated.rs:3640-3700): each net searches its window (full-board retry
 * when retry_full) with copper owned by another group blocked even at its
 * endpoints, and a foreign or mixed (-2) clearance halo blocked except at its
 * endpoints and on its group's own copper (negotiated.rs:2937-2970); steps
 * cost E*w as in the alone search.  owner / halo: w*h*layers group ids (-1
 * free), either may be NULL; values must fit 32 bits.  The via guards are not
 * modelled: this is route_legal_once with an unoccupied ViaGuard, clearance 0
 * and no per-net via-forbidden XYs.  The board must have no additive or halo-
 * owner pricing plane (mrf_board_set_additive(board, NULL, 0) drops one). */
int mrf_route_legal_batch(mrf_ctx *ctx, mrf_board *board, const mrf_net *nets, size_t n_nets,
                          const mrf_window *windows, int retry_full, const int64_t *owner,
                          const int64_t *halo, size_t n_cells, const int64_t *own_groups,
                          mrf_paths *out, mrf_error *err);

/* ---- autotuner -----------------------------------------------------------------
 * Picks the tile shape, workgroup size, in-LDS round cap (sweeps per tile
 * visit) and solve form (plain rounds, hipGraph replays of `graph_rounds`
 * rounds: the work-list granularity, or the persistent work-queue launch)
 * for a workload by measuring candidates
 * on a sample of its nets on a resident board.  No reference counterpart: the
 * Metal engine uses fixed threadgroup sizes (crates/mr-metal/src/gpu.rs). */
typedef struct {
    uint32_t tile_w, tile_h, threads, max_local;
    int mode;               /* MRF_MODE_PLAIN, MRF_MODE_GRAPH or MRF_MODE_QUEUE */
    uint32_t graph_rounds;  /* rounds per hipGraph replay (graph form) */
    uint32_t delta;         /* delta-stepping window (plain form; UINT32_MAX = off) */
    double sample_ms;       /* wall time of the chosen configuration on the sample */
    double default_ms;      /* wall time of the context's configuration before tuning */
    uint32_t tried;         /* configurations measured */
} mrf_tuning;

/* Measures candidates on every k-th net (at most `sample` nets, windows as in
 * mrf_solve with the full-board retry), then times the three fastest and the
 * context's own configuration again on every k-th net of a larger set (up to
 * max(8 x sample, 1024) nets), keeps the fastest there on the context and
 * reports it.  sample_ms is the kept configuration's time on the sample.  The
 * context's results are the same for every configuration; only the time
 * differs. */
int mrf_autotune(mrf_ctx *ctx, mrf_board *board, const mrf_net *nets, size_t n_nets, const mrf_window *windows,
                 size_t sample, mrf_tuning *out, mrf_error *err);

/* ---- serving loop ------------------------------------------------------------
 * A server keeps board planes resident, queues solve requests from any thread
 * and runs one relaxation on a worker thread.  Requests are admitted in
 * arrival order at round boundaries while the label budget (HBM) and the
 * request slots allow; the fields of every admitted request, on any of the
 * resident boards, relax together, and a request leaves as soon as its own
 * fields have converged (continuous batching).  A request's paths do not
 * depend on what it was batched with: each field is the unique fixpoint of its
 * own relaxation.  The reference's /solve path runs one request per call
 * (crates/mr-server/src/lib.rs:362-453); this is the resident, batched form of
 * that loop, with the provider's semantics per request (window pass, then
 * full-board retry of unreachable nets when asked). */
typedef struct mrf_server mrf_server;

typedef struct {
    size_t label_bytes;     /* HBM for labels of admitted requests (the admission budget);
                               0 = a third of device memory, at most 96 GiB */
    uint32_t max_boards;    /* resident board slots (0 = 64) */
    uint32_t max_layers;    /* largest layer count of any board (0 = 8) */
    uint32_t max_inflight;  /* requests on the device at once (0 = 256) */
    uint32_t tile_w, tile_h, threads; /* tile shape (0 = 16 x 16, 64 threads: the wave form) */
    /* Admission pacing: a queued request is admitted only while the running
     * requests keep fewer tiles active than this (the device is not full),
     * so a large request runs nearly alone and later ones fill its tail.
     * 0 = six workgroups per CU; UINT32_MAX = admit whenever memory allows. */
    uint32_t admit_below_tiles;
    /* Size-aware admission: a queued request whose pass covers at most this
     * many tiles is admitted even while the device is paced, and ahead of
     * earlier requests that wait for memory or pacing, so small boards relax
     * beside a large one instead of queueing behind it.  0 = 16 x the pacing
     * threshold; 1 = no request goes ahead (arrival order only). */
    uint32_t admit_small_tiles;
} mrf_server_config;

typedef struct {
    double queue_ms;    /* submit -> admitted */
    double solve_ms;    /* admitted -> paths ready */
    double total_ms;    /* submit -> paths ready */
    uint32_t rounds;    /* relaxation rounds the request took part in */
    uint32_t peers_max; /* most other requests relaxing in the same round */
} mrf_request_stats;

typedef struct {
    uint64_t submitted, completed, failed;
    uint64_t rounds, tile_visits; /* tile_visits: up to the last request completion */
    double kernel_ms;         /* 0: the serving loop brackets no round with HIP events (an event
                                 pair costs host turnaround every round) */
    size_t label_bytes;       /* admission budget */
    size_t label_bytes_peak;  /* most label bytes admitted at once */
    uint32_t inflight_peak;   /* most requests admitted at once */
    uint64_t admitted_ahead;  /* requests admitted ahead of an earlier arrival */
} mrf_server_stats;

int mrf_server_create(int device, const mrf_server_config *cfg, mrf_server **out, mrf_error *err);
/* Stops the worker and releases the device; unfinished requests are dropped.
 * Call it only after every mrf_server_wait has returned. */
void mrf_server_destroy(mrf_server *srv);
int mrf_server_add_board(mrf_server *srv, mrf_dims dims, const mrf_grid *grid, const mrf_edges *edges,
                         const mrf_pricing *pricing, uint32_t *board_id, mrf_error *err);
/* MRF_BACKEND_UNAVAILABLE while a queued or running request uses the board. */
int mrf_server_remove_board(mrf_server *srv, uint32_t board_id, mrf_error *err);
/* Queue nets on a resident board: one pass inside `windows` (NULL = full
 * board), then, when retry_full != 0 and windows were given, the unreachable
 * nets on the full board.  net_ids (or NULL) are the halo owner ids. */
int mrf_server_submit(mrf_server *srv, uint32_t board_id, const mrf_net *nets, size_t n_nets,
                      const mrf_window *windows, int retry_full, const int32_t *net_ids,
                      uint64_t *request_id, mrf_error *err);
/* Wait up to timeout_ms (< 0: no limit) for a request.  MRF_NOT_READY when
 * it has not finished (the request stays queued); otherwise the request is
 * retired and its paths or error returned. */
int mrf_server_wait(mrf_server *srv, uint64_t request_id, int64_t timeout_ms, mrf_paths *out,
                    mrf_request_stats *stats, mrf_error *err);
int mrf_server_get_stats(mrf_server *srv, mrf_server_stats *out);

/* ---- synthetic workloads --------------------------------------------------- */
/* Seeded generators for the benchmark configurations (documented in
 * metal-route_amd/csrc/synth.cpp).  Output arrays are library-owned; free with
 * mrf_free_synth. */
typedef struct {
    mrf_dims dims;
    uint32_t *cost;
    uint8_t *board_constraint; /* NULL when the config has no outline */
    uint32_t *additive;        /* NULL when the config has no history plane */
    int32_t *halo_owner;       /* NULL unless the config prices foreign-pin halos */
    uint32_t halo_penalty;
    uint32_t *x_edges, *y_edges, *via_edges;
    uint8_t *via_allowed;
    size_t n_nets;
    uint32_t *src, *dst;       /* two-pin nets (dst = UINT32_MAX for multi-pin) */
    uint64_t *pin_off;         /* multi-pin nets: pins[pin_off[i]..pin_off[i+1]) */
    uint32_t *pins;
    mrf_window *windows;       /* per-net search windows, or NULL (full board) */
} mrf_synth;

int mrf_synth_config(int config, uint64_t seed, uint32_t scale_nets, mrf_synth *out, mrf_error *err);
/* The config-2 board generator at a given side, layer count and net count
 * (serving requests). */
int mrf_synth_board(uint32_t side, uint32_t layers, uint32_t nets, uint64_t seed, mrf_synth *out);
void mrf_free_synth(mrf_synth *s);

/* ---- SimpleRouteJson boards ---------------------------------------------------
 * A tscircuit SimpleRouteJson problem rasterised into the engine's inputs: the
 * Hanan grid, its obstacle / via / board-outline planes and the chained
 * two-point nets with their own-pad cells (RasterizedProblem,
 * crates/mr-srj/src/lib.rs:1303-1320).  Host code in
 * metal-route_amd/csrc/srj.cpp; a board is built once, then solved many times
 * (`cost`, `via_forbidden`, `board_constraint` go into mrf_grid as they are). */
enum {
    /* The CLI's rule selection (route_problem_impl, crates/mr-cli/src/lib.rs:751-834):
     * the coherent typed physical rules when uniform_physical_rules resolves,
     * else rasterize_with_layers with ceil(minClearance / resolution) cells,
     * the minClearance rule (default 0.15 mm) and a 0.45 mm via pad. */
    MRF_SRJ_ROUTE = 0,
    /* rasterize_with_layers (crates/mr-srj/src/lib.rs:1364-1384) with the
     * clearance_cells / min_clearance_mm / via_pad_mm given here; all zero is
     * rasterize() (:1323-1334). */
    MRF_SRJ_LAYERS = 1,
    /* rasterize_with_uniform_physical_rules (:1400-1426); MRF_BACKEND_UNAVAILABLE
     * when the problem's typed rules do not resolve. */
    MRF_SRJ_PHYSICAL = 2
};

/* Resolved uniform physical rules (UniformPhysicalRules, crates/mr-srj/src/lib.rs:252-265);
 * a negative optional value means None. */
typedef struct mrf_srj_rules {
    double trace_width_mm, obstacle_margin_mm, trace_to_pad_clearance_mm, via_to_pad_clearance_mm;
    double pad_to_pad_clearance_mm; /* < 0: None */
    double via_pad_diameter_mm, via_hole_diameter_mm;
    double via_hole_to_hole_clearance_mm; /* < 0: None */
} mrf_srj_rules;

typedef struct {
    int mode;                 /* MRF_SRJ_* */
    double resolution;        /* <= 0: longer bounds span / 64 (default_resolution,
                                 crates/mr-cli/src/lib.rs:462-473) */
    uint32_t layers;          /* 0: the problem's layerCount */
    uint32_t clearance_cells; /* MRF_SRJ_LAYERS only */
    double min_clearance_mm;  /* MRF_SRJ_LAYERS only */
    double via_pad_mm;        /* MRF_SRJ_LAYERS only */
    uint64_t cell_budget;     /* planar cell ceiling of the fill lanes: 0 = the
                                 MR_CELL_BUDGET rule (crates/mr-srj/src/lib.rs:728-792,
                                 160,000 when unset), UINT64_MAX = no ceiling */
    int model;                /* MRF_SRJ_LAYERS: 0 = select_clearance_raster_model
                                 (lib.rs:689-726), 1 = legacy rounded, 2 = exact physical */
    const struct mrf_srj_rules *rules; /* MRF_SRJ_PHYSICAL: the rules to project, as
                                 rasterize_with_uniform_physical_rules takes them;
                                 NULL = resolve the problem's own */
} mrf_srj_params;

/* NetEndpoints (crates/mr-core/src/lib.rs:520-537) with its name. */
typedef struct {
    const char *name; /* conn | conn#k | root#conn | root#conn#k (lib.rs:2138-2155) */
    uint32_t src, dst;
    uint32_t *pads; /* passable_pads, ascending */
    size_t n_pads;
    uint32_t *via_pads; /* via_passable_pads, ascending, a subset of pads */
    size_t n_via_pads;
} mrf_srj_net;


typedef struct {
    mrf_dims dims;
    double *x_lines; /* dims.w line positions (Mapping::x_lines) */
    double *y_lines; /* dims.h */
    uint32_t *cost;  /* dims cells: 1, or UINT32_MAX (OBSTACLE) */
    uint8_t *via_forbidden;    /* NULL: the empty mask */
    uint8_t *board_constraint; /* NULL: the empty mask */
    size_t n_nets;
    mrf_srj_net *nets;
    size_t n_pins;
    uint32_t *pin_cells; /* pin_points keys, ascending */
    double *pin_xy;      /* 2 * n_pins: the exact (x, y) of each */
    /* The rules the raster was built with. */
    double resolution, trace_width_mm, min_clearance_mm, via_pad_mm;
    uint32_t clearance_cells;
    int exact_model; /* ClearanceRasterModel::ExactPhysical (lib.rs:673-726) */
    int physical;    /* 1 when the typed uniform rules were applied */
    mrf_srj_rules rules;
    uint32_t fill_dropped; /* fill lanes the cell budget removed */
} mrf_srj_problem;

/* Parse and rasterise one problem.  MRF_BACKEND_UNAVAILABLE with a message
 * for malformed JSON, a missing required field or a bad resolution. */
int mrf_srj_rasterize(const char *json, size_t len, const mrf_srj_params *params, mrf_srj_problem **out,
                      mrf_error *err);
void mrf_srj_free(mrf_srj_problem *p);
/* SimpleRouteJson::uniform_physical_rules (crates/mr-srj/src/lib.rs:286-528):
 * *ok = 1 and the rules when the typed profile is coherent, else *ok = 0. */
int mrf_srj_uniform_rules(const char *json, size_t len, mrf_srj_rules *out, int *ok, mrf_error *err);
/* The board-outline projection alone (BoardOutlineConstraint::from_srj and
 * raster_mask_plane, crates/mr-srj/src/board_outline.rs:72-108, :223-229) over
 * given sorted line arrays: *state = 0 when the problem declares no outline
 * contract (out untouched), 1 when the nx*ny plane (row-major, y outer) is
 * written, -1 when the contract is malformed (every plane bit set). */
int mrf_srj_outline_plane(const char *json, size_t len, double trace_width_mm, double via_pad_mm,
                          const double *x_lines, size_t nx, const double *y_lines, size_t ny, uint8_t *out,
                          int *state, mrf_error *err);
/* The fill-lane budget step alone (enforce_budget, crates/mr-srj/src/lib.rs:1047-1128)
 * over sorted feature lines and fill candidates; the kept fill lanes are written
 * back in place and their counts returned in *n_x_fill / *n_y_fill. */
int mrf_srj_enforce_budget(const double *x_features, size_t nxf, const double *y_features, size_t nyf,
                           double *x_fill, size_t *n_x_fill, double *y_fill, size_t *n_y_fill, uint64_t budget);

#ifdef __cplusplus
}
#endif
#endif
