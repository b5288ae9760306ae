/*
 * mrcpu -- CPU ORACLE for the batched shortest-path field engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * router's CPU search (crate mr-cpu) and of the mr-core grid predicates.  It is
 * the reference the HIP engine is compared against, and the CPU baseline that
 * bench.py times beside the GPU.  Nothing in the product library links, loads
 * or calls it; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.
 *
 * Every function cites the reference file:line it restates.  Parity pins for
 * each function are listed in oracle/mrcpu.c's header.
 */
#ifndef MRCPU_H
#define MRCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRC_MAX UINT32_MAX          /* mr-core OBSTACLE / unreachable sentinel */
#define MRC_NO_PRED UINT32_MAX      /* dijkstra.rs:21 NO_PRED */

/* mr-core/src/lib.rs:411-416 Grid::BOARD_* bits */
#define MRC_BOARD_TRACE_NODE 0x01u
#define MRC_BOARD_VIA_NODE 0x02u
#define MRC_BOARD_EDGE_NEG_Y 0x04u
#define MRC_BOARD_EDGE_NEG_X 0x08u
#define MRC_BOARD_EDGE_POS_X 0x10u
#define MRC_BOARD_EDGE_POS_Y 0x20u

typedef struct {
    uint32_t w, h, layers;
} mrc_dims;

/* mr-core/src/lib.rs:394-408 Grid.  via_forbidden / board may be NULL, which is
 * the reference's empty-Vec "no mask" representation. */
typedef struct {
    mrc_dims dims;
    const uint32_t *cost;
    const uint8_t *via_forbidden;
    const uint8_t *board;
} mrc_grid;

/* negotiated.rs:495-519 IsolatedRouteWindow (inclusive planar rectangle). */
typedef struct {
    uint32_t x0, y0, x1, y1;
} mrc_window;

/* Already-rounded edge prices (negotiated.rs:530-537 IsolatedRouteRequest):
 * x[w-1], y[h-1], via[layers-1] with via_allowed[k]==0 meaning None. */
typedef struct {
    const uint32_t *x;
    const uint32_t *y;
    const uint32_t *via;
    const uint8_t *via_allowed;
} mrc_edges;

/* Search semantics selected for astar / complete fields. */
enum {
    /* route_legal_once alone path (negotiated.rs:2885-3052) and, with an
     * additive plane, route_negotiated pricing (negotiated.rs:2746-2856). */
    MRC_MODE_ALONE = 0,
    /* LeeRouter::route_one_with_buf (lee.rs:39-65): planar, enter cost, no cap. */
    MRC_MODE_LEE = 1
};

/* One net's search specification. */
typedef struct {
    const mrc_grid *grid;
    mrc_edges edges;
    /* Optional negotiation pricing (negotiated.rs:2777-2797): the price of
     * entering v is base + additive[v] + halo_penalty * [halo_owner[v] is a
     * foreign owner].  additive == NULL and halo_owner == NULL is the alone path. */
    const uint32_t *additive;
    const int32_t *halo_owner;  /* -1 free, -2 mixed, else owning net id */
    uint32_t halo_penalty;
    int32_t own_id;
    /* Own pads (sorted ascending, deduplicated) and via-exempt pads (sorted). */
    const uint32_t *pads;
    size_t n_pads;
    const uint32_t *via_pads;
    size_t n_via_pads;
    mrc_window window;
    int mode;
    /* astar heuristic: 0 = none (Dijkstra), 1 = manhattan_scaled. */
    int use_heuristic;
    uint32_t min_via_cost;
    /* route_negotiated snapshot pricing (negotiated.rs:2777-2797), active when
     * neg_history != NULL: entering v costs passable_search_cost(base +
     * history[v] + foreign(v)) with the net's own copper and counted halo
     * subtracted from the snapshot (fused when neg_shared != NULL). */
    const uint32_t *neg_history;
    const uint32_t *neg_present;
    const uint32_t *neg_halo;    /* NULL = all zero */
    const uint64_t *neg_shared;  /* fused shared congestion, or NULL */
    uint32_t neg_pfac;
    int has_own_path;            /* own_path Some / None */
    const uint32_t *own_path;    /* sorted unique */
    size_t n_own_path;
    const uint32_t *own_halo;    /* sorted unique cells ... */
    const uint32_t *own_halo_count; /* ... and their multiplicities */
    size_t n_own_halo;
    /* route_legal_once hard blocks (negotiated.rs:2937-2970): committed copper
     * owners and clearance halo (-1 free, -2 mixed, else a group), either may
     * be NULL; the searching net's group is own_group. */
    const int64_t *legal_owner;
    const int64_t *legal_halo;
    int64_t own_group;
} mrc_spec;

/* ---- scalar helpers ---------------------------------------------------- */
uint32_t mrc_edge_cost(double length);                                  /* dijkstra.rs:43-52 */
uint32_t mrc_axis_gap_cost(const double *lines, size_t n_lines, uint32_t k); /* negotiated.rs:1197-1201 */
void mrc_axis_edge_costs(const double *lines, size_t n_lines, uint32_t count,
                         uint32_t *out);                                /* negotiated.rs:1268-1272 */
mrc_window mrc_net_window(mrc_dims d, uint32_t src, uint32_t dst, const uint32_t *pads,
                          size_t n_pads);                               /* negotiated.rs:568-594 */
mrc_window mrc_window_full(mrc_dims d);                                 /* negotiated.rs:505-512 */
uint32_t mrc_passable_search_cost(uint64_t cost);                       /* negotiated.rs:1171-1173 */
int mrc_is_step_legal_allowed(const mrc_edges *e, uint32_t layers, uint32_t a, uint32_t b);

/* ---- mr-core predicates (mr-core/src/lib.rs:434-496) --------------------- */
int mrc_is_obstacle(const mrc_grid *g, uint32_t c);
int mrc_is_via_forbidden(const mrc_grid *g, uint32_t c);
int mrc_is_board_forbidden(const mrc_grid *g, uint32_t c);
int mrc_is_board_via_forbidden(const mrc_grid *g, uint32_t c);
int mrc_is_board_planar_step_forbidden(const mrc_grid *g, uint32_t u, uint32_t v);

/* ---- complete planar fields (sweep.rs) ----------------------------------- */
/* bfs_distance_field == dijkstra(grid, src, 0).dist (sweep.rs:143-148, dijkstra.rs:70-115) */
void mrc_bfs_distance_field(const mrc_grid *g, uint32_t src, uint32_t *dist_out);
/* sweep_distance_field (sweep.rs:64-139) */
void mrc_sweep_distance_field(const mrc_grid *g, uint32_t src, uint32_t *dist_out);
/* path_from_field (sweep.rs:158-235). Returns path length, 0 = None. */
size_t mrc_path_from_field(const mrc_grid *g, const uint32_t *dist, uint32_t src, uint32_t dst,
                           uint32_t *path_out);

/* ---- targeted canonical search (dijkstra.rs:243-401) --------------------- */
typedef struct mrc_searchbuf mrc_searchbuf;
mrc_searchbuf *mrc_searchbuf_new(size_t n);
void mrc_searchbuf_free(mrc_searchbuf *b);
/* Returns path length (0 = None); path_out needs room for dims.len() cells. */
size_t mrc_astar(mrc_searchbuf *b, const mrc_spec *s, uint32_t src, uint32_t dst,
                 uint32_t *path_out, uint32_t *cost_out);

/* ---- complete canonical fields (lexicographic (dist, hops) Dijkstra) ------ */
/* Multi-source complete field under spec s: dist/hops per cell (MAX unreached).
 * Sources that are blocked under s are skipped (negotiated.rs:2946-2955 force
 * src/dst passable; `forced` lists the cells the search forces passable). */
void mrc_field_lex(const mrc_spec *s, const uint32_t *sources, size_t n_sources,
                   const uint32_t *forced, size_t n_forced, uint32_t *dist_out,
                   uint32_t *hops_out);
/* Canonical backward reconstruction: lowest-CellIdx predecessor p with
 * dist(p)+step(p,c)==dist(c) and hops(p)+1==hops(c).  Returns length, 0 = None. */
size_t mrc_reconstruct(const mrc_spec *s, const uint32_t *dist, const uint32_t *hops,
                       uint32_t dst, uint32_t *path_out);

/* ---- provider contract ---------------------------------------------------- */
/* provider_path_is_valid (negotiated.rs:1284-1331) with via legality from edges. */
int mrc_provider_path_is_valid(const mrc_grid *g, uint32_t src, uint32_t dst,
                               const uint32_t *pads, size_t n_pads, const uint32_t *via_pads,
                               size_t n_via_pads, const mrc_edges *e, const uint32_t *path,
                               size_t len);
/* grid_path_cost (negotiated.rs:1106-1115) */
uint32_t mrc_grid_path_cost(const mrc_grid *g, const uint32_t *pads, size_t n_pads,
                            const uint32_t *path, size_t len);
/* Step price of u -> v under s (cost_fn / via_step); MRC_MAX when illegal. */
uint32_t mrc_step_cost(const mrc_spec *s, uint32_t u, uint32_t v);

/* ---- batched alone search (negotiated.rs:2115-2184) ------------------------ */
typedef struct {
    uint32_t src, dst;
    const uint32_t *pads;
    size_t n_pads;
    const uint32_t *via_pads;
    size_t n_via_pads;
    int32_t own_id;
    /* negotiated searches: the net's own copper and counted halo */
    int has_own_path;
    const uint32_t *own_path;
    size_t n_own_path;
    const uint32_t *own_halo;
    const uint32_t *own_halo_count;
    size_t n_own_halo;
    int64_t own_group; /* legal searches: the net's group */
} mrc_net;

/* For each net: search its window, and on failure the full board.  Paths are
 * written into path_cells at path_off[i] (capacity per net = dims.len());
 * path_len[i] = 0 for unreachable; cost[i] = search cost.  n_threads workers
 * split the nets like rayon's into_par_iter (results are order-independent).
 * `template` supplies grid/edges/additive/mode/heuristic. */
void mrc_alone_batch(const mrc_spec *tmpl, const mrc_net *nets, size_t n_nets,
                     const mrc_window *windows, int retry_full, int n_threads,
                     uint32_t *path_cells, const uint64_t *path_off, uint32_t *path_len,
                     uint32_t *cost);

/* Complete multi-source fields for many nets on n_threads workers. */
void mrc_fields_batch(const mrc_spec *tmpl, const mrc_net *nets, size_t n_nets,
                      const uint32_t *const *sources, const size_t *n_sources,
                      const mrc_window *windows, int n_threads, uint32_t *dist_out,
                      uint32_t *hops_out);

/* mr-core/src/lib.rs:575-593 BoardRoute::congestion_from: out[c] = number of
 * paths that hold cell c, each path counting a cell once (a loopy path is one
 * route).  Path i is cells[off[i] .. off[i+1]). */
void mrc_congestion_from(mrc_dims d, size_t n_paths, const uint64_t *off, const uint32_t *cells,
                         uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
