/*
 * mrcpu.c -- CPU ORACLE (test infrastructure, never product code).
 *
 * Plain-C restatement of the reference CPU router's search core.  The product
 * engine (metal-route_amd/) never links or calls this file; tests/, smoke()
 * and bench.py's cpu_baseline leg load it as the reference the engine is
 * compared against and as the CPU baseline.
 *
 * Sources restated (paths relative to the reference root):
 *   crates/mr-core/src/lib.rs     Dims mapping, Grid predicates   :90-164, :434-512
 *   crates/mr-cpu/src/dijkstra.rs edge_cost, dijkstra, astar_buf  :43-52, :70-115, :243-401
 *   crates/mr-cpu/src/sweep.rs    sweep/bfs fields, path_from_field :48-235
 *   crates/mr-cpu/src/lee.rs      LeeRouter closures              :39-65
 *   crates/mr-cpu/src/negotiated.rs
 *       net_window :568-594, passable_search_cost :1171-1173,
 *       axis_gap_cost/axis_edge_costs :1197-1201/:1268-1272,
 *       manhattan_scaled/axis_cost_prefix/axis_leg_cost :1245-1282,
 *       provider_path_is_valid :1284-1331, grid_path_cost :1106-1115,
 *       CPU alone batch :2115-2184, route_negotiated :2746-2856
 *       (priced_with_base with self-subtraction :2777-2797),
 *       route_legal_once :2885-3052.
 *
 * Parity pins (what checks each function; tests/test_oracle*.py):
 *   mrc_edge_cost / axis costs   reference vectors [144,16,16], [16,16]
 *                                (negotiated.rs:4663-4705) + closed forms.
 *   mrc_net_window               closed form of the margin rule on hand cases.
 *   mrc_bfs/sweep_distance_field battery costs 93 / 8 / 8, asymmetric-mask
 *                                field values 3 / MAX (sweep.rs:351-370), and
 *                                sweep == bfs exhaustively on 3x3 masks.
 *   mrc_path_from_field          tie_break_2x2 [0,1,3], zero plateau [5,4,3].
 *   mrc_astar (LEE / ALONE)      every concrete path vector the reference tests
 *                                pin (mr-cpu lib.rs, astar.rs, mr-metal lib.rs,
 *                                negotiated.rs; see tests/golden/reference_vectors.json),
 *                                brute-force enumeration of simple paths on tiny
 *                                grids, and heuristic on == off.
 *   mrc_field_lex/reconstruct    equal to mrc_astar on every net of the seeded
 *                                stress batteries (two different algorithms:
 *                                label-correcting A* with early stop versus
 *                                label-setting lexicographic Dijkstra), closed
 *                                form on open grids, multi-source = min over
 *                                single-source fields.
 *   mrc_provider_path_is_valid   accepts every oracle path, rejects hand-built
 *                                illegal paths (negotiated.rs:4486-4517 cases).
 *   blocked (legal owner / halo) equal to the alone search on the board with
 *                                every blocked cell turned into a board
 *                                trace node (endpoint exemptions applied), on
 *                                random tiny boards; hand cases for the
 *                                endpoint / own-copper / mixed-halo ordering.
 *   priced_negotiated            the two properties negotiated.rs:6560-6660
 *                                asserts: a net whose own copper and halo are
 *                                the whole snapshot routes as on an empty
 *                                snapshot, and fused == term-by-term pricing
 *                                (randomized like :6665-6800); plus brute-force
 *                                simple-path enumeration under the restated
 *                                price on tiny grids.
 *   mrc_congestion_from          mr-core's vectors [1,2,1] and, with a loopy
 *                                path, [1,2,2,1] (mr-core/src/lib.rs:856-891),
 *                                and equal to a brute count of distinct cells
 *                                per path on random path sets.
 */
#include "mrcpu.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* scalar helpers                                                             */
/* ------------------------------------------------------------------------- */

static inline uint32_t sat_add32(uint32_t a, uint32_t b) {
    uint32_t r = a + b;
    return r < a ? MRC_MAX : r;
}

static inline uint32_t sat_sub32(uint32_t a, uint32_t b) { return a > b ? a - b : 0; }

static inline uint64_t sat_add64(uint64_t a, uint64_t b) {
    uint64_t r = a + b;
    return r < a ? UINT64_MAX : r;
}

static inline uint64_t sat_mul64(uint64_t a, uint64_t b) {
    if (a == 0 || b == 0) return 0;
    if (a > UINT64_MAX / b) return UINT64_MAX;
    return a * b;
}

/* dijkstra.rs:43-52: round(length * COST_SCALE), half away from zero,
 * saturating, non-finite / non-positive -> 0. */
uint32_t mrc_edge_cost(double length) {
    double scaled = round(length * 16.0);
    if (!isfinite(scaled) || scaled <= 0.0) return 0;
    if (scaled >= (double)MRC_MAX) return MRC_MAX;
    return (uint32_t)scaled;
}

/* negotiated.rs:1197-1201 with GridCoords' index-as-position fallback
 * (mr-core/src/lib.rs:235-246). */
uint32_t mrc_axis_gap_cost(const double *lines, size_t n_lines, uint32_t k) {
    double a = (size_t)k < n_lines ? lines[k] : (double)k;
    double b = (size_t)k + 1 < n_lines ? lines[k + 1] : (double)(k + 1);
    return mrc_edge_cost(fabs(b - a));
}

/* negotiated.rs:1268-1272 */
void mrc_axis_edge_costs(const double *lines, size_t n_lines, uint32_t count, uint32_t *out) {
    for (uint32_t k = 0; k + 1 < count; ++k) out[k] = mrc_axis_gap_cost(lines, n_lines, k);
}

/* negotiated.rs:568-594 */
mrc_window mrc_net_window(mrc_dims d, uint32_t src, uint32_t dst, const uint32_t *pads,
                          size_t n_pads) {
    uint32_t plane = d.w * d.h;
    uint32_t sr = src % plane, dr = dst % plane;
    uint32_t sx = sr % d.w, sy = sr / d.w, dx = dr % d.w, dy = dr / d.w;
    uint32_t x0 = sx < dx ? sx : dx, y0 = sy < dy ? sy : dy;
    uint32_t x1 = sx > dx ? sx : dx, y1 = sy > dy ? sy : dy;
    for (size_t i = 0; i < n_pads; ++i) {
        uint32_t r = pads[i] % plane, px = r % d.w, py = r / d.w;
        if (px < x0) x0 = px;
        if (py < y0) y0 = py;
        if (px > x1) x1 = px;
        if (py > y1) y1 = py;
    }
    uint32_t span = (x1 - x0) > (y1 - y0) ? (x1 - x0) : (y1 - y0);
    uint32_t m = (3 * span + 9) / 10; /* div_ceil(3*span, 10) */
    if (m < 16) m = 16;
    mrc_window win;
    win.x0 = x0 > m ? x0 - m : 0;
    win.y0 = y0 > m ? y0 - m : 0;
    win.x1 = x1 + m < d.w - 1 ? x1 + m : d.w - 1;
    win.y1 = y1 + m < d.h - 1 ? y1 + m : d.h - 1;
    return win;
}

mrc_window mrc_window_full(mrc_dims d) {
    mrc_window w = {0, 0, d.w ? d.w - 1 : 0, d.h ? d.h - 1 : 0};
    return w;
}

/* negotiated.rs:1171-1173 */
uint32_t mrc_passable_search_cost(uint64_t cost) {
    return cost < (uint64_t)(MRC_MAX - 1) ? (uint32_t)cost : MRC_MAX - 1;
}

/* mr-core/src/lib.rs:360-372 ViaModel::is_step_legal, with the model's allowed
 * set carried as the request's via_edge_costs Option (negotiated.rs:1423-1429). */
int mrc_is_step_legal_allowed(const mrc_edges *e, uint32_t layers, uint32_t a, uint32_t b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    if (hi - lo != 1 || hi >= layers) return 0;
    return e->via_allowed != NULL && e->via_allowed[lo] != 0;
}

/* ------------------------------------------------------------------------- */
/* mr-core predicates (mr-core/src/lib.rs:434-496)                            */
/* ------------------------------------------------------------------------- */

static inline size_t dims_len(mrc_dims d) { return (size_t)d.w * d.h * d.layers; }

int mrc_is_obstacle(const mrc_grid *g, uint32_t c) { return g->cost[c] == MRC_MAX; }

int mrc_is_via_forbidden(const mrc_grid *g, uint32_t c) {
    return g->via_forbidden != NULL && g->via_forbidden[c] != 0;
}

int mrc_is_board_forbidden(const mrc_grid *g, uint32_t c) {
    return g->board != NULL && (g->board[c] & MRC_BOARD_TRACE_NODE) != 0;
}

int mrc_is_board_via_forbidden(const mrc_grid *g, uint32_t c) {
    return g->board != NULL && (g->board[c] & MRC_BOARD_VIA_NODE) != 0;
}

int mrc_is_board_planar_step_forbidden(const mrc_grid *g, uint32_t u, uint32_t v) {
    if (g->board == NULL) return 0;
    mrc_dims d = g->dims;
    size_t n = dims_len(d);
    if ((size_t)u >= n || (size_t)v >= n) return 1;
    uint32_t plane = d.w * d.h;
    uint32_t ul = u / plane, ur = u % plane, ux = ur % d.w, uy = ur / d.w;
    uint32_t vl = v / plane, vr = v % plane, vx = vr % d.w, vy = vr / d.w;
    if (ul != vl) return 1;
    uint8_t sb, db;
    if (ux == vx && uy == vy + 1) {
        sb = MRC_BOARD_EDGE_NEG_Y;
        db = MRC_BOARD_EDGE_POS_Y;
    } else if (uy == vy && ux == vx + 1) {
        sb = MRC_BOARD_EDGE_NEG_X;
        db = MRC_BOARD_EDGE_POS_X;
    } else if (uy == vy && ux + 1 == vx) {
        sb = MRC_BOARD_EDGE_POS_X;
        db = MRC_BOARD_EDGE_NEG_X;
    } else if (ux == vx && uy + 1 == vy) {
        sb = MRC_BOARD_EDGE_POS_Y;
        db = MRC_BOARD_EDGE_NEG_Y;
    } else {
        return 1;
    }
    return (g->board[u] & sb) != 0 || (g->board[v] & db) != 0;
}

/* Ascending-CellIdx 4-neighbours (mr-core/src/lib.rs:129-146); returns count. */
static int neighbors4(mrc_dims d, uint32_t i, uint32_t out[4]) {
    uint32_t plane = d.w * d.h, r = i % plane, x = r % d.w, y = r / d.w;
    int n = 0;
    if (y > 0) out[n++] = i - d.w;
    if (x > 0) out[n++] = i - 1;
    if (x + 1 < d.w) out[n++] = i + 1;
    if (y + 1 < d.h) out[n++] = i + d.w;
    return n;
}

static int sorted_contains(const uint32_t *a, size_t n, uint32_t v) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && a[lo] == v;
}

/* ------------------------------------------------------------------------- */
/* generic binary min-heap of 64-bit keys                                     */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t *k;
    uint32_t *aux; /* secondary payload (hops) for the lexicographic heap */
    size_t n, cap;
} heap_t;

static void heap_init(heap_t *h, int with_aux) {
    h->cap = 1024;
    h->n = 0;
    h->k = (uint64_t *)malloc(h->cap * sizeof(uint64_t));
    h->aux = with_aux ? (uint32_t *)malloc(h->cap * sizeof(uint32_t)) : NULL;
}

static void heap_free(heap_t *h) {
    free(h->k);
    free(h->aux);
}

/* Ordering: key ascending, then aux ascending (aux only when present). */
static inline int heap_less(const heap_t *h, size_t a, size_t b) {
    if (h->k[a] != h->k[b]) return h->k[a] < h->k[b];
    if (h->aux) return h->aux[a] < h->aux[b];
    return 0;
}

static inline void heap_swap(heap_t *h, size_t a, size_t b) {
    uint64_t t = h->k[a];
    h->k[a] = h->k[b];
    h->k[b] = t;
    if (h->aux) {
        uint32_t u = h->aux[a];
        h->aux[a] = h->aux[b];
        h->aux[b] = u;
    }
}

static void heap_push(heap_t *h, uint64_t key, uint32_t aux) {
    if (h->n == h->cap) {
        h->cap *= 2;
        h->k = (uint64_t *)realloc(h->k, h->cap * sizeof(uint64_t));
        if (h->aux) h->aux = (uint32_t *)realloc(h->aux, h->cap * sizeof(uint32_t));
    }
    size_t i = h->n++;
    h->k[i] = key;
    if (h->aux) h->aux[i] = aux;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (!heap_less(h, i, p)) break;
        heap_swap(h, i, p);
        i = p;
    }
}

static int heap_pop(heap_t *h, uint64_t *key, uint32_t *aux) {
    if (h->n == 0) return 0;
    *key = h->k[0];
    if (aux && h->aux) *aux = h->aux[0];
    h->n--;
    if (h->n) {
        h->k[0] = h->k[h->n];
        if (h->aux) h->aux[0] = h->aux[h->n];
        size_t i = 0;
        for (;;) {
            size_t l = 2 * i + 1, r = l + 1, m = i;
            if (l < h->n && heap_less(h, l, m)) m = l;
            if (r < h->n && heap_less(h, r, m)) m = r;
            if (m == i) break;
            heap_swap(h, i, m);
            i = m;
        }
    }
    return 1;
}

/* ------------------------------------------------------------------------- */
/* complete planar fields: dijkstra / bfs / sweep / path_from_field            */
/* ------------------------------------------------------------------------- */

/* dijkstra.rs:70-115 with heuristic 0; only dist is needed by the callers. */
void mrc_bfs_distance_field(const mrc_grid *g, uint32_t src, uint32_t *dist) {
    size_t n = dims_len(g->dims);
    for (size_t i = 0; i < n; ++i) dist[i] = MRC_MAX;
    if ((size_t)src >= n) return; /* sweep.rs:144-146 */
    if (mrc_is_obstacle(g, src) || mrc_is_board_forbidden(g, src)) return;
    dist[src] = 0;
    heap_t h;
    heap_init(&h, 0);
    heap_push(&h, (uint64_t)src, 0); /* (prio 0, cell) */
    uint64_t key;
    while (heap_pop(&h, &key, NULL)) {
        uint32_t prio = (uint32_t)(key >> 32), u = (uint32_t)key;
        uint32_t du = dist[u];
        if (prio > du) continue;
        uint32_t nb[4];
        int k = neighbors4(g->dims, u, nb);
        for (int j = 0; j < k; ++j) {
            uint32_t v = nb[j];
            if (mrc_is_obstacle(g, v) || mrc_is_board_forbidden(g, v) ||
                mrc_is_board_planar_step_forbidden(g, u, v))
                continue;
            uint32_t nd = sat_add32(du, g->cost[v]);
            if (nd < dist[v]) {
                dist[v] = nd;
                heap_push(&h, ((uint64_t)nd << 32) | v, 0);
            }
        }
    }
    heap_free(&h);
}

/* sweep.rs:48-57 */
static uint32_t sweep_relax(const mrc_grid *g, uint32_t prev, uint32_t n, uint32_t prev_dist) {
    if (prev_dist == MRC_MAX || mrc_is_obstacle(g, n) || mrc_is_board_forbidden(g, n) ||
        mrc_is_board_planar_step_forbidden(g, prev, n))
        return MRC_MAX;
    return sat_add32(prev_dist, g->cost[n]);
}

/* sweep.rs:64-139 */
void mrc_sweep_distance_field(const mrc_grid *g, uint32_t src, uint32_t *dist) {
    mrc_dims d = g->dims;
    size_t n = dims_len(d);
    for (size_t i = 0; i < n; ++i) dist[i] = MRC_MAX;
    if (n == 0 || (size_t)src >= n || mrc_is_obstacle(g, src) || mrc_is_board_forbidden(g, src))
        return;
    dist[src] = 0;
    uint32_t layer = src / (d.w * d.h);
#define IDX3(x, y) (((layer)*d.h + (y)) * d.w + (x))
    for (;;) {
        int changed = 0;
        for (uint32_t y = 0; y < d.h; ++y) {
            for (uint32_t x = 1; x < d.w; ++x) {
                uint32_t cur = IDX3(x, y), prev = IDX3(x - 1, y);
                uint32_t c = sweep_relax(g, prev, cur, dist[prev]);
                if (c < dist[cur]) {
                    dist[cur] = c;
                    changed = 1;
                }
            }
            for (uint32_t xi = d.w; xi >= 2; --xi) {
                uint32_t x = xi - 2, cur = IDX3(x, y), prev = IDX3(x + 1, y);
                uint32_t c = sweep_relax(g, prev, cur, dist[prev]);
                if (c < dist[cur]) {
                    dist[cur] = c;
                    changed = 1;
                }
            }
        }
        for (uint32_t x = 0; x < d.w; ++x) {
            for (uint32_t y = 1; y < d.h; ++y) {
                uint32_t cur = IDX3(x, y), prev = IDX3(x, y - 1);
                uint32_t c = sweep_relax(g, prev, cur, dist[prev]);
                if (c < dist[cur]) {
                    dist[cur] = c;
                    changed = 1;
                }
            }
            for (uint32_t yi = d.h; yi >= 2; --yi) {
                uint32_t y = yi - 2, cur = IDX3(x, y), prev = IDX3(x, y + 1);
                uint32_t c = sweep_relax(g, prev, cur, dist[prev]);
                if (c < dist[cur]) {
                    dist[cur] = c;
                    changed = 1;
                }
            }
        }
        if (!changed) break;
    }
#undef IDX3
}

/* sweep.rs:158-235 */
size_t mrc_path_from_field(const mrc_grid *g, const uint32_t *dist, uint32_t src, uint32_t dst,
                           uint32_t *path) {
    mrc_dims d = g->dims;
    size_t n = dims_len(d);
    if ((size_t)src >= n || (size_t)dst >= n || mrc_is_obstacle(g, src) ||
        mrc_is_obstacle(g, dst) || mrc_is_board_forbidden(g, src) || mrc_is_board_forbidden(g, dst))
        return 0;
    if (dist[dst] == MRC_MAX) return 0;
    uint32_t *hops = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *queue = (uint32_t *)malloc(n * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) hops[i] = MRC_MAX;
    size_t qh = 0, qt = 0;
    hops[src] = 0;
    queue[qt++] = src;
    while (qh < qt) {
        uint32_t u = queue[qh++];
        uint32_t hu = hops[u], du = dist[u];
        if (du == MRC_MAX) continue;
        uint32_t nb[4];
        int k = neighbors4(d, u, nb);
        for (int j = 0; j < k; ++j) {
            uint32_t v = nb[j];
            if (mrc_is_obstacle(g, v) || mrc_is_board_forbidden(g, v) ||
                mrc_is_board_planar_step_forbidden(g, u, v) || sat_add32(du, g->cost[v]) != dist[v] ||
                hops[v] != MRC_MAX)
                continue;
            hops[v] = sat_add32(hu, 1);
            queue[qt++] = v;
        }
    }
    size_t len = 0;
    if (hops[dst] != MRC_MAX) {
        uint32_t cur = dst;
        path[len++] = dst;
        while (cur != src) {
            uint32_t need = dist[cur], step = g->cost[cur];
            if (hops[cur] == 0) {
                len = 0;
                break;
            }
            uint32_t need_hops = hops[cur] - 1;
            uint32_t nb[4], next = MRC_MAX;
            int k = neighbors4(d, cur, nb);
            for (int j = 0; j < k; ++j) {
                uint32_t p = nb[j], dp = dist[p];
                if (dp != MRC_MAX && !mrc_is_board_planar_step_forbidden(g, p, cur) &&
                    sat_add32(dp, step) == need && hops[p] == need_hops) {
                    next = p;
                    break;
                }
            }
            if (next == MRC_MAX) {
                len = 0;
                break;
            }
            path[len++] = next;
            cur = next;
        }
    }
    free(hops);
    free(queue);
    for (size_t i = 0; i < len / 2; ++i) {
        uint32_t t = path[i];
        path[i] = path[len - 1 - i];
        path[len - 1 - i] = t;
    }
    return len;
}

/* ------------------------------------------------------------------------- */
/* search-spec closures (lee.rs:39-65, negotiated.rs:2801-2847, :2914-3043)    */
/* ------------------------------------------------------------------------- */

static inline int in_window(const mrc_spec *s, uint32_t c) {
    mrc_dims d = s->grid->dims;
    uint32_t r = c % (d.w * d.h), x = r % d.w, y = r / d.w;
    return x >= s->window.x0 && x <= s->window.x1 && y >= s->window.y0 && y <= s->window.y1;
}

static inline int is_pad(const mrc_spec *s, uint32_t c) {
    return sorted_contains(s->pads, s->n_pads, c);
}

/* enter_weight (negotiated.rs:2806-2812, :2914-2920; lee.rs:55-59) */
static inline uint32_t enter_weight(const mrc_spec *s, uint32_t c) {
    const mrc_grid *g = s->grid;
    if (mrc_is_obstacle(g, c) && is_pad(s, c)) return 1;
    return g->cost[c];
}

/* blocked_fn.  `forced` cells (src/dst of the alone search) stay passable
 * unless outside the window or on a board trace-node (negotiated.rs:2942-2955). */
/* halo_is_foreign (negotiated.rs:190-192): HALO_MIXED (-2) or another group. */
static int halo_is_foreign(int64_t h, int64_t own) { return h == -2 || (h >= 0 && h != own); }

static int blocked(const mrc_spec *s, uint32_t c, const uint32_t *forced, size_t n_forced) {
    const mrc_grid *g = s->grid;
    if (s->mode == MRC_MODE_LEE) /* lee.rs:61 */
        return mrc_is_board_forbidden(g, c) || (mrc_is_obstacle(g, c) && !is_pad(s, c));
    if (!in_window(s, c)) return 1;
    if (mrc_is_board_forbidden(g, c)) return 1;
    /* route_legal_once: foreign copper blocks even the endpoints (:2946-2953) */
    if (s->legal_owner && s->legal_owner[c] >= 0 && s->legal_owner[c] != s->own_group) return 1;
    for (size_t i = 0; i < n_forced; ++i)
        if (forced[i] == c) return 0;
    /* ... the endpoints stay reachable, then a foreign halo blocks unless the
     * cell is this group's own copper (:2954-2966) */
    if (s->legal_halo && halo_is_foreign(s->legal_halo[c], s->own_group) &&
        !(s->legal_owner && s->legal_owner[c] == s->own_group))
        return 1;
    return mrc_is_obstacle(g, c) && !is_pad(s, c);
}

/* priced_with_base (negotiated.rs:2777-2797) generalised to an additive plane
 * plus a foreign-halo penalty; without them it is passable_search_cost (:2929). */
static inline uint64_t sat_sub64(uint64_t a, uint64_t b) { return a > b ? a - b : 0; }

/* CountedCellSet::count of the own halo (0 when absent). */
static uint32_t own_halo_count(const mrc_spec *s, uint32_t c) {
    size_t lo = 0, hi = s->n_own_halo;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (s->own_halo[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < s->n_own_halo && s->own_halo[lo] == c) ? s->own_halo_count[lo] : 0;
}

/* priced_with_base of route_negotiated (negotiated.rs:2777-2797): base plus
 * history plus the foreign congestion -- the snapshot less the net's own
 * copper (self_present) and own halo multiplicity (self_halo), either from
 * the fused shared term or term by term -- then passable_search_cost. */
static uint32_t priced_negotiated(const mrc_spec *s, uint32_t c, uint64_t base) {
    const uint64_t scale = 16; /* SCALE == CLEARANCE_NEG_WEIGHT (negotiated.rs:71, :99) */
    uint64_t self_present = (s->has_own_path && sorted_contains(s->own_path, s->n_own_path, c)) ? 1 : 0;
    uint64_t self_halo = own_halo_count(s, c);
    uint64_t pf = (uint64_t)s->neg_pfac * scale;
    uint64_t priced;
    if (s->neg_shared) {
        uint64_t foreign = sat_sub64(sat_sub64(s->neg_shared[c], pf * self_present), pf * self_halo);
        priced = sat_add64(sat_add64(base, s->neg_history[c]), foreign);
    } else {
        uint64_t fp = sat_sub64(s->neg_present[c], self_present);
        uint64_t fh = sat_sub64(s->neg_halo ? s->neg_halo[c] : 0, self_halo);
        priced = sat_add64(sat_add64(sat_add64(base, s->neg_history[c]), pf * fp), pf * fh);
    }
    return mrc_passable_search_cost(priced);
}

static inline uint32_t priced(const mrc_spec *s, uint32_t c, uint64_t base) {
    if (s->neg_history) return priced_negotiated(s, c, base);
    uint64_t p = base;
    if (s->additive) p = sat_add64(p, s->additive[c]);
    if (s->halo_owner) {
        int32_t o = s->halo_owner[c];
        if (o != -1 && o != s->own_id) p = sat_add64(p, s->halo_penalty);
    }
    return mrc_passable_search_cost(p);
}

/* cost_fn(u, v) for a planar neighbour and via_step(u, v) for a via neighbour;
 * MRC_MAX means "no edge" (OBSTACLE / None). */
uint32_t mrc_step_cost(const mrc_spec *s, uint32_t u, uint32_t v) {
    const mrc_grid *g = s->grid;
    mrc_dims d = g->dims;
    uint32_t plane = d.w * d.h;
    uint32_t ul = u / plane, vl = v / plane;
    if (ul == vl) {
        if (mrc_is_board_planar_step_forbidden(g, u, v)) return MRC_MAX;
        if (s->mode == MRC_MODE_LEE) return enter_weight(s, v); /* lee.rs:52-59 */
        uint32_t ur = u % plane, vr = v % plane;
        uint32_t ux = ur % d.w, uy = ur / d.w, vx = vr % d.w, vy = vr / d.w;
        uint32_t e = (ux != vx) ? s->edges.x[ux < vx ? ux : vx] : s->edges.y[uy < vy ? uy : vy];
        return priced(s, v, sat_mul64(e, enter_weight(s, v)));
    }
    if (s->mode == MRC_MODE_LEE) return MRC_MAX; /* lee.rs:63 `|_, _| None` */
    /* via_step (negotiated.rs:2995-3043 with no owners/halo/guard; :2833-2847) */
    if (!mrc_is_step_legal_allowed(&s->edges, d.layers, ul, vl)) return MRC_MAX;
    if (mrc_is_board_via_forbidden(g, u) || mrc_is_board_via_forbidden(g, v)) return MRC_MAX;
    if ((mrc_is_via_forbidden(g, u) && !sorted_contains(s->via_pads, s->n_via_pads, u)) ||
        (mrc_is_via_forbidden(g, v) && !sorted_contains(s->via_pads, s->n_via_pads, v)))
        return MRC_MAX;
    uint32_t lo = ul < vl ? ul : vl;
    return priced(s, v, sat_mul64(s->edges.via[lo], enter_weight(s, v)));
}

/* ------------------------------------------------------------------------- */
/* astar_buf (dijkstra.rs:125-401)                                            */
/* ------------------------------------------------------------------------- */

struct mrc_searchbuf {
    uint32_t *dist, *hops, *pred, *stamp;
    uint32_t gen;
    size_t n;
    heap_t heap;
};

mrc_searchbuf *mrc_searchbuf_new(size_t n) {
    mrc_searchbuf *b = (mrc_searchbuf *)calloc(1, sizeof(*b));
    b->n = n;
    b->dist = (uint32_t *)malloc(n * sizeof(uint32_t) + 4);
    b->hops = (uint32_t *)malloc(n * sizeof(uint32_t) + 4);
    b->pred = (uint32_t *)malloc(n * sizeof(uint32_t) + 4);
    b->stamp = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    b->gen = 0;
    heap_init(&b->heap, 0);
    return b;
}

void mrc_searchbuf_free(mrc_searchbuf *b) {
    if (!b) return;
    free(b->dist);
    free(b->hops);
    free(b->pred);
    free(b->stamp);
    heap_free(&b->heap);
    free(b);
}

/* SearchBuf::begin (dijkstra.rs:153-162) */
static void sb_begin(mrc_searchbuf *b) {
    if (b->gen == UINT32_MAX) {
        memset(b->stamp, 0, b->n * sizeof(uint32_t));
        b->gen = 1;
    } else {
        b->gen++;
    }
}
static inline uint32_t sb_dist(const mrc_searchbuf *b, uint32_t c) {
    return b->stamp[c] == b->gen ? b->dist[c] : MRC_MAX;
}
static inline uint32_t sb_hops(const mrc_searchbuf *b, uint32_t c) {
    return b->stamp[c] == b->gen ? b->hops[c] : MRC_MAX;
}
static inline uint32_t sb_pred(const mrc_searchbuf *b, uint32_t c) {
    return b->stamp[c] == b->gen ? b->pred[c] : MRC_NO_PRED;
}
static inline void sb_set(mrc_searchbuf *b, uint32_t c, uint32_t d, uint32_t h, uint32_t p) {
    b->dist[c] = d;
    b->hops[c] = h;
    b->pred[c] = p;
    b->stamp[c] = b->gen;
}

/* manhattan_scaled (negotiated.rs:1245-1282) over prefix sums of the per-gap
 * edge prices (axis_cost_prefix builds exactly these sums, :1255-1264). */
typedef struct {
    uint64_t *xp, *yp;
} heur_t;

static uint32_t axis_leg(const uint64_t *p, uint32_t i, uint32_t j) {
    uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
    uint64_t t = p[hi] - p[lo];
    return t < (uint64_t)MRC_MAX ? (uint32_t)t : MRC_MAX;
}

static uint32_t heuristic(const mrc_spec *s, const heur_t *h, uint32_t c, uint32_t dst) {
    if (!s->use_heuristic || s->mode == MRC_MODE_LEE) return 0;
    mrc_dims d = s->grid->dims;
    uint32_t plane = d.w * d.h;
    uint32_t cr = c % plane, dr = dst % plane;
    uint32_t planar = sat_add32(axis_leg(h->xp, cr % d.w, dr % d.w), axis_leg(h->yp, cr / d.w, dr / d.w));
    uint32_t cl = c / plane, dl = dst / plane;
    uint32_t dlay = cl > dl ? cl - dl : dl - cl;
    uint64_t vm = (uint64_t)dlay * s->min_via_cost;
    uint32_t vmc = vm < MRC_MAX ? (uint32_t)vm : MRC_MAX;
    return sat_add32(planar, vmc);
}

static void relax_one(mrc_searchbuf *b, const mrc_spec *s, const heur_t *hh, uint32_t u,
                      uint32_t v, uint32_t du, uint32_t hu, uint32_t src, uint32_t dst,
                      const uint32_t *forced) {
    if (blocked(s, v, forced, 2)) return;
    uint32_t step = mrc_step_cost(s, u, v);
    if (step == MRC_MAX) return; /* planar OBSTACLE / via None */
    uint32_t nd = sat_add32(du, step);
    uint32_t nh = sat_add32(hu, 1);
    uint32_t old = sb_dist(b, v);
    if (nd < old) {
        sb_set(b, v, nd, nh, u);
        heap_push(&b->heap, ((uint64_t)sat_add32(nd, heuristic(s, hh, v, dst)) << 32) | v, 0);
    } else if (nd == old && v != src) {
        uint32_t oh = sb_hops(b, v);
        if (nh < oh) {
            sb_set(b, v, nd, nh, u);
            heap_push(&b->heap, ((uint64_t)sat_add32(nd, heuristic(s, hh, v, dst)) << 32) | v, 0);
        } else if (nh == oh && u < sb_pred(b, v)) {
            sb_set(b, v, nd, nh, u);
        }
    }
}

size_t mrc_astar(mrc_searchbuf *b, const mrc_spec *s, uint32_t src, uint32_t dst, uint32_t *path,
                 uint32_t *cost_out) {
    mrc_dims d = s->grid->dims;
    uint32_t plane = d.w * d.h;
    uint32_t forced[2] = {src, dst};
    size_t n = dims_len(d);
    heur_t hh = {NULL, NULL};
    if (s->use_heuristic && s->mode != MRC_MODE_LEE) {
        hh.xp = (uint64_t *)malloc(((size_t)d.w + 1) * sizeof(uint64_t));
        hh.yp = (uint64_t *)malloc(((size_t)d.h + 1) * sizeof(uint64_t));
        hh.xp[0] = 0;
        for (uint32_t k = 0; k + 1 < d.w; ++k) hh.xp[k + 1] = hh.xp[k] + s->edges.x[k];
        hh.yp[0] = 0;
        for (uint32_t k = 0; k + 1 < d.h; ++k) hh.yp[k + 1] = hh.yp[k] + s->edges.y[k];
    }
    sb_begin(b);
    b->heap.n = 0;
    size_t len = 0;
    if (cost_out) *cost_out = MRC_MAX;
    if (blocked(s, src, forced, 2)) goto done;
    sb_set(b, src, 0, 0, MRC_NO_PRED);
    heap_push(&b->heap, ((uint64_t)heuristic(s, &hh, src, dst) << 32) | src, 0);
    int have_goal = 0;
    uint32_t goal = 0;
    uint64_t key;
    while (heap_pop(&b->heap, &key, NULL)) {
        uint32_t prio = (uint32_t)(key >> 32), u = (uint32_t)key;
        if (have_goal && prio > goal) break;
        uint32_t du = sb_dist(b, u), hu = sb_hops(b, u);
        if (sat_sub32(prio, heuristic(s, &hh, u, dst)) > du) continue;
        if (u == dst) {
            have_goal = 1;
            goal = du;
            continue;
        }
        uint32_t layer = u / plane, planar = u % plane, x = planar % d.w;
        if (planar >= d.w) relax_one(b, s, &hh, u, u - d.w, du, hu, src, dst, forced);
        if (x > 0) relax_one(b, s, &hh, u, u - 1, du, hu, src, dst, forced);
        if (x + 1 < d.w) relax_one(b, s, &hh, u, u + 1, du, hu, src, dst, forced);
        if (planar + d.w < plane) relax_one(b, s, &hh, u, u + d.w, du, hu, src, dst, forced);
        if (layer > 0) relax_one(b, s, &hh, u, u - plane, du, hu, src, dst, forced);
        if (layer + 1 < d.layers) relax_one(b, s, &hh, u, u + plane, du, hu, src, dst, forced);
    }
    {
        uint32_t total = sb_dist(b, dst);
        if (total == MRC_MAX) goto done;
        uint32_t cur = dst;
        path[len++] = dst;
        while (cur != src) {
            if (len > n) {
                len = 0;
                goto done;
            }
            uint32_t p = sb_pred(b, cur);
            if (p == MRC_NO_PRED) {
                len = 0;
                goto done;
            }
            path[len++] = p;
            cur = p;
        }
        for (size_t i = 0; i < len / 2; ++i) {
            uint32_t t = path[i];
            path[i] = path[len - 1 - i];
            path[len - 1 - i] = t;
        }
        if (cost_out) *cost_out = total;
    }
done:
    free(hh.xp);
    free(hh.yp);
    return len;
}

/* ------------------------------------------------------------------------- */
/* complete canonical fields: label-setting lexicographic Dijkstra             */
/* ------------------------------------------------------------------------- */

/* The canonical label of a cell is the lexicographic minimum of (search cost,
 * hops) over all paths from any source (astar_buf's label order,
 * dijkstra.rs:221-231).  Every step weight (c, 1) is strictly positive in that
 * order, so a heap keyed by (dist, hops) settles each cell exactly once. */
void mrc_field_lex(const mrc_spec *s, const uint32_t *sources, size_t n_sources,
                   const uint32_t *forced, size_t n_forced, uint32_t *dist, uint32_t *hops_out) {
    mrc_dims d = s->grid->dims;
    size_t n = dims_len(d);
    uint32_t *hops = hops_out ? hops_out : (uint32_t *)malloc(n * sizeof(uint32_t) + 4);
    uint32_t plane = d.w * d.h;
    for (size_t i = 0; i < n; ++i) {
        dist[i] = MRC_MAX;
        hops[i] = MRC_MAX;
    }
    heap_t h;
    heap_init(&h, 1);
    for (size_t i = 0; i < n_sources; ++i) {
        uint32_t c = sources[i];
        if ((size_t)c >= n || blocked(s, c, forced, n_forced)) continue;
        dist[c] = 0;
        hops[c] = 0;
        heap_push(&h, 0, c); /* key (dist<<32 | hops), aux cell */
    }
    uint64_t key;
    uint32_t u;
    while (heap_pop(&h, &key, &u)) {
        uint32_t du = (uint32_t)(key >> 32), hk = (uint32_t)key;
        if (du != dist[u] || hk != hops[u]) continue;
        uint32_t cand[6];
        int k = 0;
        uint32_t layer = u / plane, planar = u % plane, x = planar % d.w;
        if (planar >= d.w) cand[k++] = u - d.w;
        if (x > 0) cand[k++] = u - 1;
        if (x + 1 < d.w) cand[k++] = u + 1;
        if (planar + d.w < plane) cand[k++] = u + d.w;
        if (layer > 0) cand[k++] = u - plane;
        if (layer + 1 < d.layers) cand[k++] = u + plane;
        for (int j = 0; j < k; ++j) {
            uint32_t v = cand[j];
            if (blocked(s, v, forced, n_forced)) continue;
            uint32_t step = mrc_step_cost(s, u, v);
            if (step == MRC_MAX) continue;
            uint64_t nd = (uint64_t)du + step;
            if (nd >= MRC_MAX) continue; /* saturating add reaches the sentinel */
            uint32_t nh = hk + 1;
            if ((uint32_t)nd < dist[v] || ((uint32_t)nd == dist[v] && nh < hops[v])) {
                dist[v] = (uint32_t)nd;
                hops[v] = nh;
                heap_push(&h, ((uint64_t)nd << 32) | nh, v);
            }
        }
    }
    heap_free(&h);
    if (!hops_out) free(hops);
}

size_t mrc_reconstruct(const mrc_spec *s, const uint32_t *dist, const uint32_t *hops, uint32_t dst,
                       uint32_t *path) {
    mrc_dims d = s->grid->dims;
    size_t n = dims_len(d);
    uint32_t plane = d.w * d.h;
    if ((size_t)dst >= n || dist[dst] == MRC_MAX) return 0;
    size_t len = 0;
    uint32_t cur = dst;
    path[len++] = cur;
    while (hops[cur] != 0) {
        if (len > n) return 0;
        uint32_t layer = cur / plane, planar = cur % plane, x = planar % d.w;
        uint32_t cand[6];
        int k = 0;
        /* ascending CellIdx: lower layer, up, left, right, down, upper layer */
        if (layer > 0) cand[k++] = cur - plane;
        if (planar >= d.w) cand[k++] = cur - d.w;
        if (x > 0) cand[k++] = cur - 1;
        if (x + 1 < d.w) cand[k++] = cur + 1;
        if (planar + d.w < plane) cand[k++] = cur + d.w;
        if (layer + 1 < d.layers) cand[k++] = cur + plane;
        uint32_t next = MRC_MAX;
        for (int j = 0; j < k; ++j) {
            uint32_t p = cand[j];
            if (dist[p] == MRC_MAX) continue;
            uint32_t step = mrc_step_cost(s, p, cur);
            if (step == MRC_MAX) continue;
            if ((uint64_t)dist[p] + step == dist[cur] && hops[p] + 1 == hops[cur]) {
                next = p;
                break;
            }
        }
        if (next == MRC_MAX) return 0;
        path[len++] = next;
        cur = next;
    }
    for (size_t i = 0; i < len / 2; ++i) {
        uint32_t t = path[i];
        path[i] = path[len - 1 - i];
        path[len - 1 - i] = t;
    }
    return len;
}

/* ------------------------------------------------------------------------- */
/* provider contract                                                          */
/* ------------------------------------------------------------------------- */

static int u32_cmp(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y);
}

/* negotiated.rs:1284-1331 */
int mrc_provider_path_is_valid(const mrc_grid *g, uint32_t src, uint32_t dst, const uint32_t *pads,
                               size_t n_pads, const uint32_t *via_pads, size_t n_via_pads,
                               const mrc_edges *e, const uint32_t *path, size_t len) {
    mrc_dims d = g->dims;
    size_t n = dims_len(d);
    if (len == 0 || len > n || path[0] != src || path[len - 1] != dst) return 0;
    uint32_t *seen = (uint32_t *)malloc(len * sizeof(uint32_t));
    memcpy(seen, path, len * sizeof(uint32_t));
    int ok = 1;
    for (size_t i = 0; i < len && ok; ++i) {
        uint32_t c = path[i];
        if ((size_t)c >= n || mrc_is_board_forbidden(g, c) ||
            (mrc_is_obstacle(g, c) && !sorted_contains(pads, n_pads, c)))
            ok = 0;
    }
    if (ok) {
        qsort(seen, len, sizeof(uint32_t), u32_cmp);
        for (size_t i = 1; i < len; ++i)
            if (seen[i] == seen[i - 1]) ok = 0;
    }
    free(seen);
    if (!ok) return 0;
    uint32_t plane = d.w * d.h;
    for (size_t i = 0; i + 1 < len; ++i) {
        uint32_t a = path[i], b = path[i + 1];
        uint32_t al = a / plane, ar = a % plane, ax = ar % d.w, ay = ar / d.w;
        uint32_t bl = b / plane, br = b % plane, bx = br % d.w, by = br / d.w;
        if (al == bl) {
            uint32_t adx = ax > bx ? ax - bx : bx - ax, ady = ay > by ? ay - by : by - ay;
            if (!((adx == 1 && ay == by) || (ady == 1 && ax == bx))) return 0;
            if (mrc_is_board_planar_step_forbidden(g, a, b)) return 0;
        } else {
            if (!(ax == bx && ay == by)) return 0;
            if (!mrc_is_step_legal_allowed(e, d.layers, al, bl)) return 0;
            if (mrc_is_board_via_forbidden(g, a) || mrc_is_board_via_forbidden(g, b)) return 0;
            if (mrc_is_via_forbidden(g, a) && !sorted_contains(via_pads, n_via_pads, a)) return 0;
            if (mrc_is_via_forbidden(g, b) && !sorted_contains(via_pads, n_via_pads, b)) return 0;
        }
    }
    return 1;
}

/* negotiated.rs:1106-1115 */
uint32_t mrc_grid_path_cost(const mrc_grid *g, const uint32_t *pads, size_t n_pads,
                            const uint32_t *path, size_t len) {
    uint32_t acc = 0;
    for (size_t i = 1; i < len; ++i) {
        uint32_t c = path[i];
        uint32_t step = (mrc_is_obstacle(g, c) && sorted_contains(pads, n_pads, c)) ? 1 : g->cost[c];
        acc = sat_add32(acc, step);
    }
    return acc;
}

/* ------------------------------------------------------------------------- */
/* threaded batches (negotiated.rs:2115-2184 rayon into_par_iter().map_init)   */
/* ------------------------------------------------------------------------- */

typedef struct {
    const mrc_spec *tmpl;
    const mrc_net *nets;
    size_t n_nets;
    const mrc_window *windows;
    int retry_full;
    uint32_t *path_cells;
    const uint64_t *path_off;
    uint32_t *path_len;
    uint32_t *cost;
    /* fields mode */
    const uint32_t *const *sources;
    const size_t *n_sources;
    uint32_t *dist_out, *hops_out;
    atomic_size_t next;
} batch_job;

static void spec_for_net(const batch_job *j, size_t i, mrc_spec *s) {
    *s = *j->tmpl;
    s->pads = j->nets[i].pads;
    s->n_pads = j->nets[i].n_pads;
    s->via_pads = j->nets[i].via_pads;
    s->n_via_pads = j->nets[i].n_via_pads;
    s->own_id = j->nets[i].own_id;
    s->has_own_path = j->nets[i].has_own_path;
    s->own_path = j->nets[i].own_path;
    s->n_own_path = j->nets[i].n_own_path;
    s->own_halo = j->nets[i].own_halo;
    s->own_halo_count = j->nets[i].own_halo_count;
    s->n_own_halo = j->nets[i].n_own_halo;
    if (j->tmpl->legal_owner || j->tmpl->legal_halo) s->own_group = j->nets[i].own_group;
    s->window = j->windows ? j->windows[i] : mrc_window_full(j->tmpl->grid->dims);
}

static void *alone_worker(void *arg) {
    batch_job *j = (batch_job *)arg;
    size_t n = dims_len(j->tmpl->grid->dims);
    mrc_searchbuf *b = mrc_searchbuf_new(n);
    for (;;) {
        size_t i = atomic_fetch_add(&j->next, 1);
        if (i >= j->n_nets) break;
        mrc_spec s;
        spec_for_net(j, i, &s);
        uint32_t *out = j->path_cells + j->path_off[i];
        uint32_t c = MRC_MAX;
        size_t len = mrc_astar(b, &s, j->nets[i].src, j->nets[i].dst, out, &c);
        if (len == 0 && j->retry_full) { /* .or_else(full board), negotiated.rs:2155-2173 */
            s.window = mrc_window_full(s.grid->dims);
            len = mrc_astar(b, &s, j->nets[i].src, j->nets[i].dst, out, &c);
        }
        j->path_len[i] = (uint32_t)len;
        j->cost[i] = len ? c : MRC_MAX;
    }
    mrc_searchbuf_free(b);
    return NULL;
}

static void run_workers(batch_job *j, int n_threads, void *(*fn)(void *)) {
    if (n_threads < 1) n_threads = 1;
    atomic_init(&j->next, 0);
    if (n_threads == 1) {
        fn(j);
        return;
    }
    pthread_t *t = (pthread_t *)malloc((size_t)n_threads * sizeof(pthread_t));
    for (int k = 0; k < n_threads; ++k) pthread_create(&t[k], NULL, fn, j);
    for (int k = 0; k < n_threads; ++k) pthread_join(t[k], NULL);
    free(t);
}

void mrc_alone_batch(const mrc_spec *tmpl, const mrc_net *nets, size_t n_nets,
                     const mrc_window *windows, int retry_full, int n_threads,
                     uint32_t *path_cells, const uint64_t *path_off, uint32_t *path_len,
                     uint32_t *cost) {
    batch_job j;
    memset(&j, 0, sizeof(j));
    j.tmpl = tmpl;
    j.nets = nets;
    j.n_nets = n_nets;
    j.windows = windows;
    j.retry_full = retry_full;
    j.path_cells = path_cells;
    j.path_off = path_off;
    j.path_len = path_len;
    j.cost = cost;
    run_workers(&j, n_threads, alone_worker);
}

static void *field_worker(void *arg) {
    batch_job *j = (batch_job *)arg;
    size_t n = dims_len(j->tmpl->grid->dims);
    for (;;) {
        size_t i = atomic_fetch_add(&j->next, 1);
        if (i >= j->n_nets) break;
        mrc_spec s;
        spec_for_net(j, i, &s);
        size_t ns = j->n_sources[i];
        uint32_t *forced = (uint32_t *)malloc((ns + 1) * sizeof(uint32_t));
        memcpy(forced, j->sources[i], ns * sizeof(uint32_t));
        size_t nf = ns;
        if (j->nets[i].dst != MRC_MAX) forced[nf++] = j->nets[i].dst;
        mrc_field_lex(&s, j->sources[i], ns, forced, nf, j->dist_out + i * n,
                      j->hops_out ? j->hops_out + i * n : NULL);
        free(forced);
    }
    return NULL;
}

void mrc_fields_batch(const mrc_spec *tmpl, const mrc_net *nets, size_t n_nets,
                      const uint32_t *const *sources, const size_t *n_sources,
                      const mrc_window *windows, int n_threads, uint32_t *dist_out,
                      uint32_t *hops_out) {
    batch_job j;
    memset(&j, 0, sizeof(j));
    j.tmpl = tmpl;
    j.nets = nets;
    j.n_nets = n_nets;
    j.windows = windows;
    j.sources = sources;
    j.n_sources = n_sources;
    j.dist_out = dist_out;
    j.hops_out = hops_out;
    run_workers(&j, n_threads, field_worker);
}

/* mr-core/src/lib.rs:575-593: a fresh seen-set per path (the reference's
 * HashSet sized to the path), here a sorted copy of the path's cells. */
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

void mrc_congestion_from(mrc_dims d, size_t n_paths, const uint64_t *off, const uint32_t *cells,
                         uint32_t *out) {
    size_t n = dims_len(d);
    memset(out, 0, n * sizeof(uint32_t));
    for (size_t i = 0; i < n_paths; ++i) {
        size_t len = (size_t)(off[i + 1] - off[i]);
        if (!len) continue;
        uint32_t *tmp = (uint32_t *)malloc(len * sizeof(uint32_t));
        memcpy(tmp, cells + off[i], len * sizeof(uint32_t));
        qsort(tmp, len, sizeof(uint32_t), cmp_u32);
        for (size_t k = 0; k < len; ++k)
            if ((k == 0 || tmp[k] != tmp[k - 1]) && tmp[k] < n) out[tmp[k]] += 1;
        free(tmp);
    }
}
