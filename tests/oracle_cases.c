/* The oracle (oracle/mrcpu.c) exercised as a C program, so that it can run
 * under the host sanitizers: AddressSanitizer with UndefinedBehaviorSanitizer,
 * and ThreadSanitizer for the threaded batch entry points.
 *
 * On random small boards (obstacles, board masks, via masks, forbidden vias,
 * pads and via pads, windows, additive prices) it checks the oracle's own
 * invariants, each comparing two different code paths:
 *   - every A* path is accepted by mrc_provider_path_is_valid and its cost is
 *     the sum of mrc_step_cost along it;
 *   - the heuristic on and off give the same path and cost;
 *   - the complete field (mrc_field_lex) at dst equals the A* cost, and
 *     mrc_reconstruct over it returns the A* path;
 *   - mrc_alone_batch on four threads equals one A* per net;
 *   - mrc_fields_batch on three threads equals mrc_field_lex per net;
 *   - mrc_congestion_from counts each path's distinct cells once;
 *   - the planar sweep field equals the BFS field.
 * Exit status 0 when every case holds. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../oracle/mrcpu.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd(uint32_t n) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (uint32_t)(rng_state % n);
}

static int fails = 0;
#define EXPECT(c, ...)                 \
  do {                                 \
    if (!(c)) {                        \
      printf("FAIL: " __VA_ARGS__);    \
      printf("\n");                    \
      fails++;                         \
    }                                  \
  } while (0)

static int cmp_u32(const void* a, const void* b) {
  const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

int main(void) {
  for (int board = 0; board < 60; ++board) {
    const uint32_t w = 3 + rnd(14), h = 3 + rnd(14), L = 1 + rnd(3);
    const size_t n = (size_t)w * h * L;
    const mrc_dims d = {w, h, L};
    uint32_t* cost = malloc(n * 4);
    uint8_t* vf = malloc(n);
    uint8_t* bm = malloc(n);
    uint32_t* add = malloc(n * 4);
    for (size_t c = 0; c < n; ++c) {
      cost[c] = rnd(100) < 15 ? 0xFFFFFFFFu : 1 + rnd(3);
      vf[c] = rnd(100) < 10;
      bm[c] = rnd(100) < 8 ? (uint8_t)(1u << rnd(6)) : 0;
      add[c] = rnd(5);
    }
    const mrc_grid g = {d, cost, board % 2 ? vf : NULL, board % 3 ? bm : NULL};
    uint32_t ex[32], ey[32], ev[8];
    uint8_t va[8];
    for (uint32_t i = 0; i < 32; ++i) ex[i] = 1 + rnd(20), ey[i] = 1 + rnd(20);
    for (uint32_t i = 0; i < 8; ++i) ev[i] = 5 + rnd(40), va[i] = rnd(5) != 0;
    mrc_spec s;
    memset(&s, 0, sizeof s);
    s.grid = &g;
    s.edges = (mrc_edges){ex, ey, ev, va};
    s.window = mrc_window_full(d);
    s.mode = MRC_MODE_ALONE;
    s.halo_owner = NULL;
    s.additive = board % 4 == 0 ? add : NULL;
    s.own_id = -1;
    s.legal_owner = NULL;
    s.legal_halo = NULL;
    s.own_group = -1;
    mrc_searchbuf* buf = mrc_searchbuf_new(n);
    uint32_t* path = malloc(n * 4);
    uint32_t* path2 = malloc(n * 4);
    uint32_t* dist = malloc(n * 4);
    uint32_t* hops = malloc(n * 4);
    enum { NETS = 6 };
    mrc_net nets[NETS];
    uint32_t pads[NETS][3];
    uint32_t paths_len[NETS], costs[NETS];
    uint32_t* cells = malloc(NETS * n * 4);
    uint64_t off[NETS + 1];
    for (int k = 0; k < NETS; ++k) {
      memset(&nets[k], 0, sizeof nets[k]);
      nets[k].src = rnd((uint32_t)n);
      nets[k].dst = rnd((uint32_t)n);
      /* own pads hold the net's endpoints, as rasterised nets' do (a
       * provider path may cross an obstacle only on an own pad) */
      pads[k][0] = nets[k].src;
      pads[k][1] = nets[k].dst;
      pads[k][2] = rnd((uint32_t)n);
      qsort(pads[k], 3, 4, cmp_u32);
      size_t np = 1;
      for (int j = 1; j < 3; ++j)
        if (pads[k][j] != pads[k][np - 1]) pads[k][np++] = pads[k][j];
      nets[k].pads = pads[k];
      nets[k].n_pads = np;
      nets[k].via_pads = pads[k];
      nets[k].n_via_pads = 1;
      nets[k].own_id = -1;
      nets[k].own_group = -1;
      off[k] = (uint64_t)k * n;
    }
    off[NETS] = (uint64_t)NETS * n;
    for (int k = 0; k < NETS; ++k) {
      mrc_spec t = s;
      t.pads = nets[k].pads;
      t.n_pads = nets[k].n_pads;
      t.via_pads = nets[k].via_pads;
      t.n_via_pads = nets[k].n_via_pads;
      uint32_t c1 = 0, c2 = 0;
      const size_t n1 = mrc_astar(buf, &t, nets[k].src, nets[k].dst, path, &c1);
      t.use_heuristic = 1;
      t.min_via_cost = 5;
      const size_t n2 = mrc_astar(buf, &t, nets[k].src, nets[k].dst, path2, &c2);
      t.use_heuristic = 0;
      EXPECT(n1 == n2 && (n1 == 0 || (c1 == c2 && !memcmp(path, path2, n1 * 4))), "heuristic board %d net %d", board,
             k);
      paths_len[k] = (uint32_t)n1;
      costs[k] = c1;
      memcpy(cells + off[k], path, n1 * 4);
      if (n1) {
        uint32_t sum = 0;
        for (size_t i = 1; i < n1; ++i) sum += mrc_step_cost(&t, path[i - 1], path[i]);
        EXPECT(sum == c1, "step-cost sum board %d net %d", board, k);
        if (!s.additive)
          EXPECT(mrc_provider_path_is_valid(&g, nets[k].src, nets[k].dst, t.pads, t.n_pads, t.via_pads, t.n_via_pads,
                                            &t.edges, path, n1),
                 "validity board %d net %d", board, k);
      }
      const uint32_t src = nets[k].src;
      const uint32_t forced[2] = {nets[k].src, nets[k].dst};
      mrc_field_lex(&t, &src, 1, forced, 2, dist, hops);
      if (n1) {
        EXPECT(dist[nets[k].dst] == c1, "field cost board %d net %d", board, k);
        const size_t n3 = mrc_reconstruct(&t, dist, hops, nets[k].dst, path2);
        EXPECT(n3 == n1 && !memcmp(path, path2, n1 * 4), "reconstruct board %d net %d", board, k);
      } else {
        EXPECT(dist[nets[k].dst] == 0xFFFFFFFFu, "unreached field board %d net %d", board, k);
      }
    }
    {
      uint32_t* bcells = malloc(NETS * n * 4);
      uint32_t blen[NETS], bcost[NETS];
      mrc_alone_batch(&s, nets, NETS, NULL, 1, 4, bcells, off, blen, bcost);
      for (int k = 0; k < NETS; ++k)
        EXPECT(blen[k] == paths_len[k] && (!blen[k] || (bcost[k] == costs[k] &&
                                                         !memcmp(bcells + off[k], cells + off[k], blen[k] * 4))),
               "alone batch board %d net %d", board, k);
      free(bcells);
    }
    {
      const uint32_t* srcs[NETS];
      size_t ns[NETS];
      for (int k = 0; k < NETS; ++k) srcs[k] = &nets[k].src, ns[k] = 1;
      uint32_t* bd = malloc(NETS * n * 4);
      uint32_t* bh = malloc(NETS * n * 4);
      mrc_fields_batch(&s, nets, NETS, srcs, ns, NULL, 3, bd, bh);
      for (int k = 0; k < NETS; ++k) {
        mrc_spec t = s;
        t.pads = nets[k].pads;
        t.n_pads = nets[k].n_pads;
        t.via_pads = nets[k].via_pads;
        t.n_via_pads = nets[k].n_via_pads;
        const uint32_t forced[2] = {nets[k].src, nets[k].dst};
        mrc_field_lex(&t, &nets[k].src, 1, forced, 2, dist, hops);
        EXPECT(!memcmp(bd + (size_t)k * n, dist, n * 4) && !memcmp(bh + (size_t)k * n, hops, n * 4),
               "fields batch board %d net %d", board, k);
      }
      free(bd);
      free(bh);
    }
    {
      uint64_t coff[NETS + 1];
      uint32_t* flat = malloc(NETS * n * 4 + 4);
      coff[0] = 0;
      for (int k = 0; k < NETS; ++k) {
        memcpy(flat + coff[k], cells + off[k], paths_len[k] * 4);
        coff[k + 1] = coff[k] + paths_len[k];
      }
      uint32_t* cong = calloc(n, 4);
      uint32_t* want = calloc(n, 4);
      mrc_congestion_from(d, NETS, coff, flat, cong);
      uint8_t* seen = calloc(n, 1);
      for (int k = 0; k < NETS; ++k) {
        memset(seen, 0, n);
        for (uint32_t i = 0; i < paths_len[k]; ++i) {
          const uint32_t c = cells[off[k] + i];
          if (!seen[c]) seen[c] = 1, want[c]++;
        }
      }
      EXPECT(!memcmp(cong, want, n * 4), "congestion board %d", board);
      free(seen);
      free(cong);
      free(want);
      free(flat);
    }
    if (L == 1) {
      const mrc_grid g1 = {d, cost, NULL, NULL};
      mrc_sweep_distance_field(&g1, nets[0].src, dist);
      mrc_bfs_distance_field(&g1, nets[0].src, hops);
      EXPECT(!memcmp(dist, hops, n * 4), "sweep == bfs board %d", board);
    }
    free(cells);
    free(path);
    free(path2);
    free(dist);
    free(hops);
    mrc_searchbuf_free(buf);
    free(cost);
    free(vf);
    free(bm);
    free(add);
  }
  printf("%s\n", fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}
