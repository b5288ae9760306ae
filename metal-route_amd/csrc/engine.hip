// Batched shortest-path field engine for gfx950 (MI355X).
//
// A field is the complete canonical label plane of one net: for every cell of
// the net's window (all layers), the lexicographic minimum of (search cost,
// hops) over paths from the net's sources.  This is the label astar_buf
// settles (crates/mr-cpu/src/dijkstra.rs:221-231, :296-374); every step weight
// (c, 1) is strictly positive in that order, so any fair relaxation order
// converges to the same unique fixpoint, which is what makes a parallel,
// tile-local, asynchronous relaxation bit-exact against the CPU search.
//
// Layout in HBM
//   * board planes (shared by every field of a board, resident across calls):
//     cost u32 enter weight, optional additive u32 and halo-owner i32 planes,
//     bmask u8 (trace/obstacle + 6 directed-edge bits), raw u8 (via bits).
//   * labels: one u64 per cell, (dist << 32) | hops, so the lexicographic
//     minimum is a plain unsigned minimum.  Stored tile-major per field
//     ([tile][layer][y][x], 256-byte aligned blocks) so no cache line is shared
//     by two tiles.
//   * per-tile state word (touched / pending / active bits).
//
// Kernels
//   relax_round       one visit of every active (field, tile) item; host loop
//                     until no item changes (the plain per-batch-launch form).
//   relax_persistent  device work queue over (field, tile) items with a
//                     pending/active state machine; no host round trips.
//   path_lengths / walk_paths
//                     canonical reconstruction: from dst, the lowest-CellIdx
//                     predecessor p with label(p) + (step(p, c), 1) == label(c)
//                     (mr-metal/src/lib.rs:304-396 order; masks re-checked).
//   extract_dist      tile-major labels -> window-major u32 dist / hops.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "engine.h"

namespace mrf {

#define HIP_OK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw Error(1, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
  } while (0)

constexpr uint64_t kInf = ~0ull;
constexpr uint64_t kLabLimit = 0xFFFFFFFF00000000ull;  // labels at or above carry dist == MAX

// LDS cell mask bits
enum : uint8_t {
  K_PASS = 0x01,
  K_NY = 0x02,
  K_NX = 0x04,
  K_PX = 0x08,
  K_PY = 0x10,
  K_DN = 0x20,
  K_UP = 0x40,
  K_CHG = 0x80,
};
// per-item state bits
enum : uint32_t { S_TOUCHED = 1u, S_PEND = 2u, S_ACTIVE = 4u };
// device counters
enum : int {
  C_CHANGED = 0,
  C_VISITS = 1,
  C_ROUNDS = 2,
  C_ERROR = 3,
  C_QHEAD = 4,
  C_QTAIL = 5,
  C_PENDING = 6,
  C_NEXTN = 7,
  C_BAR_ARRIVE = 8,
  C_BAR_GEN = 9,
  C_TAKEN0 = 10,  // 10..12
  C_NEXT0 = 13,   // 13..15
  C_PROUNDS = 16,
  C_MISC = 17,
  C_GROUND = 18,  // graph form: current round index
  C_CURN = 19,    // graph form: length of the current round's list
  C_NCOUNTERS = 20
};
// Each device counter sits on its own 128-byte line: polled words and the
// work-list atomics must not share a line, or every pop and push queues
// behind the pollers.
constexpr uint32_t kCtrStride = 32;
constexpr uint32_t kCtrWords = C_NCOUNTERS * kCtrStride;
enum : int { C_PAD_ = 0
};

BoardView DevBoard::view() const {
  BoardView v;
  v.w = w;
  v.h = h;
  v.layers = layers;
  v.plane = w * h;
  v.cost = cost;
  v.add = add;
  v.halo = halo;
  v.halo_pen = halo_pen;
  v.bmask = bmask;
  v.raw = raw;
  v.xe = xe;
  v.ye = ye;
  v.ve = ve;
  for (int i = 0; i < 16; ++i) v.va[i] = via_allowed.size() > (size_t)i ? via_allowed[i] : 0;
  return v;
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// Step price of entering a cell of weight w with additive a over an edge
// price e: min(e*w + a, MAX-1) (negotiated.rs:1171-1173, :2777-2797,
// mr-metal gpu.rs:254-257).  e*w + a < 2^64 for all u32 inputs.
__device__ __forceinline__ uint32_t step_price(uint32_t e, uint32_t w, uint32_t a) {
  uint64_t p = (uint64_t)e * (uint64_t)w + (uint64_t)a;
  return p < 0xFFFFFFFEull ? (uint32_t)p : 0xFFFFFFFEu;
}

// Candidate label through an edge of price s; kInf when the saturating sum
// would reach MAX (dijkstra.rs:303-313 rejects that candidate).
__device__ __forceinline__ uint64_t extend(uint64_t prev, uint32_t s) {
  uint64_t cand = prev + ((((uint64_t)s) << 32) | 1ull);
  return (cand > prev && cand < kLabLimit) ? cand : kInf;
}

// Workgroup-uniform read of an LDS word.  Loop exits and branches around
// __syncthreads() must be uniform for the compiler, not only at run time: a
// value loaded from LDS lives in a vector register, and a loop whose exit
// depends on it may be restructured so that lanes leave it at different
// points (lanes 1..63 re-running the barrier while lane 0 is parked).
// readfirstlane turns the value into a scalar so every such branch stays a
// scalar branch.
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ bool sorted_has(const uint32_t* a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && a[lo] == v;
}

__device__ __forceinline__ uint32_t lower_bound(const uint32_t* a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t add_price(const BoardView& b, const FieldDesc& F, uint32_t g) {
  uint64_t a = b.add ? b.add[g] : 0u;
  if (b.halo) {
    int32_t o = b.halo[g];
    if (o != -1 && o != F.own_id) a += b.halo_pen;
  }
  return a < 0xFFFFFFFFull ? (uint32_t)a : 0xFFFFFFFFu;
}

// Via bits of global cell g for field F, including per-net via exemptions
// (negotiated.rs:3010-3043).
__device__ uint8_t via_bits(const BoardView& b, const FieldDesc& F, const uint32_t* pool, uint32_t g) {
  const uint8_t* va = b.va;
  uint32_t l = g / b.plane;
  const uint32_t* vp = pool + F.vpad_off;
  uint8_t r = b.raw[g];
  bool self_ok = !(r & R_VIANODE) && (!(r & R_VF) || sorted_has(vp, F.n_vpad, g));
  uint8_t bits = 0;
  if (!self_ok) return 0;
  if (l > 0 && va[l - 1]) {
    uint32_t q = g - b.plane;
    uint8_t rq = b.raw[q];
    if (!(rq & R_VIANODE) && (!(rq & R_VF) || sorted_has(vp, F.n_vpad, q))) bits |= K_DN;
  }
  if (l + 1 < b.layers && va[l]) {
    uint32_t q = g + b.plane;
    uint8_t rq = b.raw[q];
    if (!(rq & R_VIANODE) && (!(rq & R_VF) || sorted_has(vp, F.n_vpad, q))) bits |= K_UP;
  }
  return bits;
}

// Complete per-field view of one cell (must match the tile loader exactly).
struct CellInfo {
  uint8_t mask;
  uint32_t w, a;
};

__device__ CellInfo cell_info(const BoardView& b, const FieldDesc& F, const uint32_t* pool, uint32_t g) {
  CellInfo ci;
  uint8_t bm = b.bmask[g];
  uint8_t m = (uint8_t)((bm >> 1) & 0x7E);
  bool trace = bm & M_TRACE, obst = bm & M_OBST;
  bool pad = F.n_pad && obst && sorted_has(pool + F.pad_off, F.n_pad, g);
  bool pass = !trace && (!obst || pad);
  if (!pass && !trace && (F.flags & F_FORCE)) {
    for (uint32_t k = 0; k < F.n_forced; ++k)
      if (pool[F.forced_off + k] == g) pass = true;
  }
  if (pass) m |= K_PASS;
  if (F.n_vpad) m = (uint8_t)((m & ~(K_DN | K_UP)) | via_bits(b, F, pool, g));
  ci.mask = m;
  ci.w = (obst && pad) ? 1u : b.cost[g];
  ci.a = (b.add || b.halo) ? add_price(b, F, g) : 0u;
  return ci;
}

struct RelaxArgs {
  const BoardView* boards;
  const FieldDesc* fields;
  const uint32_t* item_field;
  const uint32_t* pool;
  uint64_t* labels;
  uint32_t* state;
  uint32_t* act_cur;
  uint32_t* act_next;
  const uint32_t* list_cur;
  uint32_t* list_next;
  uint32_t* counters;
  uint32_t* queue;
  uint32_t n_items;
  uint32_t tw, th, max_local;
  uint32_t use_add;
  uint32_t qcap;
  uint32_t spin_limit;
  uint64_t spin_ticks;  // real-time bound of any spin wait (wall_clock64 ticks)
};

// LDS carve-up for one tile (all sizes in bytes, 16-byte aligned sections).
struct Lds {
  uint32_t* d;   // [L][th+2][rs] dist incl. halo ring
  uint32_t* h;   // [L][th+2][rs] hops
  uint32_t* w;   // [L][th][tw+1] enter weight (odd row stride: no bank conflicts)
  uint32_t* a;   // [L][th][tw+1] additive price
  uint8_t* m;    // [L][th][tw+1] mask
  uint32_t* xg;  // [tw+1] x gap prices entering column i from the left
  uint32_t* yg;  // [th+1]
  uint32_t* vg;  // [16]
  uint32_t* misc;
  uint32_t rs, rh, ps;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline size_t lds_bytes(uint32_t tw, uint32_t th, uint32_t L, bool use_add) {
  size_t rs = tw + 3, rh = th + 2, ncell = (size_t)L * th * (tw + 1), nlab = (size_t)L * rh * rs;
  size_t s = 0;
  s += align16(nlab * 4) * 2;
  s += align16(ncell * 4);
  if (use_add) s += align16(ncell * 4);
  s += align16(ncell);
  s += align16((tw + 1) * 4) + align16((th + 1) * 4) + align16(16 * 4) + 64;
  return s;
}

__device__ inline Lds carve(char* base, uint32_t tw, uint32_t th, uint32_t L, bool use_add) {
  Lds s;
  s.rs = tw + 3;
  s.rh = th + 2;
  s.ps = tw + 1;
  size_t ncell = (size_t)L * th * s.ps, nlab = (size_t)L * s.rh * s.rs;
  size_t o = 0;
  s.d = (uint32_t*)(base + o);
  o += align16(nlab * 4);
  s.h = (uint32_t*)(base + o);
  o += align16(nlab * 4);
  s.w = (uint32_t*)(base + o);
  o += align16(ncell * 4);
  if (use_add) {
    s.a = (uint32_t*)(base + o);
    o += align16(ncell * 4);
  } else {
    s.a = nullptr;
  }
  s.m = (uint8_t*)(base + o);
  o += align16(ncell);
  s.xg = (uint32_t*)(base + o);
  o += align16((tw + 1) * 4);
  s.yg = (uint32_t*)(base + o);
  o += align16((th + 1) * 4);
  s.vg = (uint32_t*)(base + o);
  o += align16(16 * 4);
  s.misc = (uint32_t*)(base + o);
  return s;
}

__device__ __forceinline__ uint64_t ld_lab(const Lds& s, uint32_t i) {
  return ((uint64_t)s.d[i] << 32) | s.h[i];
}
__device__ __forceinline__ void st_lab(const Lds& s, uint32_t i, uint64_t v) {
  s.d[i] = (uint32_t)(v >> 32);
  s.h[i] = (uint32_t)v;
}

// Step price of entering LDS cell c over an edge priced e (kMax when the
// edge or the cell is closed: `bits` must all be set in the cell mask).
__device__ __forceinline__ uint32_t gated_step(const Lds& s, uint32_t c, uint8_t bits, uint32_t e, bool use_add) {
  uint8_t m = s.m[c];
  return ((m & bits) == bits) ? step_price(e, s.w[c], use_add ? s.a[c] : 0u) : kMax;
}

// Sweep one line held in registers: forward (index increasing) then backward.
// lab[0] and lab[N+1] are the halo ends; sf[i] / sb[i] are the prices of
// entering cell i from i-1 / i+1.  Returns a bit mask of improved cells.
template <int N>
__device__ __forceinline__ uint64_t sweep_line(uint64_t (&lab)[N + 2], const uint32_t (&sf)[N],
                                               const uint32_t (&sb)[N]) {
  uint64_t chg = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    uint64_t c = extend(lab[i], sf[i]);
    if (c < lab[i + 1]) {
      lab[i + 1] = c;
      chg |= 1ull << i;
    }
  }
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    uint64_t c = extend(lab[i + 2], sb[i]);
    if (c < lab[i + 1]) {
      lab[i + 1] = c;
      chg |= 1ull << i;
    }
  }
  return chg;
}

// In-LDS relaxation of one tile to a local fixpoint: rows, columns and via
// columns are each swept forward and backward with the whole line held in
// registers (one lane per line).  Returns true when the round cap was hit.
template <int TW, int TH, int LM>
__device__ bool relax_local(const Lds& s, uint32_t L, bool use_add, uint32_t max_local, uint32_t& rounds) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  const uint32_t rs = s.rs, rh = s.rh;
  constexpr uint32_t PS = TW + 1;  // padded interior row stride
  constexpr uint32_t plane_p = TH * PS;
  for (rounds = 0; rounds < max_local; ++rounds) {
    bool ch = false;
    // rows (lane per (layer, y)); forward then backward, each with its own
    // step array so only one is live at a time
    for (uint32_t line = tid; line < L * TH; line += nt) {
      uint32_t l = line / TH, y = line - l * TH;
      uint32_t lb = (l * rh + y + 1) * rs, cb = (l * TH + y) * PS;
      uint64_t lab[TW + 2];
      uint32_t st[TW];
      uint64_t chg = 0;
#pragma unroll
      for (int i = 0; i < TW + 2; ++i) lab[i] = ld_lab(s, lb + i);
#pragma unroll
      for (int x = 0; x < TW; ++x) st[x] = gated_step(s, cb + x, K_PASS | K_NX, s.xg[x], use_add);
#pragma unroll
      for (int x = 0; x < TW; ++x) {
        uint64_t c = extend(lab[x], st[x]);
        if (c < lab[x + 1]) {
          lab[x + 1] = c;
          chg |= 1ull << x;
        }
      }
#pragma unroll
      for (int x = 0; x < TW; ++x) st[x] = gated_step(s, cb + x, K_PASS | K_PX, s.xg[x + 1], use_add);
#pragma unroll
      for (int x = TW - 1; x >= 0; --x) {
        uint64_t c = extend(lab[x + 2], st[x]);
        if (c < lab[x + 1]) {
          lab[x + 1] = c;
          chg |= 1ull << x;
        }
      }
      if (chg) {
        ch = true;
#pragma unroll
        for (int x = 0; x < TW; ++x)
          if ((chg >> x) & 1) {
            st_lab(s, lb + x + 1, lab[x + 1]);
            s.m[cb + x] |= K_CHG;
          }
      }
    }
    __syncthreads();
    // columns (lane per (layer, x))
    for (uint32_t line = tid; line < L * TW; line += nt) {
      uint32_t l = line / TW, x = line - l * TW;
      uint32_t lb = l * rh * rs + x + 1, cb = l * plane_p + x;
      uint64_t lab[TH + 2];
      uint32_t st[TH];
      uint64_t chg = 0;
#pragma unroll
      for (int i = 0; i < TH + 2; ++i) lab[i] = ld_lab(s, lb + i * rs);
#pragma unroll
      for (int y = 0; y < TH; ++y) st[y] = gated_step(s, cb + y * PS, K_PASS | K_NY, s.yg[y], use_add);
#pragma unroll
      for (int y = 0; y < TH; ++y) {
        uint64_t c = extend(lab[y], st[y]);
        if (c < lab[y + 1]) {
          lab[y + 1] = c;
          chg |= 1ull << y;
        }
      }
#pragma unroll
      for (int y = 0; y < TH; ++y) st[y] = gated_step(s, cb + y * PS, K_PASS | K_PY, s.yg[y + 1], use_add);
#pragma unroll
      for (int y = TH - 1; y >= 0; --y) {
        uint64_t c = extend(lab[y + 2], st[y]);
        if (c < lab[y + 1]) {
          lab[y + 1] = c;
          chg |= 1ull << y;
        }
      }
      if (chg) {
        ch = true;
#pragma unroll
        for (int y = 0; y < TH; ++y)
          if ((chg >> y) & 1) {
            st_lab(s, lb + (y + 1) * rs, lab[y + 1]);
            s.m[cb + y * PS] |= K_CHG;
          }
      }
    }
    __syncthreads();
    // via columns (lane per (y, x)); at most LM layers
    if (L > 1) {
      for (uint32_t col = tid; col < TW * TH; col += nt) {
        uint32_t y = col / TW, x = col - y * TW;
        uint32_t lb = (y + 1) * rs + x + 1, cb = y * PS + x;
        uint64_t lab[LM + 2];
        uint32_t sf[LM], sb[LM];
#pragma unroll
        for (int l = 0; l < LM + 2; ++l) lab[l] = kInf;
#pragma unroll
        for (int l = 0; l < LM; ++l) {
          sf[l] = kMax;
          sb[l] = kMax;
          if ((uint32_t)l < L) {
            uint32_t c = l * plane_p + cb;
            lab[l + 1] = ld_lab(s, l * rh * rs + lb);
            if (l > 0) sf[l] = gated_step(s, c, K_PASS | K_DN, s.vg[l - 1], use_add);
            if ((uint32_t)l + 1 < L) sb[l] = gated_step(s, c, K_PASS | K_UP, s.vg[l], use_add);
          }
        }
        uint64_t chg = sweep_line<LM>(lab, sf, sb);
        if (chg) {
          ch = true;
#pragma unroll
          for (int l = 0; l < LM; ++l)
            if ((chg >> l) & 1) {
              st_lab(s, l * rh * rs + lb, lab[l + 1]);
              s.m[l * plane_p + cb] |= K_CHG;
            }
        }
      }
    }
    if (ch) s.misc[2] = rounds + 1;
    __syncthreads();
    if (uniform_u32(s.misc[2]) != rounds + 1) {
      ++rounds;
      return false;
    }
  }
  return true;
}

// Result flags of one tile visit.
enum : uint32_t { T_LEFT = 1, T_RIGHT = 2, T_TOP = 4, T_BOTTOM = 8, T_ANY = 16, T_CAPPED = 32 };

// Load, relax to an in-LDS fixpoint, and write back one (field, tile) item.
// Returns T_* flags (valid in every thread).
template <int TW, int TH, int LM>
__device__ uint32_t process_tile(const RelaxArgs& A, uint32_t item, char* smem) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  const uint32_t f = A.item_field[item];
  const FieldDesc F = A.fields[f];
  const BoardView& B = A.boards[F.board];
  const uint32_t t = item - F.item0;
  const uint32_t txi = t % F.tx, tyi = t / F.tx;
  constexpr uint32_t tw = TW, th = TH;
  const uint32_t L = B.layers;
  const uint32_t ncell = L * th * tw, plane_t = th * tw;
  const uint32_t wx0 = txi * tw, wy0 = tyi * th;
  const bool use_add = A.use_add != 0;  // LDS holds an additive plane in this launch
  const bool has_add = B.add != nullptr || B.halo != nullptr;
  Lds s = carve(smem, tw, th, L, use_add);
  const uint32_t rs = s.rs, rh = s.rh;
  constexpr uint32_t PS = TW + 1;  // padded interior row stride (mask / weight / additive)
  const uint64_t blockbase = F.lab_off + (uint64_t)t * ncell;

  if (tid == 0) {
    s.misc[0] = __hip_atomic_load(&A.state[item], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s.misc[1] = 0;  // result flags
    s.misc[2] = 0;  // round changed
  }
  // tile edge prices
  for (uint32_t i = tid; i <= tw; i += nt) {
    uint32_t gx = F.x0 + wx0 + i;  // entering column gx from gx-1
    s.xg[i] = (gx >= 1 && gx - 1 + 1 < B.w) ? B.xe[gx - 1] : 0u;
  }
  for (uint32_t i = tid; i <= th; i += nt) {
    uint32_t gy = F.y0 + wy0 + i;
    s.yg[i] = (gy >= 1 && gy - 1 + 1 < B.h) ? B.ye[gy - 1] : 0u;
  }
  for (uint32_t i = tid; i < 16; i += nt) s.vg[i] = (i + 1 < L) ? B.ve[i] : 0u;
  __syncthreads();
  const bool touched = (uniform_u32(s.misc[0]) & S_TOUCHED) != 0;

  // interior cells
  for (uint32_t c = tid; c < ncell; c += nt) {
    uint32_t l = c / plane_t, r = c - l * plane_t, y = r / tw, x = r - y * tw;
    uint32_t wx = wx0 + x, wy = wy0 + y;
    uint32_t li = (l * rh + y + 1) * rs + x + 1;
    uint8_t m = 0;
    uint32_t wgt = 0, ap = 0;
    uint64_t lab = kInf;
    if (wx < F.ww && wy < F.wh) {
      uint32_t g = (l * B.h + F.y0 + wy) * B.w + F.x0 + wx;
      uint8_t bm = B.bmask[g];
      m = (uint8_t)((bm >> 1) & 0x7E);
      if (!(bm & (M_TRACE | M_OBST))) m |= K_PASS;
      wgt = B.cost[g];
      if (has_add) ap = add_price(B, F, g);
      if (touched)
        lab = __hip_atomic_load(&A.labels[blockbase + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    uint32_t pc = (l * th + y) * PS + x;
    s.m[pc] = m;
    s.w[pc] = wgt;
    if (use_add) s.a[pc] = ap;
    st_lab(s, li, lab);
  }
  // halo strips (neighbour tiles of the same field)
  for (uint32_t i = tid; i < L * th; i += nt) {
    uint32_t l = i / th, y = i - l * th;
    uint64_t left = kInf, right = kInf;
    if (wy0 + y < F.wh) {
      if (txi > 0) {
        uint32_t nb = item - 1;
        if (__hip_atomic_load(&A.state[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & S_TOUCHED)
          left = __hip_atomic_load(&A.labels[F.lab_off + (uint64_t)(t - 1) * ncell + (l * th + y) * tw + tw - 1],
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (txi + 1 < F.tx) {
        uint32_t nb = item + 1;
        if (__hip_atomic_load(&A.state[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & S_TOUCHED)
          right = __hip_atomic_load(&A.labels[F.lab_off + (uint64_t)(t + 1) * ncell + (l * th + y) * tw],
                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    st_lab(s, (l * rh + y + 1) * rs + 0, left);
    st_lab(s, (l * rh + y + 1) * rs + tw + 1, right);
  }
  for (uint32_t i = tid; i < L * tw; i += nt) {
    uint32_t l = i / tw, x = i - l * tw;
    uint64_t top = kInf, bot = kInf;
    if (wx0 + x < F.ww) {
      if (tyi > 0) {
        uint32_t nb = item - F.tx;
        if (__hip_atomic_load(&A.state[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & S_TOUCHED)
          top = __hip_atomic_load(&A.labels[F.lab_off + (uint64_t)(t - F.tx) * ncell + (l * th + th - 1) * tw + x],
                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (tyi + 1 < F.ty) {
        uint32_t nb = item + F.tx;
        if (__hip_atomic_load(&A.state[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & S_TOUCHED)
          bot = __hip_atomic_load(&A.labels[F.lab_off + (uint64_t)(t + F.tx) * ncell + (l * th) * tw + x],
                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    st_lab(s, (l * rh + 0) * rs + x + 1, top);
    st_lab(s, (l * rh + th + 1) * rs + x + 1, bot);
  }
  __syncthreads();

  // per-net overrides: own pads (weight 1, passable) and forced endpoints
  const uint32_t plane = B.plane;
  if (F.n_pad) {
    const uint32_t* pads = A.pool + F.pad_off;
    for (uint32_t seg = tid; seg < L * th; seg += nt) {
      uint32_t l = seg / th, y = seg - l * th, wy = wy0 + y;
      if (wy >= F.wh) continue;
      uint32_t row0 = (l * B.h + F.y0 + wy) * B.w + F.x0 + wx0;
      uint32_t lim = min(tw, F.ww - wx0);
      for (uint32_t k = lower_bound(pads, F.n_pad, row0); k < F.n_pad && pads[k] < row0 + lim; ++k) {
        uint32_t x = pads[k] - row0, c = (l * th + y) * PS + x;
        uint8_t bm = B.bmask[pads[k]];
        if (!(bm & M_TRACE)) s.m[c] |= K_PASS;
        if (bm & M_OBST) s.w[c] = 1u;
      }
    }
  }
  if (F.flags & F_FORCE) {
    for (uint32_t k = tid; k < F.n_forced; k += nt) {
      uint32_t g = A.pool[F.forced_off + k];
      uint32_t l = g / plane, r = g - l * plane, gy = r / B.w, gx = r - gy * B.w;
      if (gx < F.x0 + wx0 || gy < F.y0 + wy0) continue;
      uint32_t x = gx - F.x0 - wx0, y = gy - F.y0 - wy0;
      if (x >= tw || y >= th || wx0 + x >= F.ww || wy0 + y >= F.wh) continue;
      if (!(B.bmask[g] & M_TRACE)) s.m[(l * th + y) * PS + x] |= K_PASS;
    }
  }
  __syncthreads();
  if (F.n_vpad) {
    const uint32_t* vp = A.pool + F.vpad_off;
    for (uint32_t seg = tid; seg < L * th; seg += nt) {
      uint32_t l = seg / th, y = seg - l * th, wy = wy0 + y;
      if (wy >= F.wh) continue;
      uint32_t row0 = (l * B.h + F.y0 + wy) * B.w + F.x0 + wx0;
      uint32_t lim = min(tw, F.ww - wx0);
      for (uint32_t k = lower_bound(vp, F.n_vpad, row0); k < F.n_vpad && vp[k] < row0 + lim; ++k) {
        uint32_t x = vp[k] - row0;
        for (int dl = -1; dl <= 1; ++dl) {
          int ll = (int)l + dl;
          if (ll < 0 || ll >= (int)L) continue;
          uint32_t g = vp[k] + (uint32_t)((int)plane * dl);
          uint32_t c = ((uint32_t)ll * th + y) * PS + x;
          uint8_t vb = via_bits(B, F, A.pool, g);
          s.m[c] = (uint8_t)((s.m[c] & ~(K_DN | K_UP)) | vb);
        }
      }
    }
    __syncthreads();
  }
  // seeds (first visit only): passable sources start at label (0, 0)
  if (!touched) {
    for (uint32_t k = tid; k < F.n_src; k += nt) {
      uint32_t g = A.pool[F.src_off + k];
      uint32_t l = g / plane, r = g - l * plane, gy = r / B.w, gx = r - gy * B.w;
      if (gx < F.x0 + wx0 || gy < F.y0 + wy0) continue;
      uint32_t x = gx - F.x0 - wx0, y = gy - F.y0 - wy0;
      if (x >= tw || y >= th || wx0 + x >= F.ww || wy0 + y >= F.wh) continue;
      uint32_t c = (l * th + y) * PS + x;
      if (s.m[c] & K_PASS) {
        st_lab(s, (l * rh + y + 1) * rs + x + 1, 0ull);
        s.m[c] |= K_CHG;
      }
    }
    __syncthreads();
  }

  // ---- in-LDS relaxation to a local fixpoint ----------------------------------
  uint32_t rounds = 0;
  bool capped = relax_local<TW, TH, LM>(s, L, use_add, A.max_local, rounds);
  // ---- write back -----------------------------------------------------------
  uint32_t flags = 0;
  for (uint32_t c = tid; c < ncell; c += nt) {
    uint32_t l = c / plane_t, r = c - l * plane_t, y = r / tw, x = r - y * tw;
    uint8_t m = s.m[(l * th + y) * PS + x];
    if (!touched || (m & K_CHG)) {
      uint64_t v = ld_lab(s, (l * rh + y + 1) * rs + x + 1);
      __hip_atomic_store(&A.labels[blockbase + c], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (m & K_CHG) {
      flags |= T_ANY;
      if (x == 0) flags |= T_LEFT;
      if (x == tw - 1) flags |= T_RIGHT;
      if (y == 0) flags |= T_TOP;
      if (y == th - 1) flags |= T_BOTTOM;
    }
  }
  if (flags) atomicOr(&s.misc[1], flags);
  __syncthreads();  // also orders this workgroup's label stores before the flag / queue updates
  uint32_t res = uniform_u32(s.misc[1]) | (capped ? T_CAPPED : 0u);
  if (tid == 0) {
    if (!touched) __hip_atomic_fetch_or(&A.state[item], S_TOUCHED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd(&A.counters[(C_VISITS) * kCtrStride], 1u);
    atomicAdd(&A.counters[(C_ROUNDS) * kCtrStride], rounds);
  }
  __syncthreads();
  return res;
}

// Map a workgroup to an item so consecutive items (tiles of one field) stay on
// one XCD: hardware dispatches workgroup b to XCD b % 8.
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t n) {
  uint32_t per = (n + 7) / 8;
  uint32_t xcd = b % 8, k = b / 8;
  return xcd * per + k;
}

// Activate `nb` for the next round once (flag + append to the next list).
__device__ __forceinline__ void activate(const RelaxArgs& A, uint32_t nb) {
  if (atomicExch(&A.act_next[nb], 1u) == 0u) A.list_next[atomicAdd(&A.counters[(C_NEXTN) * kCtrStride], 1u)] = nb;
}

template <int TW, int TH, int LM>
__global__ void __launch_bounds__(256, 2) relax_round(RelaxArgs A) {
  extern __shared__ __align__(16) char smem[];
  uint32_t item = A.list_cur[blockIdx.x];
  uint32_t res = process_tile<TW, TH, LM>(A, item, smem);
  if (threadIdx.x == 0) {
    A.act_cur[item] = 0;
    const FieldDesc F = A.fields[A.item_field[item]];
    uint32_t t = item - F.item0, txi = t % F.tx, tyi = t / F.tx;
    if ((res & T_LEFT) && txi > 0) activate(A, item - 1);
    if ((res & T_RIGHT) && txi + 1 < F.tx) activate(A, item + 1);
    if ((res & T_TOP) && tyi > 0) activate(A, item - F.tx);
    if ((res & T_BOTTOM) && tyi + 1 < F.ty) activate(A, item + F.tx);
    if (res & T_CAPPED) activate(A, item);
    if (res & (T_ANY | T_CAPPED)) atomicAdd(&A.counters[(C_CHANGED) * kCtrStride], 1u);
  }
}

// ---------------------------------------------------------------------------
// persistent form: device-side rounds over a work list
// ---------------------------------------------------------------------------
//
// All workgroups stay resident.  A round pops items from the current active
// list with one atomicAdd per item (a device work queue), processes them and
// appends activated neighbours to the next list (deduplicated by flags).  A
// grid barrier with agent-scope fences separates rounds, so every round reads
// labels written by earlier rounds exactly as the host-driven form does, and
// the loop ends when a round activates nothing.  Per-round counters are
// triple-buffered so a slow workgroup can still read round r's counters while
// round r+1 runs.  Every wait is bounded by spin_limit (error flag + exit).

__device__ bool grid_barrier(const RelaxArgs& A, uint32_t nblocks, uint32_t& my_gen) {
  __syncthreads();
  bool ok = true;
  if (threadIdx.x == 0) {
    __threadfence();  // release this workgroup's label stores
    uint32_t g = my_gen;
    if (atomicAdd(&A.counters[(C_BAR_ARRIVE) * kCtrStride], 1u) == nblocks - 1) {
      atomicExch(&A.counters[(C_BAR_ARRIVE) * kCtrStride], 0u);
      __hip_atomic_fetch_add(&A.counters[(C_BAR_GEN) * kCtrStride], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      // bounded by the real-time clock, with exponential backoff
      uint64_t t0 = wall_clock64();
      uint32_t naps = 1;
      while (__hip_atomic_load(&A.counters[(C_BAR_GEN) * kCtrStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (wall_clock64() - t0 > A.spin_ticks) {
          atomicOr(&A.counters[(C_ERROR) * kCtrStride], 8u);
          ok = false;
          break;
        }
        for (uint32_t i = 0; i < naps; ++i) __builtin_amdgcn_s_sleep(8);
        naps = naps < 16 ? naps * 2 : 16;
      }
    }
    __threadfence();  // acquire labels written before the barrier
    my_gen = g + 1;
  }
  __shared__ uint32_t s_ok;
  if (threadIdx.x == 0) s_ok = ok ? 1u : 0u;
  __syncthreads();
  return uniform_u32(s_ok) != 0;
}

template <int TW, int TH, int LM>
__global__ void __launch_bounds__(256, 2) relax_persistent(RelaxArgs A, uint32_t n_first) {
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t s_k, s_n;
  RelaxArgs R = A;
  uint32_t my_gen = 0;
  uint32_t n_cur = n_first;
  for (uint32_t round = 0;; ++round) {
    uint32_t* taken = &A.counters[(C_TAKEN0 + round % 3) * kCtrStride];
    uint32_t* nextn = &A.counters[(C_NEXT0 + round % 3) * kCtrStride];
    R.counters = A.counters;
    // pop / process this round's list
    for (;;) {
      if (threadIdx.x == 0) s_k = atomicAdd(taken, 1u);
      __syncthreads();
      uint32_t k = uniform_u32(s_k);
      __syncthreads();
      if (k >= n_cur) break;
      uint32_t item = uniform_u32(R.list_cur[k]);
      uint32_t res = process_tile<TW, TH, LM>(R, item, smem);
      if (threadIdx.x == 0) {
        R.act_cur[item] = 0;
        const FieldDesc F = R.fields[R.item_field[item]];
        uint32_t t = item - F.item0, txi = t % F.tx, tyi = t / F.tx;
        auto act = [&](uint32_t nb) {
          if (atomicExch(&R.act_next[nb], 1u) == 0u) R.list_next[atomicAdd(nextn, 1u)] = nb;
        };
        if ((res & T_LEFT) && txi > 0) act(item - 1);
        if ((res & T_RIGHT) && txi + 1 < F.tx) act(item + 1);
        if ((res & T_TOP) && tyi > 0) act(item - F.tx);
        if ((res & T_BOTTOM) && tyi + 1 < F.ty) act(item + F.tx);
        if (res & T_CAPPED) act(item);
      }
    }
    // the counters of round + 1 were last used in round - 2: clear them now
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      atomicExch(&A.counters[(C_TAKEN0 + (round + 1) % 3) * kCtrStride], 0u);
      atomicExch(&A.counters[(C_NEXT0 + (round + 1) % 3) * kCtrStride], 0u);
    }
    if (!grid_barrier(A, gridDim.x, my_gen)) return;
    if (threadIdx.x == 0)
      s_n = __hip_atomic_load(nextn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    n_cur = uniform_u32(s_n);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&A.counters[(C_PROUNDS) * kCtrStride], 1u);
    if (n_cur == 0 ||
        uniform_u32(__hip_atomic_load(&A.counters[(C_ERROR) * kCtrStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
      return;
    uint32_t* ta = R.act_cur;
    R.act_cur = R.act_next;
    R.act_next = ta;
    const uint32_t* tl = R.list_cur;
    R.list_cur = R.list_next;
    R.list_next = const_cast<uint32_t*>(tl);
  }
}

// ---------------------------------------------------------------------------
// graph form: one grid-stride launch per round, rounds replayed from a hipGraph
// ---------------------------------------------------------------------------
//
// A launch reads the round index and list length that the previous
// advance_round wrote, pops items from the round's list with one atomicAdd
// per item (the device work queue) and appends activated neighbours to the
// other list.  Kernel boundaries separate rounds, so every round reads
// coherent labels; no workgroup ever waits on another.  A launch whose list is
// empty returns at once, so a captured sequence can overshoot the fixpoint.

template <int TW, int TH, int LM>
__global__ void __launch_bounds__(256, 2) relax_list(RelaxArgs E, RelaxArgs O) {
  extern __shared__ __align__(16) char smem[];
  const uint32_t r =
      uniform_u32(__hip_atomic_load(&E.counters[C_GROUND * kCtrStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  const uint32_t n_cur =
      uniform_u32(__hip_atomic_load(&E.counters[C_CURN * kCtrStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  // E holds the even-round list/flag pointers, O the same pointers swapped;
  // both stay kernel arguments (no private copy).
  const RelaxArgs& R = (r & 1) ? O : E;
  // Grid-stride loop: the index is a workgroup-uniform scalar, so the loop
  // (which contains __syncthreads) stays convergent.  A pop loop of the form
  // `if (tid == 0) k = atomicAdd(..); __syncthreads(); ..` is restructured by
  // the compiler so lanes other than 0 re-run the body without a new pop.
  for (uint32_t k = blockIdx.x; k < n_cur; k += gridDim.x) {
    uint32_t item = uniform_u32(R.list_cur[k]);
    uint32_t res = process_tile<TW, TH, LM>(R, item, smem);
    if (threadIdx.x == 0) {
      R.act_cur[item] = 0;
      const FieldDesc F = R.fields[R.item_field[item]];
      uint32_t t = item - F.item0, txi = t % F.tx, tyi = t / F.tx;
      if ((res & T_LEFT) && txi > 0) activate(R, item - 1);
      if ((res & T_RIGHT) && txi + 1 < F.tx) activate(R, item + 1);
      if ((res & T_TOP) && tyi > 0) activate(R, item - F.tx);
      if ((res & T_BOTTOM) && tyi + 1 < F.ty) activate(R, item + F.tx);
      if (res & T_CAPPED) activate(R, item);
    }
    __syncthreads();
  }
}

// Single-thread round step between two relax_list launches.
__global__ void advance_round(uint32_t* ctr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t n = ctr[C_NEXTN * kCtrStride];
  if (ctr[C_CURN * kCtrStride]) ctr[C_PROUNDS * kCtrStride] += 1;
  ctr[C_CURN * kCtrStride] = n;
  ctr[C_NEXTN * kCtrStride] = 0;
  ctr[C_TAKEN0 * kCtrStride] = 0;
  ctr[C_GROUND * kCtrStride] += 1;
}

__global__ void fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void fill_u8(uint8_t* p, uint8_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Activate items idx[0..n) for the first round: flags and list.
__global__ void seed_list(uint32_t* act, uint32_t* list, const uint32_t* idx, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    act[idx[i]] = 1u;
    list[i] = idx[i];
  }
}

// Activate every item: flags and the identity list.
__global__ void all_list(uint32_t* act, uint32_t* list, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    act[i] = 1u;
    list[i] = i;
  }
}

// ---------------------------------------------------------------------------
// field access + reconstruction
// ---------------------------------------------------------------------------

struct FieldRead {
  const BoardView* boards;
  const FieldDesc* fields;
  const uint64_t* labels;
  const uint32_t* state;
  uint32_t tw, th;
};

// Label of global cell g in field F (kInf outside the window or untouched tile).
__device__ uint64_t field_label(const FieldRead& R, const BoardView& B, const FieldDesc& F, uint32_t g) {
  uint32_t l = g / B.plane, r = g - l * B.plane, gy = r / B.w, gx = r - gy * B.w;
  if (gx < F.x0 || gy < F.y0) return kInf;
  uint32_t wx = gx - F.x0, wy = gy - F.y0;
  if (wx >= F.ww || wy >= F.wh) return kInf;
  uint32_t txi = wx / R.tw, tyi = wy / R.th, t = tyi * F.tx + txi;
  if (!(R.state[F.item0 + t] & S_TOUCHED)) return kInf;
  uint32_t x = wx - txi * R.tw, y = wy - tyi * R.th;
  uint64_t ncell = (uint64_t)B.layers * R.th * R.tw;
  return R.labels[F.lab_off + t * ncell + (l * R.th + y) * R.tw + x];
}

__global__ void path_lengths(FieldRead R, uint32_t n, uint32_t* plen, uint32_t* pcost) {
  uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const FieldDesc F = R.fields[f];
  const BoardView B = R.boards[F.board];
  uint32_t len = 0, cost = kMax;
  if (F.dst != kMax) {
    uint64_t lab = field_label(R, B, F, F.dst);
    if ((uint32_t)(lab >> 32) != kMax) {
      len = (uint32_t)lab + 1;
      cost = (uint32_t)(lab >> 32);
    }
  }
  plen[f] = len;
  pcost[f] = cost;
}

struct WalkArgs {
  FieldRead R;
  const uint32_t* pool;
  const uint64_t* poff;
  const uint32_t* plen;
  uint32_t* cells;
  uint32_t* counters;
  uint32_t n;
};

// One wavefront per field: lanes 0..5 test the six candidate predecessors in
// ascending CellIdx order (lower layer, up, left, right, down, upper layer).
__global__ void walk_paths(WalkArgs W) {
  uint32_t f = blockIdx.x;
  uint32_t lane = threadIdx.x;
  if (f >= W.n || W.plen[f] == 0) return;
  const FieldDesc F = W.R.fields[f];
  const BoardView B = W.R.boards[F.board];
  uint32_t* out = W.cells + W.poff[f];
  uint32_t cur = F.dst;
  uint64_t lab = field_label(W.R, B, F, cur);
  for (uint32_t guard = 0; guard <= W.plen[f]; ++guard) {
    uint32_t hops = (uint32_t)lab;
    if (lane == 0) out[hops] = cur;
    if (hops == 0) return;
    CellInfo ci = cell_info(B, F, W.pool, cur);
    uint32_t l = cur / B.plane, r = cur - l * B.plane, y = r / B.w, x = r - y * B.w;
    bool ok = false;
    uint32_t p = 0, e = 0;
    uint8_t bit = 0;
    switch (lane) {
      case 0: ok = l > 0; p = cur - B.plane; e = ok ? B.ve[l - 1] : 0; bit = K_DN; break;
      case 1: ok = y > 0; p = cur - B.w; e = ok ? B.ye[y - 1] : 0; bit = K_NY; break;
      case 2: ok = x > 0; p = cur - 1; e = ok ? B.xe[x - 1] : 0; bit = K_NX; break;
      case 3: ok = x + 1 < B.w; p = cur + 1; e = ok ? B.xe[x] : 0; bit = K_PX; break;
      case 4: ok = y + 1 < B.h; p = cur + B.w; e = ok ? B.ye[y] : 0; bit = K_PY; break;
      case 5: ok = l + 1 < B.layers; p = cur + B.plane; e = ok ? B.ve[l] : 0; bit = K_UP; break;
      default: ok = false;
    }
    ok = ok && (ci.mask & bit);
    if (ok) {
      uint64_t lp = field_label(W.R, B, F, p);
      ok = (uint32_t)(lp >> 32) != kMax && extend(lp, step_price(e, ci.w, ci.a)) == lab;
      if (ok) lab = lp;  // only meaningful in the chosen lane
    }
    uint64_t mask = __ballot(ok);
    if (mask == 0) {
      if (lane == 0) atomicOr(&W.counters[(C_ERROR) * kCtrStride], 16u);
      return;
    }
    uint32_t win = __ffsll((long long)mask) - 1;
    cur = __shfl(p, win);
    uint32_t lo = __shfl((uint32_t)lab, win), hi = __shfl((uint32_t)(lab >> 32), win);
    lab = ((uint64_t)hi << 32) | lo;
  }
  if (lane == 0) atomicOr(&W.counters[(C_ERROR) * kCtrStride], 32u);
}

// Tile-major labels of one field -> window-major u32 dist (and hops).
__global__ void extract_field(FieldRead R, uint32_t f, uint32_t* dist, uint32_t* hops) {
  const FieldDesc F = R.fields[f];
  const BoardView B = R.boards[F.board];
  size_t n = (size_t)F.ww * F.wh * B.layers;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint32_t l = (uint32_t)(i / ((size_t)F.ww * F.wh));
    uint32_t r = (uint32_t)(i - (size_t)l * F.ww * F.wh);
    uint32_t wy = r / F.ww, wx = r - wy * F.ww;
    uint32_t g = (l * B.h + F.y0 + wy) * B.w + F.x0 + wx;
    uint64_t lab = field_label(R, B, F, g);
    dist[i] = (uint32_t)(lab >> 32);
    if (hops) hops[i] = (uint32_t)(lab >> 32) == kMax ? kMax : (uint32_t)lab;
  }
}

// ---------------------------------------------------------------------------
// host engine
// ---------------------------------------------------------------------------

template <typename T>
static void grow(T*& p, size_t& cap, size_t need) {
  if (need <= cap) return;
  if (p) HIP_OK(hipFree(p));
  p = nullptr;
  size_t n = std::max(need, cap + cap / 2);
  HIP_OK(hipMalloc((void**)&p, n * sizeof(T)));
  cap = n;
}

template <typename T>
static T* dev_copy(const T* h, size_t n, hipStream_t s) {
  T* d = nullptr;
  HIP_OK(hipMalloc((void**)&d, std::max<size_t>(n, 1) * sizeof(T) + 16));
  if (n) HIP_OK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

// Instantiated tile shapes (width x height) and layer bounds.
struct KernelSet {
  uint32_t tw, th, lm;
  const void* round;
  const void* persistent;
  const void* list;
};

template <int TW, int TH, int LM>
static KernelSet make_set() {
  return KernelSet{TW, TH, LM, (const void*)relax_round<TW, TH, LM>, (const void*)relax_persistent<TW, TH, LM>,
                   (const void*)relax_list<TW, TH, LM>};
}

template <int TW, int TH>
static void add_shape(std::vector<KernelSet>& v) {
  v.push_back(make_set<TW, TH, 2>());
  v.push_back(make_set<TW, TH, 4>());
  v.push_back(make_set<TW, TH, 8>());
  v.push_back(make_set<TW, TH, 16>());
}

static const std::vector<KernelSet>& kernel_sets() {
  static const std::vector<KernelSet> v = [] {
    std::vector<KernelSet> k;
    add_shape<32, 32>(k);
    add_shape<32, 16>(k);
    add_shape<16, 16>(k);
    add_shape<16, 8>(k);
    add_shape<8, 4>(k);
    add_shape<4, 1>(k);
    return k;
  }();
  return v;
}

bool tile_shape_supported(uint32_t tw, uint32_t th) {
  for (const KernelSet& k : kernel_sets())
    if (k.tw == tw && k.th == th) return true;
  return false;
}

static const KernelSet& pick_kernels(uint32_t tw, uint32_t th, uint32_t layers) {
  uint32_t lm = layers <= 2 ? 2 : layers <= 4 ? 4 : layers <= 8 ? 8 : 16;
  for (const KernelSet& k : kernel_sets())
    if (k.tw == tw && k.th == th && k.lm == lm) return k;
  throw Error(1, "tile shape " + std::to_string(tw) + "x" + std::to_string(th) + " is not instantiated");
}

void ctx_init(Ctx& c, int device) {
  int n = 0;
  HIP_OK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) throw Error(1, "no such HIP device");
  c.device = device;
  HIP_OK(hipSetDevice(device));
  HIP_OK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, device));
  c.cus = prop.multiProcessorCount;
  int rate_khz = 0;
  if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || rate_khz <= 0)
    rate_khz = 100000;
  c.spin_ticks = (uint64_t)rate_khz * 1000ull * 4ull;  // 4 s
  HIP_OK(hipEventCreate(&c.ws.ev0));
  HIP_OK(hipEventCreate(&c.ws.ev1));
  HIP_OK(hipEventCreate(&c.ws.ev2));
  size_t total = prop.totalGlobalMem;
  c.label_budget = std::min<size_t>(c.label_budget, total / 2);
  HIP_OK(hipMalloc((void**)&c.ws.counters, kCtrWords * sizeof(uint32_t)));
  HIP_OK(hipHostMalloc((void**)&c.ws.pinned, kCtrWords * sizeof(uint32_t), hipHostMallocDefault));
  for (const KernelSet& k : kernel_sets()) {
    HIP_OK(hipFuncSetAttribute(k.round, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64));
    HIP_OK(hipFuncSetAttribute(k.persistent, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64));
    HIP_OK(hipFuncSetAttribute(k.list, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64));
  }
}

void ctx_free(Ctx& c) {
  Workspace& w = c.ws;
  void* ptrs[] = {w.labels, w.state, w.act, w.list, w.queue, w.fields, w.item_field, w.pool, w.counters,
                  w.plen, w.poff, w.pcost, w.pcells, w.fout, w.boards};
  (void)hipSetDevice(c.device);
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (w.pinned) (void)hipHostFree(w.pinned);
  if (c.graph.exec) (void)hipGraphExecDestroy(c.graph.exec);
  c.graph = GraphCache();
  if (w.ev0) (void)hipEventDestroy(w.ev0);
  if (w.ev1) (void)hipEventDestroy(w.ev1);
  if (w.ev2) (void)hipEventDestroy(w.ev2);
  w = Workspace();
  if (c.stream) (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

void board_free(DevBoard& b) {
  void* ptrs[] = {b.cost, b.add, b.halo, b.bmask, b.raw, b.xe, b.ye, b.ve};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  b = DevBoard();
}

void board_upload(Ctx& c, DevBoard& b, uint32_t w, uint32_t h, uint32_t layers, const uint32_t* cost,
                  const uint8_t* via_forbidden, const uint8_t* board, const uint32_t* xe,
                  const uint32_t* ye, const uint32_t* ve, const uint8_t* via_allowed,
                  const uint32_t* add, const int32_t* halo, uint32_t halo_pen) {
  if (layers > 16) throw Error(1, "at most 16 layers are supported");
  HIP_OK(hipSetDevice(c.device));
  size_t n = (size_t)w * h * layers;
  uint32_t plane = w * h;
  std::vector<uint8_t> bm(n), raw(n);
  // Directed-edge bitset: mr-core is_board_planar_step_forbidden (lib.rs:466-496)
  // is symmetric, so each cell stores whether each of its four planar edges is
  // open; via bits fold ViaModel legality, BOARD_VIA_NODE and via_forbidden.
  for (uint32_t l = 0; l < layers; ++l)
    for (uint32_t y = 0; y < h; ++y)
      for (uint32_t x = 0; x < w; ++x) {
        size_t i = ((size_t)l * h + y) * w + x;
        uint8_t bc = board ? board[i] : 0;
        uint8_t m = 0;
        if (bc & 0x01) m |= M_TRACE;
        if (cost[i] == kMax) m |= M_OBST;
        auto open = [&](size_t u, size_t v, uint8_t sb, uint8_t db) {
          if (!board) return true;
          return !((board[u] & sb) || (board[v] & db));
        };
        if (y > 0 && open(i, i - w, 0x04, 0x20)) m |= M_NY;
        if (x > 0 && open(i, i - 1, 0x08, 0x10)) m |= M_NX;
        if (x + 1 < w && open(i, i + 1, 0x10, 0x08)) m |= M_PX;
        if (y + 1 < h && open(i, i + w, 0x20, 0x04)) m |= M_PY;
        uint8_t r = 0;
        if (bc & 0x02) r |= R_VIANODE;
        if (via_forbidden && via_forbidden[i]) r |= R_VF;
        raw[i] = r;
        bm[i] = m;
      }
  for (uint32_t l = 0; l < layers; ++l)
    for (uint32_t k = 0; k < plane; ++k) {
      size_t i = (size_t)l * plane + k;
      auto ok = [&](size_t q) { return !(raw[q] & (R_VIANODE | R_VF)); };
      if (l > 0 && via_allowed[l - 1] && ok(i) && ok(i - plane)) bm[i] |= M_DN;
      if (l + 1 < layers && via_allowed[l] && ok(i) && ok(i + plane)) bm[i] |= M_UP;
    }
  board_free(b);
  b.w = w;
  b.h = h;
  b.layers = layers;
  b.cost = dev_copy(cost, n, c.stream);
  b.bmask = dev_copy(bm.data(), n, c.stream);
  b.raw = dev_copy(raw.data(), n, c.stream);
  std::vector<uint32_t> x(w, 0), y(h, 0), v(16, 0);
  for (uint32_t k = 0; k + 1 < w; ++k) x[k] = xe[k];
  for (uint32_t k = 0; k + 1 < h; ++k) y[k] = ye[k];
  b.via_allowed.assign(16, 0);
  for (uint32_t k = 0; k + 1 < layers; ++k) {
    v[k] = ve ? ve[k] : 0;
    b.via_allowed[k] = via_allowed[k] ? 1 : 0;
  }
  b.xe = dev_copy(x.data(), x.size(), c.stream);
  b.ye = dev_copy(y.data(), y.size(), c.stream);
  b.ve = dev_copy(v.data(), v.size(), c.stream);
  b.bytes = n * (4 + 1 + 1);
  if (add) {
    b.add = dev_copy(add, n, c.stream);
    b.bytes += n * 4;
  }
  if (halo) {
    b.halo = dev_copy(halo, n, c.stream);
    b.bytes += n * 4;
  }
  b.halo_pen = halo_pen;
  HIP_OK(hipStreamSynchronize(c.stream));
}

size_t field_label_words(const TileShape& s, uint32_t ww, uint32_t wh, uint32_t layers) {
  size_t tx = (ww + s.tw - 1) / s.tw, ty = (wh + s.th - 1) / s.th;
  return tx * ty * (size_t)layers * s.th * s.tw;
}

static RelaxArgs make_args(Ctx& c, uint32_t n_items, bool use_add) {
  RelaxArgs A;
  memset(&A, 0, sizeof(A));
  A.boards = c.ws.boards;
  A.fields = c.ws.fields;
  A.item_field = c.ws.item_field;
  A.pool = c.ws.pool;
  A.labels = c.ws.labels;
  A.state = c.ws.state;
  A.act_cur = c.ws.act;
  A.act_next = c.ws.act + n_items;
  A.list_cur = c.ws.list;
  A.list_next = c.ws.list + n_items;
  A.counters = c.ws.counters;
  A.queue = c.ws.queue;
  A.n_items = n_items;
  A.tw = c.shape.tw;
  A.th = c.shape.th;
  A.max_local = c.shape.max_local;
  A.use_add = use_add ? 1 : 0;
  A.qcap = n_items;
  A.spin_limit = 1u << 26;
  A.spin_ticks = c.spin_ticks;
  return A;
}

static void launch_fill32(hipStream_t s, uint32_t* p, uint32_t v, size_t n) {
  if (!n) return;
  uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_u32, dim3(blocks), dim3(256), 0, s, p, v, n);
}
static void launch_fill8(hipStream_t s, uint8_t* p, uint8_t v, size_t n) {
  if (!n) return;
  uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_u8, dim3(blocks), dim3(256), 0, s, p, v, n);
}

// Capture (or reuse) `graph_rounds` rounds of relax_list + advance_round.
static hipGraphExec_t graph_for(Ctx& c, const KernelSet& K, uint32_t grid, size_t lds, const RelaxArgs& A) {
  GraphCache& g = c.graph;
  std::vector<uint8_t> key(sizeof(RelaxArgs) + 32);
  memcpy(key.data(), &A, sizeof(RelaxArgs));
  uint64_t extra[4] = {(uint64_t)(uintptr_t)K.list, grid, lds, (uint64_t)c.shape.threads * 1000 + c.graph_rounds};
  memcpy(key.data() + sizeof(RelaxArgs), extra, sizeof(extra));
  if (g.exec && g.key == key) return g.exec;
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t s = c.stream;
  HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  RelaxArgs e = A, o = A;
  o.act_cur = A.act_next;
  o.act_next = A.act_cur;
  o.list_cur = A.list_next;
  o.list_next = const_cast<uint32_t*>(A.list_cur);
  void* args[] = {&e, &o};
  for (uint32_t r = 0; r < c.graph_rounds; ++r) {
    HIP_OK(hipLaunchKernel(K.list, dim3(grid), dim3(c.shape.threads), args, lds, s));
    hipLaunchKernelGGL(advance_round, dim3(1), dim3(64), 0, s, c.ws.counters);
  }
  HIP_OK(hipStreamEndCapture(s, &graph));
  HIP_OK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
  HIP_OK(hipGraphDestroy(graph));
  g.key = key;
  return g.exec;
}

void solve_fields(Ctx& c, const std::vector<const DevBoard*>& boards, const std::vector<FieldReq>& reqs,
                  SolveStats* st) {
  HIP_OK(hipSetDevice(c.device));
  auto t0 = std::chrono::steady_clock::now();
  const TileShape& S = c.shape;
  uint32_t Lmax = 1;
  bool use_add = false;
  c.views.clear();
  for (const DevBoard* b : boards) {
    Lmax = std::max(Lmax, b->layers);
    use_add = use_add || b->add != nullptr || b->halo != nullptr;
    c.views.push_back(b->view());
  }
  if ((size_t)Lmax * S.th * S.tw > 16384 || S.tw < 4 || S.th < 1)
    throw Error(1, "tile shape too large for LDS");
  size_t lds = lds_bytes(S.tw, S.th, Lmax, use_add);
  if (lds > 160 * 1024 - 64) throw Error(1, "tile shape exceeds 160 KiB of LDS");

  std::vector<FieldDesc>& layout = c.layout;
  layout.clear();
  std::vector<uint32_t> pool, item_field, seeds;
  uint64_t lab_off = 0, field_cells = 0;
  uint32_t item = 0;
  for (size_t i = 0; i < reqs.size(); ++i) {
    const FieldReq& q = reqs[i];
    if (q.board >= boards.size()) throw Error(1, "field names a board outside the batch");
    const DevBoard& b = *boards[q.board];
    const uint32_t L = b.layers;
    FieldDesc F;
    memset(&F, 0, sizeof(F));
    F.x0 = q.x0;
    F.y0 = q.y0;
    F.ww = q.x1 - q.x0 + 1;
    F.wh = q.y1 - q.y0 + 1;
    F.tx = (F.ww + S.tw - 1) / S.tw;
    F.ty = (F.wh + S.th - 1) / S.th;
    F.item0 = item;
    F.lab_off = lab_off;
    auto put = [&](const std::vector<uint32_t>& v, uint32_t& off, uint32_t& n) {
      off = (uint32_t)pool.size();
      n = (uint32_t)v.size();
      pool.insert(pool.end(), v.begin(), v.end());
    };
    put(q.sources, F.src_off, F.n_src);
    put(q.forced, F.forced_off, F.n_forced);
    put(q.pads, F.pad_off, F.n_pad);
    put(q.via_pads, F.vpad_off, F.n_vpad);
    F.dst = q.dst;
    F.own_id = q.own_id;
    F.flags = q.flags;
    F.board = q.board;
    uint32_t nt = F.tx * F.ty;
    for (uint32_t k = 0; k < nt; ++k) item_field.push_back((uint32_t)i);
    // seed tiles: tiles holding a source inside the window
    std::vector<uint32_t> st_items;
    for (uint32_t s : q.sources) {
      uint32_t plane = b.w * b.h, r = s % plane, gy = r / b.w, gx = r % b.w;
      if ((size_t)s >= (size_t)plane * L || gx < F.x0 || gy < F.y0 || gx > q.x1 || gy > q.y1) continue;
      uint32_t t = ((gy - F.y0) / S.th) * F.tx + (gx - F.x0) / S.tw;
      st_items.push_back(item + t);
    }
    std::sort(st_items.begin(), st_items.end());
    st_items.erase(std::unique(st_items.begin(), st_items.end()), st_items.end());
    seeds.insert(seeds.end(), st_items.begin(), st_items.end());
    item += nt;
    lab_off += (uint64_t)nt * L * S.th * S.tw;
    field_cells += (uint64_t)F.ww * F.wh * L;
    layout.push_back(F);
  }
  if (lab_off * 8 > c.label_budget) throw Error(1, "label working set exceeds the device budget");

  Workspace& W = c.ws;
  grow(W.labels, W.labels_cap, std::max<size_t>(lab_off, 1));
  grow(W.state, W.state_cap, std::max<size_t>(item, 1));
  grow(W.queue, W.queue_cap, std::max<size_t>(item, 1));
  grow(W.act, W.act_cap, std::max<size_t>((size_t)item * 2, 2));
  grow(W.list, W.list_cap, std::max<size_t>((size_t)item * 2, 2));
  grow(W.fields, W.fields_cap, std::max<size_t>(reqs.size(), 1));
  grow(W.item_field, W.item_field_cap, std::max<size_t>(item, 1));
  grow(W.pool, W.pool_cap, std::max<size_t>(pool.size() + seeds.size(), 1));
  grow(W.boards, W.boards_cap, std::max<size_t>(c.views.size(), 1));
  hipStream_t s = c.stream;
  HIP_OK(hipMemcpyAsync(W.boards, c.views.data(), c.views.size() * sizeof(BoardView), hipMemcpyHostToDevice, s));
  if (!layout.empty())
    HIP_OK(hipMemcpyAsync(W.fields, layout.data(), layout.size() * sizeof(FieldDesc),
                          hipMemcpyHostToDevice, s));
  if (item)
    HIP_OK(hipMemcpyAsync(W.item_field, item_field.data(), item * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  std::vector<uint32_t> pool_all = pool;
  pool_all.insert(pool_all.end(), seeds.begin(), seeds.end());
  if (!pool_all.empty())
    HIP_OK(hipMemcpyAsync(W.pool, pool_all.data(), pool_all.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  const uint32_t* d_seeds = W.pool + pool.size();
  launch_fill32(s, W.state, 0u, item);
  launch_fill32(s, W.counters, 0u, kCtrWords);

  RelaxArgs A = make_args(c, item, use_add);
  SolveStats local;
  local.field_cells = field_cells;
  bool timed = false;
  if (item && !seeds.empty()) {
    const KernelSet& K = pick_kernels(S.tw, S.th, Lmax);
    uint32_t* pinned = W.pinned;
    void* args[] = {&A};
    // Rounds of relax_round over the active-item list until no item is
    // activated.  Every round starts at a kernel boundary, so it reads
    // coherent labels; an empty next list certifies the fixpoint.
    auto plain_rounds = [&](uint32_t& rounds, uint32_t n_active) {
      while (n_active) {
        launch_fill32(s, W.counters + C_NEXTN * kCtrStride, 0u, 1);
        HIP_OK(hipLaunchKernel(K.round, dim3(n_active), dim3(S.threads), args, lds, s));
        HIP_OK(hipMemcpyAsync(pinned, W.counters, kCtrWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        ++rounds;
        if (pinned[C_ERROR * kCtrStride]) break;
        n_active = pinned[C_NEXTN * kCtrStride];
        std::swap(A.act_cur, A.act_next);
        uint32_t* lc = const_cast<uint32_t*>(A.list_cur);
        A.list_cur = A.list_next;
        A.list_next = lc;
        if (rounds > 1000000) throw Error(1, "relaxation did not converge");
      }
    };
    uint32_t rounds = 0;
    launch_fill32(s, W.act, 0u, (size_t)item * 2);
    if (c.mode == MODE_PLAIN) {
      hipLaunchKernelGGL(seed_list, dim3(((uint32_t)seeds.size() + 255) / 256), dim3(256), 0, s, W.act, W.list,
                         d_seeds, (uint32_t)seeds.size());
      HIP_OK(hipEventRecord(W.ev0, s));
      timed = true;
      plain_rounds(rounds, (uint32_t)seeds.size());
    } else if (c.mode == MODE_GRAPH) {
      hipLaunchKernelGGL(seed_list, dim3(((uint32_t)seeds.size() + 255) / 256), dim3(256), 0, s, W.act, W.list,
                         d_seeds, (uint32_t)seeds.size());
      launch_fill32(s, W.counters + C_CURN * kCtrStride, (uint32_t)seeds.size(), 1);
      int occ = 0;
      HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, K.list, S.threads, lds));
      if (occ < 1) throw Error(1, "relaxation kernel does not fit on a CU");
      uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)occ * c.cus, item));
      hipGraphExec_t exec = graph_for(c, K, grid, lds, A);
      HIP_OK(hipEventRecord(W.ev0, s));
      timed = true;
      uint32_t replays = 0;
      for (;;) {
        HIP_OK(hipGraphLaunch(exec, s));
        HIP_OK(hipMemcpyAsync(pinned, W.counters, kCtrWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        ++replays;
        if (pinned[C_ERROR * kCtrStride] || pinned[C_CURN * kCtrStride] == 0) break;
        if (replays > 1000000) throw Error(1, "relaxation did not converge");
      }
      rounds = pinned[C_PROUNDS * kCtrStride];
      local.replays = replays;
    } else {
      throw Error(1, "the persistent grid-barrier form is disabled on this build (it did not return on MI355X); "
                     "use the plain or graph form");
    }
    local.rounds = rounds;
  }
  if (timed) HIP_OK(hipEventRecord(W.ev1, s));
  uint32_t* cnt = W.pinned;
  HIP_OK(hipMemcpyAsync(cnt, W.counters, kCtrWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (cnt[C_ERROR * kCtrStride]) throw Error(1, "device relaxation reported error flags " + std::to_string(cnt[C_ERROR * kCtrStride]));
  float ms = 0;
  if (timed) HIP_OK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
  local.relax_ms = ms;
  local.tile_visits = cnt[C_VISITS * kCtrStride];
  local.local_rounds = cnt[C_ROUNDS * kCtrStride];
  local.cells_visited = (uint64_t)cnt[C_VISITS * kCtrStride] * Lmax * S.th * S.tw;
  local.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (st) *st = local;
  c.last = local;
}

static FieldRead make_read(Ctx& c) {
  FieldRead R;
  R.boards = c.ws.boards;
  R.fields = c.ws.fields;
  R.labels = c.ws.labels;
  R.state = c.ws.state;
  R.tw = c.shape.tw;
  R.th = c.shape.th;
  return R;
}

void reconstruct_paths(Ctx& c, size_t n, std::vector<uint8_t>& routed,
                       std::vector<uint32_t>& cost, std::vector<uint64_t>& off, std::vector<uint32_t>& cells,
                       SolveStats* st) {
  HIP_OK(hipSetDevice(c.device));
  Workspace& W = c.ws;
  hipStream_t s = c.stream;
  routed.assign(n, 0);
  cost.assign(n, kMax);
  off.assign(n + 1, 0);
  cells.clear();
  if (n == 0) return;
  grow(W.plen, W.plen_cap, n);
  grow(W.pcost, W.pcost_cap, n);
  grow(W.poff, W.poff_cap, n + 1);
  FieldRead R = make_read(c);
  HIP_OK(hipEventRecord(W.ev1, s));
  hipLaunchKernelGGL(path_lengths, dim3(((uint32_t)n + 255) / 256), dim3(256), 0, s, R, (uint32_t)n, W.plen,
                     W.pcost);
  std::vector<uint32_t> plen(n);
  HIP_OK(hipMemcpyAsync(plen.data(), W.plen, n * 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(cost.data(), W.pcost, n * 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (size_t i = 0; i < n; ++i) {
    off[i + 1] = off[i] + plen[i];
    routed[i] = plen[i] ? 1 : 0;
  }
  size_t total = off[n];
  cells.resize(total);
  if (total) {
    grow(W.pcells, W.pcells_cap, total);
    HIP_OK(hipMemcpyAsync(W.poff, off.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
    WalkArgs WA;
    memset(&WA, 0, sizeof(WA));
    WA.R = R;
    WA.pool = W.pool;
    WA.poff = W.poff;
    WA.plen = W.plen;
    WA.cells = W.pcells;
    WA.counters = W.counters;
    WA.n = (uint32_t)n;
    launch_fill32(s, W.counters + C_ERROR * kCtrStride, 0u, 1);
    hipLaunchKernelGGL(walk_paths, dim3((uint32_t)n), dim3(64), 0, s, WA);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(cells.data(), W.pcells, total * 4, hipMemcpyDeviceToHost, s));
    uint32_t err = 0;
    HIP_OK(hipMemcpyAsync(&err, W.counters + C_ERROR * kCtrStride, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipEventRecord(W.ev2, s));
    HIP_OK(hipStreamSynchronize(s));
    if (err) throw Error(1, "device path reconstruction failed (flags " + std::to_string(err) + ")");
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, W.ev1, W.ev2));
    if (st) st->path_ms = ms;
    c.last.path_ms = ms;
  }
}

void read_field(Ctx& c, size_t f, uint32_t* dist, uint32_t* hops) {
  HIP_OK(hipSetDevice(c.device));
  if (f >= c.layout.size()) throw Error(1, "field index out of range");
  const FieldDesc& F = c.layout[f];
  const BoardView& b = c.views[F.board];
  size_t n = (size_t)F.ww * F.wh * b.layers;
  Workspace& W = c.ws;
  grow(W.fout, W.fout_cap, 2 * n);
  FieldRead R = make_read(c);
  uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(extract_field, dim3(blocks), dim3(256), 0, c.stream, R, (uint32_t)f, W.fout,
                     hops ? W.fout + n : nullptr);
  std::vector<uint32_t> tmp(hops ? 2 * n : n);
  HIP_OK(hipMemcpyAsync(tmp.data(), W.fout, tmp.size() * 4, hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  size_t total = (size_t)b.w * b.h * b.layers;
  std::fill(dist, dist + total, kMax);
  if (hops) std::fill(hops, hops + total, kMax);
  for (uint32_t l = 0; l < b.layers; ++l)
    for (uint32_t wy = 0; wy < F.wh; ++wy) {
      size_t src = ((size_t)l * F.wh + wy) * F.ww;
      size_t dst = ((size_t)l * b.h + F.y0 + wy) * b.w + F.x0;
      memcpy(dist + dst, tmp.data() + src, F.ww * 4);
      if (hops) memcpy(hops + dst, tmp.data() + n + src, F.ww * 4);
    }
}

}  // namespace mrf
