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
//     ([tile][layer][y][x]), each tile block followed by copies of the tile's
//     first and last columns so the side neighbours read their halo from two
//     contiguous runs.
//   * per-tile state word (touched / pending / active bits).
//
// Tile visit (process_tile): one workgroup loads a TW x TH x layers tile of one
// field into LDS with a one-cell halo ring from the neighbour tiles, computes
// the six step prices of entering every cell once (kMax = closed), then
// sweeps rows, columns and via columns (one line per lane, the whole line in
// registers, forward then backward) until a local fixpoint, writes back the
// improved labels and activates the neighbours whose shared border changed.
//
// Kernels
//   relax_round       one visit of every item on the round's active list;
//                     host loop until no item is activated (plain form).
//   relax_list        the same visits as a grid-stride loop over a device
//                     work list; rounds captured in a hipGraph with
//                     advance_round between them (graph form).
//   relax_queue       persistent workgroups over device work queues, one
//                     launch per solve (queue form).
//   relax_wave        (wavetile.h) one wavefront per 16 x 16 tile visit, two
//                     for 5-8 layers, u64 or packed u32 labels: the round
//                     kernel of the plain form and the serving loop where
//                     the batch qualifies.
//   board_scan / board_build
//                     a board's masks and bounds from the caller's planes.
//   path_scan / walk_paths
//                     canonical reconstruction: from dst, the lowest-CellIdx
//                     predecessor p with label(p) + (step(p, c), 1) == label(c)
//                     (mr-metal/src/lib.rs:304-396 order; masks re-checked).
//   extract_field     tile-major labels -> window-major u32 dist / hops.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <exception>
#include <numeric>
#include <stdexcept>
#include <thread>
#include <type_traits>

#include "engine.h"
#include "hostpool.h"
#include "workqueue.h"

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
// Item state: 0 until the tile's first visit, which stores S_TOUCHED with the
// index of its round above it (state >> 1).
enum : uint32_t { S_TOUCHED = 1u };
// device counters
enum : int {
  C_CHANGED = 0,
  C_VISITS = 1,
  C_ROUNDS = 2,
  C_ERROR = 3,
  C_NEXTN = 4,    // graph form and round driver: length of the next round's list
  C_NEXT0 = 5,    // 5..6: plain form, next list length, alternating per round
  C_PROUNDS = 7,  // graph form: rounds that had work
  C_GROUND = 8,   // graph form: current round index
  C_CURN = 9,     // graph form: length of the current round's list
  C_DONE = 10,    // plain form: workgroups of the launch that finished
  C_KEY0 = 11,    // 11..12: smallest activation key of the next list (delta stepping)
  C_THRESH = 13,  // delta stepping: the round processes items whose key is at most this
  C_PMAX = 14,    // packed labels: largest stored label with a real dist, over every item
  C_NCOUNTERS = 15
};
// Each device counter sits on its own 128-byte line: polled words and the
// work-list atomics must not share a line, or every pop and push queues
// behind the pollers.
constexpr uint32_t kCtrStride = 32;
constexpr uint32_t kCtrWords = C_NCOUNTERS * kCtrStride;
// Counters every visit touches are sharded: workgroup b counts into shard
// b % kShards of each kind (one 128-byte line per shard), so the visits of a
// launch spread their atomics over kShards addresses instead of queueing on
// one.  The host sums the shards.  SH_DONE holds each shard's count of
// finished workgroups of the running launch (see round_tail).
//
// The plain form's active lists are sharded the same way: item i goes to
// sub-list i % kShards (capacity `lcap` each), whose length is counted in
// shard i % kShards of one of three count sets that rotate with the round
// (appended to in round r: set r % 3; read by round r + 1; cleared by round
// r + 2's first workgroup).  The launch's last workgroup writes the running
// sums of the next list's sub-list lengths (kShards words after the shards,
// the list table); workgroup b of the next launch takes entry b of the
// sub-lists laid end to end.
constexpr uint32_t kShards = 64;
enum : uint32_t { SH_VISITS = 0, SH_ROUNDS = 1, SH_CHANGED = 2, SH_DONE = 3, SH_L0 = 4, SH_KINDS = 7 };
constexpr uint32_t kShardCounterWords = SH_KINDS * kShards * kCtrStride;
constexpr uint32_t kShardWords = kShardCounterWords + kShards;  // + the list table
static_assert(kShards == 64, "a launch's last workgroup reads the sub-list counts one per lane of a wavefront");
// Capacity of each sub-list of a list over n items (item i goes to sub-list
// i % kShards) and the words of one sharded list.
inline uint32_t lshard_cap(uint64_t n) { return (uint32_t)((n + kShards - 1) / kShards); }
inline uint64_t lshard_words(uint64_t n) { return (uint64_t)kShards * lshard_cap(n); }
inline uint64_t shard_sum(const uint32_t* shards, uint32_t kind) {
  uint64_t s = 0;
  for (uint32_t k = 0; k < kShards; ++k) s += shards[((size_t)kind * kShards + k) * kCtrStride];
  return s;
}
// Host-mapped words a plain-form launch publishes: [0] the next list's
// length, [1] error flags, [2] stamp; the round driver's group flags from
// here on.
constexpr uint32_t kDriverFlagsAt = 16;
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
  v.add16 = add16;
  v.halo = halo;
  v.halo_pen = halo_pen;
  v.bmask = bmask;
  v.raw = raw;
  v.xe = xe;
  v.ye = ye;
  v.ve = ve;
  v.va = 0;
  v.pad_ = 0;
  v.owner = legal && legal_owner ? owner : nullptr;
  v.lhalo = legal && legal_halo ? lhalo : nullptr;
  v.steps = steps;
  for (int i = 0; i < 16; ++i)
    if (via_allowed.size() > (size_t)i && via_allowed[i]) v.va |= 1u << i;
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

// Workgroup-uniform copy of a descriptor read from memory: every dword goes
// through readfirstlane, so the fields (and the plane pointers in them) live in
// scalar registers for the whole tile visit instead of being reloaded.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T& v) {
  static_assert(sizeof(T) % 4 == 0, "descriptor size must be a whole number of dwords");
  uint32_t w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// A workgroup-uniform load through the scalar data cache (constant address
// space): for data no workgroup of the running kernel writes -- descriptors
// the host uploaded, work lists and item states written by earlier launches
// (each launch starts with invalidated caches).  The value lands in scalar
// registers directly.
template <typename T>
__device__ __forceinline__ T const_load(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "descriptor size must be a whole number of dwords");
  const __attribute__((address_space(4))) uint32_t* q = (const __attribute__((address_space(4))) uint32_t*)p;
  uint32_t w[sizeof(T) / 4];
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; ++i) w[i] = q[i];
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// The same pointer in the global address space: loads through it are
// global_load (no flat address check, no LDS counter in the wait).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gbl(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}

// Four consecutive plane elements in one load (element alignment only).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef u32x4 u32x4_a4 __attribute__((aligned(4)));
typedef uint8_t u8x4 __attribute__((ext_vector_type(4)));
typedef u8x4 u8x4_a1 __attribute__((aligned(1)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
typedef u64x2 u64x2_a8 __attribute__((aligned(8)));

typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
typedef u16x4 u16x4_a2 __attribute__((aligned(2)));

__device__ __forceinline__ u32x4 load4(const uint32_t* p) { return *(const __attribute__((address_space(1))) u32x4_a4*)p; }
__device__ __forceinline__ u16x4 load4(const uint16_t* p) { return *(const __attribute__((address_space(1))) u16x4_a2*)p; }
__device__ __forceinline__ u8x4 load4(const uint8_t* p) { return *(const __attribute__((address_space(1))) u8x4_a1*)p; }

// Enter weight of cell g whose board mask is bm: the weight plane's value,
// or, on a board without one (every passable cell weighs 1), 1, and kMax for
// an obstacle.
__device__ __forceinline__ uint32_t weight_at(const BoardView& B, uint32_t g, uint8_t bm) {
  return B.cost ? gbl(B.cost)[g] : ((bm & M_OBST) ? kMax : 1u);
}

// Additive price of cell g from the board's plane (u32 or u16), 0 without one.
__device__ __forceinline__ uint32_t add_at(const BoardView& B, uint32_t g) {
  return B.add ? gbl(B.add)[g] : B.add16 ? (uint32_t)gbl(B.add16)[g] : 0u;
}

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

// Index of g in the field's additive overrides (cells ascending), or kMax.
__device__ __forceinline__ uint32_t override_index(const FieldDesc& F, const uint32_t* pool, uint32_t g) {
  const uint32_t* c = pool + F.ovr_off;
  uint32_t lo = 0, hi = F.n_ovr;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (c[mid] < g)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < F.n_ovr && c[lo] == g ? lo : kMax;
}

// Additive price of entering g for field F: the net's own override when it
// has one (negotiated batches price the net's own footprint apart,
// negotiated.rs:2781-2795), else the board plane; plus the foreign-halo
// penalty.
__device__ __forceinline__ uint32_t add_price(const BoardView& b, const FieldDesc& F, const uint32_t* pool,
                                              uint32_t g) {
  uint64_t a = b.add ? b.add[g] : b.add16 ? b.add16[g] : 0u;
  if (F.n_ovr) {
    const uint32_t k = override_index(F, pool, g);
    if (k != kMax) a = pool[F.ovr_off + F.n_ovr + k];
  }
  if (b.halo) {
    int32_t o = b.halo[g];
    if (o != -1 && o != F.own_id) a += b.halo_pen;
  }
  return a < 0xFFFFFFFFull ? (uint32_t)a : 0xFFFFFFFFu;
}

// Via bits of global cell g for field F, including per-net via exemptions
// (negotiated.rs:3010-3043).
__device__ uint8_t via_bits(const BoardView& b, const FieldDesc& F, const uint32_t* pool, uint32_t g) {
  uint32_t l = g / b.plane;
  const uint32_t* vp = pool + F.vpad_off;
  uint8_t r = b.raw[g];
  bool self_ok = !(r & R_VIANODE) && (!(r & R_VF) || sorted_has(vp, F.n_vpad, g));
  uint8_t bits = 0;
  if (!self_ok) return 0;
  if (l > 0 && ((b.va >> (l - 1)) & 1u)) {
    uint32_t q = g - b.plane;
    uint8_t rq = b.raw[q];
    if (!(rq & R_VIANODE) && (!(rq & R_VF) || sorted_has(vp, F.n_vpad, q))) bits |= K_DN;
  }
  if (l + 1 < b.layers && ((b.va >> l) & 1u)) {
    uint32_t q = g + b.plane;
    uint8_t rq = b.raw[q];
    if (!(rq & R_VIANODE) && (!(rq & R_VF) || sorted_has(vp, F.n_vpad, q))) bits |= K_UP;
  }
  return bits;
}

// Legal-search blocks of a cell with copper owner o and halo h for group grp
// (route_legal_once, negotiated.rs:2946-2966): bit 0 foreign copper (blocks
// even the endpoints), bit 1 a foreign or mixed halo on a cell that is not
// the group's own copper (blocks all but the endpoints).
__device__ __forceinline__ uint32_t legal_code(int32_t o, int32_t h, int32_t grp) {
  const bool copper = o >= 0 && o != grp;
  const bool halo = (h == -2 || (h >= 0 && h != grp)) && o != grp;
  return (copper ? 1u : 0u) | (halo ? 2u : 0u);
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
  const uint32_t legal =
      (b.owner || b.lhalo) ? legal_code(b.owner ? b.owner[g] : -1, b.lhalo ? b.lhalo[g] : -1, F.own_group) : 0u;
  if (legal & 2u) pass = false;
  if (!pass && !trace && (F.flags & F_FORCE)) {
    for (uint32_t k = 0; k < F.n_forced; ++k)
      if (pool[F.forced_off + k] == g) pass = true;
  }
  if (legal & 1u) pass = false;
  if (pass) m |= K_PASS;
  if (F.n_vpad) m = (uint8_t)((m & ~(K_DN | K_UP)) | via_bits(b, F, pool, g));
  ci.mask = m;
  ci.w = obst ? (pad ? 1u : kMax) : (b.cost ? b.cost[g] : 1u);
  ci.a = (b.add || b.add16 || b.halo || F.n_ovr) ? add_price(b, F, pool, g) : 0u;
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
  uint32_t max_local;
  // Plain form: the round's activation count goes to counter next_ctr;
  // counter clear_ctr (the previous round's, already read) is zeroed for the
  // next round; the last workgroup to finish publishes the counts to the
  // host-mapped words host_out (next count, error flags), then the launch's
  // stamp in host_out[2] with release semantics.
  uint32_t next_ctr, clear_ctr;
  uint32_t* host_out;
  uint32_t stamp;
  // Plain form and round driver: index of this launch's round (the graph
  // form reads it from counters[C_GROUND]).
  uint32_t round;
  // Round driver with several groups live: a visit that activates any tile
  // for the next round stores round + 1 in group_active[item_group[item]]
  // (its neighbours are in its own group), and the launch's last workgroup
  // copies the n_groups words to host_flags before it publishes the stamp;
  // null otherwise.
  uint32_t* group_active;
  const uint32_t* item_group;
  uint32_t* host_flags;
  uint32_t n_groups;
  // Delta stepping (key_cur non-null): every activation carries a key, the
  // smallest search cost among the border labels that changed; a round
  // processes the items whose key is at most counters[C_THRESH] and carries
  // the others over with their keys.  The launch's last workgroup sets the
  // next threshold to the next list's smallest key plus delta.
  uint32_t* key_cur;
  uint32_t* key_next;
  uint32_t key_ctr, key_clear;
  uint32_t delta;
  // Packed labels (wave form): per item, the largest label its last visit
  // stored with a real dist (the host's certificate); null otherwise.
  uint32_t* item_max;
  // Sharded counters (kShardWords words).
  uint32_t* shards;
  // Plain form: sharded lists of lcap entries per shard; the count sets of
  // the round's list, of the next list and of the set to clear.
  uint32_t lcap;
  uint32_t lset_cur, lset_next, lset_clear;
};

// Count of sub-list s in count set `set`.
__device__ __forceinline__ uint32_t* lcount(const RelaxArgs& A, uint32_t set, uint32_t s) {
  return A.shards + ((size_t)(SH_L0 + set) * kShards + s) * kCtrStride;
}

// Append item nb to sub-list nb % kShards of the next list.
__device__ __forceinline__ void list_append(const RelaxArgs& A, uint32_t nb) {
  const uint32_t s = nb % kShards;
  const uint32_t pos = atomicAdd(lcount(A, A.lset_next, s), 1u);
  A.list_next[(size_t)s * A.lcap + pos] = nb;
}

// The item of this workgroup in a plain-form round (the grid is the list's
// length): entry blockIdx of the sub-lists laid end to end, found from the
// list table (one running sum per lane; the sub-list is the number of sums
// at or below blockIdx).  Workgroup 0 clears the count set two rounds old.
__device__ __forceinline__ uint32_t round_item(const RelaxArgs& A) {
  const uint32_t lane = threadIdx.x & 63u;
  if (blockIdx.x == 0 && threadIdx.x < kShards)
    __hip_atomic_store(lcount(A, A.lset_clear, threadIdx.x), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t b = blockIdx.x;
  const uint32_t run = gbl(A.shards + kShardCounterWords)[lane];
  const uint32_t s = (uint32_t)__popcll(__ballot(run <= b));
  const uint32_t start = s ? (uint32_t)__builtin_amdgcn_readlane((int)run, (int)(s - 1)) : 0u;
  return const_load(A.list_cur + (size_t)s * A.lcap + (b - start));
}

// The workgroup's shard of counter kind `kind`.
__device__ __forceinline__ uint32_t* shard_of(const RelaxArgs& A, uint32_t kind) {
  return A.shards + ((size_t)kind * kShards + blockIdx.x % kShards) * kCtrStride;
}

// Words of one tile block in the label buffer: the L x TH x TW labels
// ([layer][y][x]) followed by copies of the tile's first and last columns
// ([layer][y] each), which the left and right neighbours read as their halo
// without touching one cache line per cell.
__host__ __device__ inline uint64_t tile_block_words(uint32_t L, uint32_t tw, uint32_t th) {
  return (uint64_t)L * th * tw + 2ull * L * th;
}

// LDS carve-up for one tile of TW x TH cells on L layers.
//   lab   u64 labels [L][TH+2][RS] with a one-cell halo ring; RS = TW+3 is odd
//         so lanes that each own a row hit distinct banks.
//   sxl / sxr / syt / syb / svd / svu
//         step prices [L][TH][PS] for entering each cell from x-1, x+1, y-1,
//         y+1, layer-1, layer+1.  Computed once per visit while the tile
//         loads; the relaxation rounds only read them.  PS = TW+1 is odd.
//         Wide form: u32 with kMax = closed.  Narrow form (boards whose every
//         step price is below 0x7FFF): i16 with -1 = closed, so a
//         sign-extending load yields the same u32 value (kMax when closed)
//         from half the LDS.
//   rchg / cchg / vchg
//         per-line change masks (row lines, column lines, via columns): bit i
//         set when cell i of the line improved during this visit.
template <bool NARROW>
struct Lds {
  using Step = typename std::conditional<NARROW, int16_t, uint32_t>::type;
  uint64_t* lab;
  Step* sxl;
  Step* sxr;
  Step* syt;
  Step* syb;
  Step* svd;
  Step* svu;
  uint32_t* rchg;
  uint32_t* cchg;
  uint32_t* vchg;
  Step* hstep;  // [2*L*(TH+TW)] lower bound of the price of entering each halo cell from the tile
  uint32_t* xg;  // [TW+1] x gap prices entering column i from the left
  uint32_t* yg;  // [TH+1]
  uint32_t* vg;  // [16]
  uint32_t* misc;
};

// Largest step price the narrow form holds (0x7FFF is reserved).
constexpr uint32_t kNarrowMaxStep = 0x7FFE;

__device__ __forceinline__ uint32_t step_of(uint32_t v) { return v; }
__device__ __forceinline__ uint32_t step_of(int16_t v) { return (uint32_t)(int32_t)v; }
template <typename ST>
__device__ __forceinline__ ST to_step(uint32_t v) {
  return (ST)v;  // narrow: kMax -> -1, prices <= kNarrowMaxStep unchanged
}

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline size_t lds_bytes(uint32_t tw, uint32_t th, uint32_t L, bool narrow) {
  const size_t rs = tw + 3, rh = th + 2, ps = tw + 1;
  const size_t ncell = (size_t)L * th * ps, nlab = (size_t)L * rh * rs;
  const size_t nstep = L > 1 ? 6 : 4;
  size_t s = 0;
  s += align16(nlab * 8);
  s += nstep * align16(ncell * (narrow ? 2 : 4));
  s += align16((size_t)L * th * 4) + align16((size_t)L * tw * 4) + align16((size_t)th * tw * 4);
  s += align16((size_t)2 * L * (th + tw) * (narrow ? 2 : 4));
  s += align16((tw + 1) * 4) + align16((th + 1) * 4) + align16(16 * 4) + 64;
  return s;
}

template <int TW, int TH, bool NARROW>
__device__ inline Lds<NARROW> carve(char* base, uint32_t L) {
  using ST = typename Lds<NARROW>::Step;
  constexpr size_t RS = TW + 3, RH = TH + 2, PS = TW + 1;
  const size_t ncell = (size_t)L * TH * PS, nlab = (size_t)L * RH * RS;
  Lds<NARROW> s;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    char* p = base + o;
    o += align16(bytes);
    return p;
  };
  s.lab = (uint64_t*)take(nlab * 8);
  s.sxl = (ST*)take(ncell * sizeof(ST));
  s.sxr = (ST*)take(ncell * sizeof(ST));
  s.syt = (ST*)take(ncell * sizeof(ST));
  s.syb = (ST*)take(ncell * sizeof(ST));
  if (L > 1) {
    s.svd = (ST*)take(ncell * sizeof(ST));
    s.svu = (ST*)take(ncell * sizeof(ST));
  } else {
    s.svd = s.svu = nullptr;
  }
  s.rchg = (uint32_t*)take((size_t)L * TH * 4);
  s.cchg = (uint32_t*)take((size_t)L * TW * 4);
  s.vchg = (uint32_t*)take((size_t)TH * TW * 4);
  s.hstep = (ST*)take((size_t)2 * L * (TH + TW) * sizeof(ST));
  s.xg = (uint32_t*)take((TW + 1) * 4);
  s.yg = (uint32_t*)take((TH + 1) * 4);
  s.vg = (uint32_t*)take(16 * 4);
  s.misc = (uint32_t*)take(64);
  return s;
}

// In-LDS label of a cell no path reaches: (MAX, 0).  Every label held in LDS
// is at most this value, and a candidate whose search cost saturates to MAX
// is (MAX, hops >= 1), never below it, so the saturating relaxation step
// below needs no separate reject test.
constexpr uint64_t kUnreached = kLabLimit;

// Relax `cur` through an edge of price s from label `from`: the candidate is
// (min(dist + s, MAX), hops + 1).  A sum that reaches MAX is the candidate
// astar_buf rejects (dijkstra.rs:303-313); saturated, it cannot win.
__device__ __forceinline__ void relax_to(uint64_t from, uint32_t s, uint64_t& cur, uint32_t& chg, uint32_t bit) {
  const uint32_t hi = __builtin_elementwise_add_sat((uint32_t)(from >> 32), s);
  const uint32_t lo = (uint32_t)from + 1u;
  const uint64_t cand = ((uint64_t)hi << 32) | lo;
  const bool better = cand < cur;
  cur = better ? cand : cur;
  chg = better ? (chg | bit) : chg;
}

// Gated step price of entering a cell: `open` is the cell's passability and
// the directed-edge bit of the move.
__device__ __forceinline__ uint32_t gate(bool open, uint32_t e, uint32_t w, uint32_t a) {
  return open ? step_price(e, w, a) : kMax;
}

// Per-phase change flags.  Phase k writes slot k & 1 with stamp k + 1; the
// slot is next written in phase k + 2, after the barrier that ends phase
// k + 1, so every thread reads phase k's flag before anyone overwrites it.
template <bool NARROW>
__device__ __forceinline__ bool phase_changed(const Lds<NARROW>& s, bool ch, uint32_t k) {
  if (ch) s.misc[8 + (k & 1)] = k + 1;
  __syncthreads();
  return uniform_u32(s.misc[8 + (k & 1)]) == k + 1;
}

// In-LDS relaxation of one tile to a local fixpoint.  Rows, columns and via
// columns are swept in turn, forward then backward, one line per lane with
// the whole line in registers; a forward + backward sweep leaves a line at its
// own fixpoint, so the tile is at a fixpoint once every other phase after a
// sweep has changed nothing.  Returns true when the round cap was hit.
template <int TW, int TH, int LM, bool NARROW>
__device__ bool relax_local(const Lds<NARROW>& s, uint32_t L, uint32_t max_local, uint32_t& rounds) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  constexpr uint32_t RS = TW + 3, RH = TH + 2, PS = TW + 1;
  constexpr uint32_t LAB_PLANE = RH * RS, STEP_PLANE = TH * PS;
  const uint32_t nphase = L > 1 ? 3u : 2u;
  uint32_t k = 0, quiet = 0;
  for (rounds = 0; rounds < max_local; ++rounds) {
    // rows: lane per (layer, y)
    bool ch = false;
    for (uint32_t line = tid; line < L * TH; line += nt) {
      const uint32_t l = line / TH, y = line - l * TH;
      uint64_t* lp = s.lab + l * LAB_PLANE + (y + 1) * RS;
      const auto* sp = s.sxl + l * STEP_PLANE + y * PS;
      uint64_t v[TW + 2];
      uint32_t st[TW];
      uint32_t chg = 0;
#pragma unroll
      for (int i = 0; i < TW + 2; ++i) v[i] = lp[i];
#pragma unroll
      for (int x = 0; x < TW; ++x) st[x] = step_of(sp[x]);
#pragma unroll
      for (int x = 0; x < TW; ++x) relax_to(v[x], st[x], v[x + 1], chg, 1u << x);
      sp = s.sxr + l * STEP_PLANE + y * PS;
#pragma unroll
      for (int x = 0; x < TW; ++x) st[x] = step_of(sp[x]);
#pragma unroll
      for (int x = TW - 1; x >= 0; --x) relax_to(v[x + 2], st[x], v[x + 1], chg, 1u << x);
      if (chg) {
        ch = true;
#pragma unroll
        for (int x = 0; x < TW; ++x) lp[x + 1] = v[x + 1];
        s.rchg[line] |= chg;
      }
    }
    quiet = phase_changed(s, ch, k++) ? 0u : quiet + 1;
    if (k >= nphase && quiet + 1 >= nphase) return false;
    // columns: lane per (layer, x)
    ch = false;
    for (uint32_t line = tid; line < L * TW; line += nt) {
      const uint32_t l = line / TW, x = line - l * TW;
      uint64_t* lp = s.lab + l * LAB_PLANE + x + 1;
      const auto* sp = s.syt + l * STEP_PLANE + x;
      uint64_t v[TH + 2];
      uint32_t st[TH];
      uint32_t chg = 0;
#pragma unroll
      for (int i = 0; i < TH + 2; ++i) v[i] = lp[i * RS];
#pragma unroll
      for (int y = 0; y < TH; ++y) st[y] = step_of(sp[y * PS]);
#pragma unroll
      for (int y = 0; y < TH; ++y) relax_to(v[y], st[y], v[y + 1], chg, 1u << y);
      sp = s.syb + l * STEP_PLANE + x;
#pragma unroll
      for (int y = 0; y < TH; ++y) st[y] = step_of(sp[y * PS]);
#pragma unroll
      for (int y = TH - 1; y >= 0; --y) relax_to(v[y + 2], st[y], v[y + 1], chg, 1u << y);
      if (chg) {
        ch = true;
#pragma unroll
        for (int y = 0; y < TH; ++y) lp[(y + 1) * RS] = v[y + 1];
        s.cchg[line] |= chg;
      }
    }
    quiet = phase_changed(s, ch, k++) ? 0u : quiet + 1;
    if (k >= nphase && quiet + 1 >= nphase) return false;
    if (L > 1) {
      // via columns: lane per (y, x), at most LM layers
      ch = false;
      for (uint32_t col = tid; col < TW * TH; col += nt) {
        const uint32_t y = col / TW, x = col - y * TW;
        uint64_t* lp = s.lab + (y + 1) * RS + x + 1;
        const auto* dp = s.svd + y * PS + x;
        const auto* up = s.svu + y * PS + x;
        uint64_t v[LM + 2];
        uint32_t sf[LM], sb[LM];
        uint32_t chg = 0;
        v[0] = kUnreached;
        v[LM + 1] = kUnreached;
#pragma unroll
        for (int l = 0; l < LM; ++l) {
          const bool in = (uint32_t)l < L;
          v[l + 1] = in ? lp[l * LAB_PLANE] : kUnreached;
          sf[l] = in ? step_of(dp[l * STEP_PLANE]) : kMax;
          sb[l] = in ? step_of(up[l * STEP_PLANE]) : kMax;
        }
#pragma unroll
        for (int l = 0; l < LM; ++l) relax_to(v[l], sf[l], v[l + 1], chg, 1u << l);
#pragma unroll
        for (int l = LM - 1; l >= 0; --l) relax_to(v[l + 2], sb[l], v[l + 1], chg, 1u << l);
        if (chg) {
          ch = true;
#pragma unroll
          for (int l = 0; l < LM; ++l)
            if ((uint32_t)l < L) lp[l * LAB_PLANE] = v[l + 1];
          s.vchg[col] |= chg;
        }
      }
      quiet = phase_changed(s, ch, k++) ? 0u : quiet + 1;
      if (k >= nphase && quiet + 1 >= nphase) return false;
    }
  }
  return true;
}

// Result flags of one tile visit.
enum : uint32_t { T_LEFT = 1, T_RIGHT = 2, T_TOP = 4, T_BOTTOM = 8, T_ANY = 16, T_CAPPED = 32 };

// Result of one tile visit (uniform across the workgroup): flags and, per
// side (left, right, top, bottom) and over the tile (ka), the smallest search
// cost among the cells that improved there (kMax when none did).
struct TileResult {
  uint32_t flags;  // side flags only where that neighbour exists
  uint32_t kl, kr, kt, kb, ka;
  uint32_t tx;  // tiles per row of the field (item distance of top / bottom)
};

__device__ __forceinline__ uint64_t lab_in(uint64_t v) { return v < kUnreached ? v : kUnreached; }

// Load, relax to an in-LDS fixpoint, and write back one (field, tile) item.
// Returns T_* flags (valid in every thread).
// LEAN: the launch's boards carry no halo-owner or legal-search planes and its
// fields no additive overrides, own pads or via-exempt pads (forced
// endpoints remain), so none of that code or its descriptor fields exist in
// the kernel: fewer scalar registers held across the visit, smaller code.
// QUEUE: the visit runs inside the persistent work-queue kernel, where other
// workgroups write labels and states during the same launch: every label of
// the solve starts unreached (filled before the launch), so any label read is
// an upper bound of its final value, whoever wrote it and when; labels, the
// own state and neighbour borders are read and written with device-scope
// atomics (never through a cache another workgroup's writes bypass); labels
// are lowered with an atomic minimum, since two visits of one tile may
// overlap; the first visit is the one whose atomic OR sets S_TOUCHED.
template <int TW, int TH, int LM, bool NARROW, bool LEAN, bool QUEUE = false>
__device__ TileResult process_tile(const RelaxArgs& A, uint32_t item, char* smem, uint32_t round) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  const uint32_t f = const_load(A.item_field + item);
  // the item's own state travels with the descriptor loads: whether the tile
  // holds labels is known before any plane load is issued
  // (an item is visited at most once per launch, so no workgroup of this
  // launch has written its state yet); in the queue form every visit loads
  // the tile's labels, and the first visit is known after the first barrier
  const bool touched = QUEUE ? true : (const_load(A.state + item) & S_TOUCHED) != 0;
  uint32_t first_was = S_TOUCHED;
  if (QUEUE && tid == 0)
    first_was = __hip_atomic_fetch_or(&A.state[item], S_TOUCHED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const FieldDesc F = const_load(A.fields + f);
  const BoardView B = const_load(A.boards + F.board);
  const uint32_t t = item - F.item0;
  const uint32_t txi = t % F.tx, tyi = t / F.tx;
  constexpr uint32_t RS = TW + 3, RH = TH + 2, PS = TW + 1;
  constexpr uint32_t LAB_PLANE = RH * RS, STEP_PLANE = TH * PS, PT = TW * TH;
  const uint32_t L = B.layers;
  const uint32_t ncell = L * PT;
  const uint32_t wx0 = txi * TW, wy0 = tyi * TH;
  Lds<NARROW> s = carve<TW, TH, NARROW>(smem, L);
  using ST = typename Lds<NARROW>::Step;
  const uint64_t TB = tile_block_words(L, TW, TH);
  const uint64_t blockbase = F.lab_off + (uint64_t)t * TB;

  // ---- loads that do not depend on the tile states go first -----------------
  // The states of the neighbours decide whether their border labels exist;
  // the board planes and labels of the first batch of cells and the
  // neighbours' border labels (speculatively: the address is valid whenever
  // the neighbour exists, the value is kept only if it was visited) are
  // issued before the barrier that publishes the states, so one memory round
  // trip covers them.
  uint32_t stv = 0;
  if (!QUEUE && tid >= 1 && tid < 5) {
    uint32_t nb = item;
    bool ok = true;
    if (tid == 1) ok = txi > 0, nb = item - 1;
    if (tid == 2) ok = txi + 1 < F.tx, nb = item + 1;
    if (tid == 3) ok = tyi > 0, nb = item - F.tx;
    if (tid == 4) ok = tyi + 1 < F.ty, nb = item + F.tx;
    stv = ok ? __hip_atomic_load(&A.state[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
  }
  // halo ring, two entries per thread: sides 0..3 = left, right, top, bottom
  const uint32_t nhalo = 2u * L * (TH + TW);
  // Entry j: the neighbour's label block word `src`, its ring slot `at`, the
  // side, whether the neighbour exists, and the board cell `gc` it holds.
  auto halo_entry = [&](uint32_t j, uint64_t& src, uint32_t& at, uint32_t& side, bool& exists, uint32_t& gc) {
    if (j < 2u * L * TH) {
      side = j / (L * TH);
      const uint32_t k = j - side * L * TH, l = k / TH, y = k - l * TH;
      const bool row_in = wy0 + y < F.wh;
      if (side == 0) {
        exists = row_in && txi > 0;
        src = blockbase - TB + ncell + L * TH + k;  // left tile's last column
      } else {
        exists = row_in && txi + 1 < F.tx;
        src = blockbase + TB + ncell + k;  // right tile's first column
      }
      at = l * LAB_PLANE + (y + 1) * RS + (side ? TW + 1 : 0);
      gc = (l * B.h + F.y0 + wy0 + y) * B.w + F.x0 + wx0 + (side ? TW : 0u) - (side ? 0u : 1u);
    } else {
      const uint32_t jj = j - 2u * L * TH, sd = jj / (L * TW), k = jj - sd * L * TW;
      const uint32_t l = k / TW, x = k - l * TW;
      const bool col_in = wx0 + x < F.ww;
      side = 2 + sd;
      if (sd == 0) {
        exists = col_in && tyi > 0;
        src = blockbase - (uint64_t)F.tx * TB + l * PT + (TH - 1) * TW + x;
      } else {
        exists = col_in && tyi + 1 < F.ty;
        src = blockbase + (uint64_t)F.tx * TB + l * PT + x;
      }
      at = l * LAB_PLANE + (sd ? (TH + 1) * RS : 0) + x + 1;
      gc = (l * B.h + F.y0 + wy0 + (sd ? TH : 0u) - (sd ? 0u : 1u)) * B.w + F.x0 + wx0 + x;
    }
    if (!exists) src = blockbase, gc = 0;
  };
  uint64_t hv[2];
  uint32_t hat[2], hside[2], hgc[2], hw[2], ha[2];
  uint8_t hb[2];
  bool hok[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const uint32_t j = u * nt + tid;
    uint64_t src = blockbase;
    hat[u] = 0;
    hside[u] = 0;
    hgc[u] = 0;
    hok[u] = false;
    if (j < nhalo) halo_entry(j, src, hat[u], hside[u], hok[u], hgc[u]);
    hv[u] = __hip_atomic_load(&A.labels[src], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    hb[u] = gbl(B.bmask)[hgc[u]];
    hw[u] = weight_at(B, hgc[u], hb[u]);
    ha[u] = add_at(B, hgc[u]);
  }
  // Interior planes by quads: four consecutive cells of one row per thread
  // per batch, loaded with one vector load per plane when the quad lies
  // inside the window (clamped scalar loads at the window's edge; cells
  // outside the window are closed afterwards).
  constexpr uint32_t QPR = TW / 4;  // quads per tile row
  const uint32_t nquad = L * TH * QPR;
  uint32_t qrowg = 0, qwxb = 0, qc0 = 0, ql = 0, qy = 0, qx0 = 0;
  bool qvalid = false;
  uint32_t wv[4], av[4];
  int32_t ov[4];
  uint8_t bv[4];
  uint8_t lg[4];  // legal-search block codes
  bool inw[4];
  uint64_t lv[4];
  const bool legal_planes = !LEAN && (B.owner || B.lhalo);
  auto load_planes = [&](uint32_t q) {
    qvalid = q < nquad;
    const uint32_t qq = qvalid ? q : 0;
    ql = qq / (TH * QPR);
    const uint32_t rq = qq - ql * TH * QPR;
    qy = rq / QPR;
    qx0 = (rq - qy * QPR) * 4;
    qc0 = ql * PT + qy * TW + qx0;
    const uint32_t wy = wy0 + qy;
    qwxb = wx0 + qx0;
#pragma unroll
    for (int k = 0; k < 4; ++k) inw[k] = qvalid && wy < F.wh && qwxb + k < F.ww;
    qrowg = (ql * B.h + F.y0 + min(wy, F.wh - 1)) * B.w + F.x0;
    if (inw[3]) {
      const uint32_t g0 = qrowg + qwxb;
      const u8x4 b4 = load4(B.bmask + g0);
      const u32x4 c4 = B.cost ? load4(B.cost + g0) : u32x4{1, 1, 1, 1};
      u32x4 a4 = u32x4{0, 0, 0, 0};
      if (B.add) {
        a4 = load4(B.add + g0);
      } else if (B.add16) {
        const u16x4 h4 = load4(B.add16 + g0);
        a4 = u32x4{h4[0], h4[1], h4[2], h4[3]};
      }
      const u32x4 o4 = (!LEAN && B.halo) ? load4((const uint32_t*)B.halo + g0) : u32x4{kMax, kMax, kMax, kMax};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        bv[k] = b4[k];
        wv[k] = (B.cost || !(bv[k] & M_OBST)) ? c4[k] : kMax;
        av[k] = a4[k];
        ov[k] = (int32_t)o4[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t g = qrowg + min(qwxb + k, F.ww - 1);
        bv[k] = gbl(B.bmask)[g];
        wv[k] = weight_at(B, g, bv[k]);
        av[k] = add_at(B, g);
        ov[k] = (!LEAN && B.halo) ? gbl(B.halo)[g] : -1;
      }
    }
    if (legal_planes) {
      const uint32_t g0 = qrowg + (inw[3] ? qwxb : 0u);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t g = inw[3] ? g0 + k : qrowg + min(qwxb + k, F.ww - 1);
        lg[k] = (uint8_t)legal_code(B.owner ? gbl(B.owner)[g] : -1, B.lhalo ? gbl(B.lhalo)[g] : -1, F.own_group);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) lg[k] = 0;
    }
    if (QUEUE) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        lv[k] = __hip_atomic_load(&A.labels[blockbase + qc0 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (touched) {
      // a quad is contiguous in the tile block whatever the window
      const auto* lp = (const __attribute__((address_space(1))) u64x2_a8*)(A.labels + blockbase + qc0);
      const u64x2 l01 = lp[0], l23 = lp[1];
      lv[0] = l01[0];
      lv[1] = l01[1];
      lv[2] = l23[0];
      lv[3] = l23[1];
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) lv[k] = kUnreached;
    }
  };
  load_planes(tid);
  if (tid >= 1 && tid < 5) s.misc[tid] = stv;
  if (QUEUE && tid == 0) s.misc[6] = first_was;
  if (tid == 0) {
    s.misc[5] = 0;  // result flags
    s.misc[8] = 0;  // phase change stamps
    s.misc[9] = 0;
    for (int k = 10; k < 15; ++k) s.misc[k] = kMax;  // side keys
  }
  // tile edge prices
  for (uint32_t i = tid; i <= TW; i += nt) {
    uint32_t gx = F.x0 + wx0 + i;  // entering column gx from gx-1
    s.xg[i] = (gx >= 1 && gx < B.w) ? gbl(B.xe)[gx - 1] : 0u;
  }
  for (uint32_t i = tid; i <= TH; i += nt) {
    uint32_t gy = F.y0 + wy0 + i;
    s.yg[i] = (gy >= 1 && gy < B.h) ? gbl(B.ye)[gy - 1] : 0u;
  }
  for (uint32_t i = tid; i < 16; i += nt) s.vg[i] = (i + 1 < L) ? gbl(B.ve)[i] : 0u;
  for (uint32_t i = tid; i < L * TH; i += nt) s.rchg[i] = 0;
  for (uint32_t i = tid; i < L * TW; i += nt) s.cchg[i] = 0;
  for (uint32_t i = tid; i < PT; i += nt) s.vchg[i] = 0;
  __syncthreads();
  // A neighbour's border labels count only if its first visit ran in an
  // earlier round: those were stored before a kernel boundary.  A neighbour
  // whose first visit runs in this round may have published its state before
  // the labels this tile loaded alongside it (the two loads are not ordered),
  // so its border is taken as unreached, an upper bound; that visit then
  // offers its border to this tile, unreached to it, and activates it for the
  // next round.
  // (the queue form takes every existing neighbour's border: its labels
  // start unreached, so any value read is an upper bound)
  auto earlier = [&](uint32_t st) { return (st & S_TOUCHED) && (st >> 1) < round; };
  const bool first_visit = QUEUE ? (uniform_u32(s.misc[6]) & S_TOUCHED) == 0 : !touched;
  const uint32_t nbv = QUEUE ? 15u : (earlier(uniform_u32(s.misc[1])) ? 1u : 0u) | (earlier(uniform_u32(s.misc[2])) ? 2u : 0u) |
                       (earlier(uniform_u32(s.misc[3])) ? 4u : 0u) | (earlier(uniform_u32(s.misc[4])) ? 8u : 0u);
  // halo labels into the ring (a neighbour never visited has none), and a
  // lower bound of the price of stepping from the tile into each halo cell:
  // the neighbour's gap price times the cell's weight (1 for an obstacle
  // that is an own pad) plus its additive price (0 when the net carries
  // overrides, which only lower it); closed when the cell is a trace node,
  // a foreign obstacle or its edge towards the tile is closed.
  const uint32_t* hpads = A.pool + F.pad_off;
  auto entry_price = [&](uint32_t side, uint32_t gc, uint8_t bm, uint32_t w, uint32_t a) -> uint32_t {
    constexpr uint8_t kEdge[4] = {K_PX, K_NX, K_PY, K_NY};  // the cell's edge facing the tile
    const uint8_t m = (uint8_t)((bm >> 1) & 0x7E);
    if ((bm & M_TRACE) || !(m & kEdge[side])) return kMax;
    uint32_t wgt = w;
    if (bm & M_OBST) {
      bool forced = false;
      if (F.flags & F_FORCE)
        for (uint32_t k = 0; k < F.n_forced; ++k) forced = forced || A.pool[F.forced_off + k] == gc;
      if (!LEAN && F.n_pad && sorted_has(hpads, F.n_pad, gc))
        wgt = 1u;
      else if (!forced)
        return kMax;
    }
    const uint32_t e = side == 0 ? s.xg[0] : side == 1 ? s.xg[TW] : side == 2 ? s.yg[0] : s.yg[TH];
    const uint32_t p = step_price(e, wgt, (!LEAN && F.n_ovr) ? 0u : a);
    return NARROW ? min(p, kNarrowMaxStep) : p;  // a smaller bound stays a bound
  };
#pragma unroll
  for (int u = 0; u < 2; ++u)
    if (u * nt + tid < nhalo) {
      s.lab[hat[u]] = (hok[u] && ((nbv >> hside[u]) & 1u)) ? lab_in(hv[u]) : kUnreached;
      s.hstep[u * nt + tid] = to_step<ST>(hok[u] ? entry_price(hside[u], hgc[u], hb[u], hw[u], ha[u]) : kMax);
    }
  for (uint32_t j = 2 * nt + tid; j < nhalo; j += nt) {
    uint64_t src;
    uint32_t at, side, gc;
    bool ex;
    halo_entry(j, src, at, side, ex, gc);
    s.lab[at] = (ex && ((nbv >> side) & 1u))
                    ? lab_in(__hip_atomic_load(&A.labels[src], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    : kUnreached;
    s.hstep[j] =
        to_step<ST>(ex ? entry_price(side, gc, gbl(B.bmask)[gc], weight_at(B, gc, gbl(B.bmask)[gc]), add_at(B, gc))
                       : kMax);
  }

  // Interior cells: labels (when the tile was visited before), the per-net
  // view of each cell (cell_info: own pads, forced endpoints, via exemptions,
  // additive and halo pricing) and the six step prices of entering it.
  const uint32_t* pads = A.pool + F.pad_off;
  const bool unit_w = B.cost == nullptr;
  // The six step prices of entering cell (l, y, x) = global g with board mask
  // bm, cost weight w0, additive price a0 and halo owner o.
  auto cell_steps = [&](uint32_t l, uint32_t y, uint32_t x, uint32_t g, uint8_t bm, uint32_t w0, uint32_t a0,
                        int32_t o, uint32_t legal) {
    const bool trace = (bm & M_TRACE) != 0, obst = (bm & M_OBST) != 0;
    uint8_t m = (uint8_t)((bm >> 1) & 0x7E);
    bool pass = !trace && !obst;
    uint32_t wgt = w0;
    if (!LEAN && obst && !trace && F.n_pad && sorted_has(pads, F.n_pad, g)) {
      pass = true;
      wgt = 1u;
    }
    if (legal & 2u) pass = false;  // foreign halo: only the endpoints stay open
    if (!pass && !trace && (F.flags & F_FORCE)) {
      for (uint32_t j = 0; j < F.n_forced; ++j)
        if (A.pool[F.forced_off + j] == g) pass = true;
    }
    if (legal & 1u) pass = false;  // foreign copper: closed even at the endpoints
    if (!LEAN && pass && F.n_vpad) m = (uint8_t)((m & ~(K_DN | K_UP)) | via_bits(B, F, A.pool, g));
    uint64_t a64 = a0;
    if (!LEAN && o != -1 && o != F.own_id) a64 += B.halo_pen;
    const uint32_t ap = a64 < 0xFFFFFFFFull ? (uint32_t)a64 : 0xFFFFFFFFu;
    const uint32_t pc = l * STEP_PLANE + y * PS + x;
    // Narrow boards bound every open step price below 0x7FFF, so it is
    // exact in 32-bit arithmetic without the saturation of step_price; on a
    // board without a weight plane every open cell of a narrow field weighs
    // 1 (a forced endpoint on an obstacle makes the field wide), so the
    // price is e + a.
    auto g6 = [&](bool open, uint32_t e) -> ST {
      if (NARROW) return to_step<ST>(open ? (unit_w ? e : e * wgt) + ap : kMax);
      return to_step<ST>(gate(open, e, wgt, ap));
    };
    s.sxl[pc] = g6(pass && (m & K_NX), s.xg[x]);
    s.sxr[pc] = g6(pass && (m & K_PX), s.xg[x + 1]);
    s.syt[pc] = g6(pass && (m & K_NY), s.yg[y]);
    s.syb[pc] = g6(pass && (m & K_PY), s.yg[y + 1]);
    if (L > 1) {
      s.svd[pc] = g6(pass && (m & K_DN) && l > 0, s.vg[l > 0 ? l - 1 : 0]);
      s.svu[pc] = g6(pass && (m & K_UP) && l + 1 < L, s.vg[l]);
    }
  };
  for (uint32_t qbase = 0; qbase < nquad; qbase += nt) {
    if (qbase) load_planes(qbase + tid);
    if (!qvalid) continue;
    const uint32_t l = ql, y = qy;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t x = qx0 + k;
      const uint8_t bm = inw[k] ? bv[k] : (uint8_t)M_TRACE;  // outside the window: closed
      cell_steps(l, y, x, qrowg + min(qwxb + k, F.ww - 1), bm, wv[k], av[k], ov[k], lg[k]);
      s.lab[l * LAB_PLANE + (y + 1) * RS + x + 1] = inw[k] ? lab_in(lv[k]) : kUnreached;
    }
  }
  __syncthreads();
  // The net's own additive overrides inside this tile (negotiated batches):
  // their cells are priced again over the board-plane prices just written.
  if (!LEAN && F.n_ovr) {
    const uint32_t* tab = A.pool + F.ovr_off + 2 * F.n_ovr;
    const uint32_t o0 = uniform_u32(tab[t]), o1 = uniform_u32(tab[t + 1]);
    if (o1 > o0) {
      const uint32_t* tc = tab + F.tx * F.ty + 1;
      const uint32_t plane = B.plane;
      for (uint32_t j = o0 + tid; j < o1; j += nt) {
        const uint32_t g = tc[j], a = tc[F.n_ovr + j];
        const uint32_t l = g / plane, r = g - l * plane, gy = r / B.w, gx = r - gy * B.w;
        cell_steps(l, gy - F.y0 - wy0, gx - F.x0 - wx0, g, gbl(B.bmask)[g], weight_at(B, g, gbl(B.bmask)[g]), a,
                   B.halo ? gbl(B.halo)[g] : -1,
                   legal_planes ? legal_code(B.owner ? gbl(B.owner)[g] : -1, B.lhalo ? gbl(B.lhalo)[g] : -1,
                                             F.own_group)
                                : 0u);
      }
      __syncthreads();
    }
  }
  // seeds (first visit only): passable sources start at label (0, 0)
  if (first_visit) {
    const uint32_t plane = B.plane;
    for (uint32_t k = tid; k < F.n_src; k += nt) {
      uint32_t g = A.pool[F.src_off + k];
      uint32_t l = g / plane, r = g - l * plane, gy = r / B.w, gx = r - gy * B.w;
      if (gx < F.x0 + wx0 || gy < F.y0 + wy0) continue;
      uint32_t x = gx - F.x0 - wx0, y = gy - F.y0 - wy0;
      if (x >= TW || y >= TH || wx0 + x >= F.ww || wy0 + y >= F.wh) continue;
      if (cell_info(B, F, A.pool, g).mask & K_PASS) {
        s.lab[l * LAB_PLANE + (y + 1) * RS + x + 1] = 0ull;
        atomicOr(&s.rchg[l * TH + y], 1u << x);
      }
    }
    __syncthreads();
  }

  // ---- in-LDS relaxation to a local fixpoint ----------------------------------
  uint32_t rounds = 0;
  const bool capped = relax_local<TW, TH, LM, NARROW>(s, L, A.max_local, rounds);
  // ---- write back: labels, border-column copies, side flags and keys -----------
  // Four cells per thread per batch: the change masks of a batch are read
  // together before any of them is used.
  uint32_t flags = 0;
  uint32_t kl = kMax, kr = kMax, kt = kMax, kb = kMax, ka = kMax;
  for (uint32_t base = 0; base < ncell; base += 4u * nt) {
    bool chg[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t c = min(base + u * nt + tid, ncell - 1);
      const uint32_t l = c / PT, r = c - l * PT, y = r / TW, x = r - y * TW;
      chg[u] = ((s.rchg[l * TH + y] >> x) | (s.cchg[l * TW + x] >> y) | (s.vchg[r] >> l)) & 1u;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t c = base + u * nt + tid;
      if (c >= ncell || !(chg[u] || !touched)) continue;
      const uint32_t l = c / PT, r = c - l * PT, y = r / TW, x = r - y * TW;
      const uint64_t v = s.lab[l * LAB_PLANE + (y + 1) * RS + x + 1];
      auto put = [&](uint64_t w) {
        if (QUEUE)
          __hip_atomic_fetch_min(&A.labels[w], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
          __hip_atomic_store(&A.labels[w], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      };
      put(blockbase + c);
      if (x == 0) put(blockbase + ncell + l * TH + y);
      if (x == TW - 1) put(blockbase + ncell + L * TH + l * TH + y);
      if (chg[u]) {
        // A changed border cell wakes the neighbour across that side only if
        // it can improve the neighbour's adjacent cell: the candidate it
        // offers is at least (dist + p, hops + 1) for p the lower bound of the
        // entry price (hstep), so a halo label (the neighbour's label at load
        // time, which only decreases) at or below that cannot improve, and
        // no path into the neighbour tile through this cell can either.
        // The planar edges are symmetric: when entering this (passable) cell
        // from a side is closed, leaving it towards that side is too.
        const uint32_t d = (uint32_t)(v >> 32);
        auto offers = [&](uint32_t p) {
          return ((uint64_t)__builtin_elementwise_add_sat(d, p) << 32) | (uint32_t)((uint32_t)v + 1u);
        };
        const uint64_t* row = s.lab + l * LAB_PLANE + (y + 1) * RS;
        const uint32_t pc = l * STEP_PLANE + y * PS + x;
        flags |= T_ANY;
        ka = min(ka, d);
        if (x == 0 && offers(step_of(s.hstep[l * TH + y])) < row[0] && step_of(s.sxl[pc]) != kMax) flags |= T_LEFT, kl = min(kl, d);
        if (x == TW - 1 && offers(step_of(s.hstep[L * TH + l * TH + y])) < row[TW + 1] && step_of(s.sxr[pc]) != kMax)
          flags |= T_RIGHT, kr = min(kr, d);
        if (y == 0 && offers(step_of(s.hstep[2 * L * TH + l * TW + x])) < s.lab[l * LAB_PLANE + x + 1] &&
            step_of(s.syt[pc]) != kMax)
          flags |= T_TOP, kt = min(kt, d);
        if (y == TH - 1 && offers(step_of(s.hstep[2 * L * TH + L * TW + l * TW + x])) < s.lab[l * LAB_PLANE + (TH + 1) * RS + x + 1] &&
            step_of(s.syb[pc]) != kMax)
          flags |= T_BOTTOM, kb = min(kb, d);
      }
    }
  }
  if (flags) {
    atomicOr(&s.misc[5], flags);
    atomicMin(&s.misc[14], ka);
    if (kl != kMax) atomicMin(&s.misc[10], kl);
    if (kr != kMax) atomicMin(&s.misc[11], kr);
    if (kt != kMax) atomicMin(&s.misc[12], kt);
    if (kb != kMax) atomicMin(&s.misc[13], kb);
  }
  // queue form: this workgroup's label writes are performed before any of
  // its activations can make another workgroup visit a neighbour
  if (QUEUE) __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();  // also orders this workgroup's label stores before the flag / queue updates
  TileResult res;
  const uint32_t sides = (txi > 0 ? T_LEFT : 0u) | (txi + 1 < F.tx ? T_RIGHT : 0u) | (tyi > 0 ? T_TOP : 0u) |
                         (tyi + 1 < F.ty ? T_BOTTOM : 0u) | T_ANY;
  res.flags = (uniform_u32(s.misc[5]) & sides) | (capped ? T_CAPPED : 0u);
  res.tx = F.tx;
  res.kl = uniform_u32(s.misc[10]);
  res.kr = uniform_u32(s.misc[11]);
  res.kt = uniform_u32(s.misc[12]);
  res.kb = uniform_u32(s.misc[13]);
  res.ka = uniform_u32(s.misc[14]);
  if (tid == 0) {
    if (!QUEUE && !touched)
      __hip_atomic_fetch_or(&A.state[item], S_TOUCHED | (round << 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd(shard_of(A, SH_VISITS), 1u);
    atomicAdd(shard_of(A, SH_ROUNDS), rounds);
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
__device__ __forceinline__ void activate(const RelaxArgs& A, uint32_t nb, uint32_t key) {
  if (A.key_next) atomicMin(&A.key_next[nb], key);
  if (atomicExch(&A.act_next[nb], 1u) == 0u) {
    if (A.lcap)
      list_append(A, nb);
    else
      A.list_next[atomicAdd(&A.counters[A.next_ctr * kCtrStride], 1u)] = nb;
  }
}

// Activations after a visit, run by every lane of the first wave: lanes 0..3
// take the left, right, top and bottom neighbour, lane 4 the tile itself when
// its visit hit the round cap (key `self_key` when no cell improved).  The
// flag exchanges run side by side and one atomicAdd reserves the list slots of
// all new entries, so the visit ends after two memory round trips instead of
// two per neighbour.  Returns the smallest key activated (kMax when none);
// every atomic whose result the list depends on has returned on exit.
__device__ __forceinline__ uint32_t activate_sides(const RelaxArgs& A, uint32_t item, const TileResult& res,
                                                   uint32_t self_key) {
  const uint32_t lane = threadIdx.x;
  const uint32_t f = res.flags;
  bool want = false;
  uint32_t nb = item, k = kMax;
  if (lane == 0) want = (f & T_LEFT) != 0, nb = item - 1, k = res.kl;
  if (lane == 1) want = (f & T_RIGHT) != 0, nb = item + 1, k = res.kr;
  if (lane == 2) want = (f & T_TOP) != 0, nb = item - res.tx, k = res.kt;
  if (lane == 3) want = (f & T_BOTTOM) != 0, nb = item + res.tx, k = res.kb;
  if (lane == 4) want = (f & T_CAPPED) != 0, k = res.ka == kMax ? self_key : res.ka;
  bool fresh = false;
  if (want) {
    if (A.key_next) atomicMin(&A.key_next[nb], k);
    fresh = atomicExch(&A.act_next[nb], 1u) == 0u;
  }
  const uint64_t m = __ballot(fresh);
  if (A.lcap) {
    // sharded lists: each new entry reserves its slot in its own sub-list
    if (fresh) list_append(A, nb);
  } else if (m) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&A.counters[A.next_ctr * kCtrStride], (uint32_t)__popcll(m));
    base = uniform_u32(base);
    if (fresh) A.list_next[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = nb;
  }
  uint32_t kmin = kMax;
#pragma unroll
  for (int j = 0; j < 5; ++j) kmin = min(kmin, (uint32_t)__builtin_amdgcn_readlane(want ? k : kMax, j));
  return kmin;
}

// Waves per SIMD the kernels are compiled for, the most that fit without
// register spills: tiles up to 16 wide with up to 8 layers fit six (at most
// 80 VGPRs); 16-layer via columns five; 32-wide line sweeps three.  The
// grid-stride form keeps a few more values live.
template <int TW, int LM>
struct MinWaves {
  static constexpr int round = TW >= 32 ? 3 : (LM >= 16 ? 5 : 6);
  static constexpr int list = TW >= 32 ? 2 : 4;
};

// After a visit (or a deferred item): activations, the item's flags and
// counters, and, in the launch's last workgroup, the counts, group flags and
// stamp published to the host.  Run by the first wavefront of the workgroup.
__device__ __forceinline__ void round_tail(const RelaxArgs& A, uint32_t item, uint32_t grp, const TileResult& res,
                                           bool defer, uint32_t key) {
  uint32_t kmin = kMax;
  if (defer) {
    if (threadIdx.x == 0) activate(A, item, key);  // carried over to the next round with its key
    kmin = key;
  } else {
    kmin = activate_sides(A, item, res, key);
  }
  uint32_t last = 0;  // this workgroup is the launch's last (lane 0)
  if (threadIdx.x == 0) {
    A.act_cur[item] = 0;
    // The group flag: an exchange whose result the done increment below
    // depends on, so it has completed before the last workgroup reads it.
    // Flags only hold round stamps up to this one, so the increment is 1.
    uint32_t was = 0;
    if (A.group_active && (defer || (res.flags & (T_LEFT | T_RIGHT | T_TOP | T_BOTTOM | T_CAPPED))))
      was = __hip_atomic_exchange(&A.group_active[grp], A.round + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (A.key_cur) A.key_cur[item] = kMax;
    if (!defer && (res.flags & (T_ANY | T_CAPPED))) atomicAdd(shard_of(A, SH_CHANGED), 1u);
    if (A.key_cur && kmin != kMax) atomicMin(&A.counters[A.key_ctr * kCtrStride], kmin);
    if (A.host_out) {
      // The last workgroup of the launch publishes the counts.  Every
      // workgroup's activation atomics returned (their results index the
      // list) before its done increment was issued, and atomics are performed
      // at the device's coherence point, so the last one reads final counts;
      // a relaxed increment keeps the caches untouched (an acquire-release
      // one would write back and invalidate L2 in every workgroup).
      // Two levels: a workgroup counts itself in its shard; the one that
      // completes its shard (whose size follows from the grid) resets it and
      // counts the shard in C_DONE; the one that completes C_DONE is last.
      // Each count is issued after the counts it completes have returned.
      const uint32_t slot = blockIdx.x % kShards;
      const uint32_t in_slot = (gridDim.x - slot + kShards - 1) / kShards;
      uint32_t* sd = shard_of(A, SH_DONE);
      const uint32_t prev = __hip_atomic_fetch_add(sd, was > A.round + 1u ? 2u : 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
      bool top = false;
      if (prev + 1 == in_slot) {
        __hip_atomic_store(sd, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t slots = gridDim.x < kShards ? gridDim.x : kShards;
        top = __hip_atomic_fetch_add(&A.counters[C_DONE * kCtrStride], 1u, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) + 1 == slots;
      }
      if (top) {
        last = 1;
        __hip_atomic_store(&A.counters[C_DONE * kCtrStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (A.key_cur) {
          const uint64_t kn = __hip_atomic_load(&A.counters[A.key_ctr * kCtrStride], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
          const uint64_t th = kn + A.delta;
          __hip_atomic_store(&A.counters[C_THRESH * kCtrStride], th < kMax ? (uint32_t)th : kMax, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
  }
  if (__shfl(last, 0)) {
    // the launch's last workgroup publishes the next round's size (sharded
    // lists: the longest sub-list, one lane per count), the error flags and,
    // without group flags, the stamp
    uint32_t nn;
    if (A.lcap) {
      // running sums of the sub-list lengths (one per lane) into the list table
      uint32_t run = __hip_atomic_load(lcount(A, A.lset_next, threadIdx.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t up = (uint32_t)__shfl_up((int)run, o);
        if (threadIdx.x >= (uint32_t)o) run += up;
      }
      A.shards[kShardCounterWords + threadIdx.x] = run;
      nn = (uint32_t)__builtin_amdgcn_readlane((int)run, 63);
    } else {
      nn = __hip_atomic_load(&A.counters[A.next_ctr * kCtrStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) {
      const uint32_t er = __hip_atomic_load(&A.counters[C_ERROR * kCtrStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&A.host_out[0], nn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&A.host_out[1], er, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      if (!A.host_flags) {
        __hip_atomic_store(&A.host_out[2], A.stamp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __threadfence_system();
      }
    }
  }
  if (A.host_flags) {
    // the first wave of the last workgroup copies the group flags, then
    // publishes the stamp
    if (__shfl(last, 0)) {
      for (uint32_t g = threadIdx.x; g < A.n_groups; g += 64)
        __hip_atomic_store(&A.host_flags[g],
                           __hip_atomic_load(&A.group_active[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __threadfence_system();
      if (threadIdx.x == 0) {
        __hip_atomic_store(&A.host_out[2], A.stamp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __threadfence_system();
      }
    }
  }
}

template <int TW, int TH, int LM, bool NARROW, bool LEAN>
__global__ void __launch_bounds__(256, (MinWaves<TW, LM>::round)) relax_round(RelaxArgs A) {
  extern __shared__ __align__(16) char smem[];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (A.key_cur && A.key_clear != A.key_ctr)
      __hip_atomic_store(&A.counters[A.key_clear * kCtrStride], kMax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const uint32_t item = round_item(A);
  const uint32_t grp = A.group_active ? const_load(A.item_group + item) : 0u;
  uint32_t key = 0;
  bool defer = false;
  if (A.key_cur) {
    key = uniform_u32(A.key_cur[item]);
    defer = key > uniform_u32(__hip_atomic_load(&A.counters[C_THRESH * kCtrStride], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT));
  }
  TileResult res{};
  if (!defer) res = process_tile<TW, TH, LM, NARROW, LEAN>(A, item, smem, A.round);
  if (threadIdx.x >= 64) return;
  round_tail(A, item, grp, res, defer, key);
}

#include "wavetile.h"

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

template <int TW, int TH, int LM, bool NARROW, bool LEAN>
__global__ void __launch_bounds__(256, (MinWaves<TW, LM>::list)) relax_list(RelaxArgs E, RelaxArgs O) {
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
    uint32_t item = const_load(R.list_cur + k);
    const TileResult res = process_tile<TW, TH, LM, NARROW, LEAN>(R, item, smem, r);
    if (threadIdx.x < 64) {
      if (threadIdx.x == 0) R.act_cur[item] = 0;
      activate_sides(R, item, res, res.ka);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// persistent form: one launch, a device work queue per XCD
// ---------------------------------------------------------------------------
//
// Every field belongs to one of kQueues queues (the host balances their
// tiles); workgroup b serves queue b % kQueues, the XCD the hardware
// dispatches it to, so a field's labels, states and queue traffic stay in
// one XCD's L2.  Correctness does not depend on that placement: every access
// the workgroups share is a device-scope atomic (workqueue.h, process_tile's
// queue form).  A workgroup takes a ticket, waits for its item, visits it,
// activates the neighbours its visit can improve (each a push unless already
// queued) and marks the visit done; it stops when a wait finds nothing
// pending, or, with an error flag, when a wait or the workgroup's whole run
// outlasts its limit.  Control flow is uniform: lane 0 decides, the decision
// reaches every lane through LDS and readfirstlane.

constexpr uint32_t kQueues = 8;

struct QueueArgs {
  uint64_t* slots;             // every queue's ring
  uint32_t* ctl;               // queue k: head, tail, pend at ctl[(3k + j) * kCtrStride]
  uint32_t slot0[kQueues];     // first slot of each queue's ring
  uint32_t cap[kQueues];
  uint64_t wait_ticks;         // longest wait for a slot or for work (s_memrealtime ticks)
  uint64_t run_ticks;          // longest run of one workgroup
};

struct DevOps {
  __device__ static uint32_t load(uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
  __device__ static uint64_t load64(uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static void store64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static uint32_t exch(uint32_t* p, uint32_t v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static uint32_t add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static uint32_t orr(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static bool cas(uint32_t* p, uint32_t expect, uint32_t want) {
    return __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static uint64_t now() { return __builtin_amdgcn_s_memrealtime(); }
  // about 2^k x 64 clocks
  __device__ static void pause(uint32_t k) {
    for (uint32_t i = 0; i < (1u << k); ++i) __builtin_amdgcn_s_sleep(1);
  }
  __device__ static void order() { __builtin_amdgcn_s_waitcnt(0); }
};

// Error flags of the persistent form (device counter C_ERROR).
enum : uint32_t { E_QWAIT = 64u, E_QTAKE = 128u, E_QPUSH = 256u, E_QRUN = 512u };

template <int TW, int TH, int LM, bool NARROW, bool LEAN>
__global__ void __launch_bounds__(256, (MinWaves<TW, LM>::round)) relax_queue(RelaxArgs A, QueueArgs Q) {
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t cmd;
  constexpr uint32_t kExit = 0xFFFFFFFFu;
  const uint32_t qi = blockIdx.x % kQueues;
  const wq::Queue q{Q.slots + Q.slot0[qi], Q.cap[qi], Q.ctl + (3 * qi + 0) * kCtrStride,
                    Q.ctl + (3 * qi + 1) * kCtrStride, Q.ctl + (3 * qi + 2) * kCtrStride};
  uint32_t* err = A.counters + C_ERROR * kCtrStride;
  const uint64_t start = DevOps::now();
  for (;;) {
    if (threadIdx.x == 0) {
      uint32_t d = kExit;
      if (DevOps::now() - start > Q.run_ticks) {
        atomicOr(err, E_QRUN);
      } else {
        uint32_t it = 0;
        const uint32_t r = wq::wait_item<DevOps>(q, wq::take_ticket<DevOps>(q), Q.wait_ticks, it);
        if (r == wq::GOT) {
          if (!wq::take<DevOps>(A.act_cur, it)) atomicOr(err, E_QTAKE);
          d = it;
        } else if (r == wq::TIMEOUT) {
          atomicOr(err, E_QWAIT);
        }  // STOP: nothing pending, no item can come
      }
      cmd = d;
    }
    __syncthreads();
    const uint32_t item = uniform_u32(cmd);
    __syncthreads();
    if (item == kExit) break;
    const TileResult res = process_tile<TW, TH, LM, NARROW, LEAN, true>(A, item, smem, 0u);
    if (threadIdx.x < 64) {
      // lanes 0..3: left, right, top, bottom neighbour; lane 4: the tile
      // itself when its visit hit the in-LDS round cap
      const uint32_t lane = threadIdx.x, fl = res.flags;
      bool want = false;
      uint32_t nb = item;
      if (lane == 0) want = (fl & T_LEFT) != 0, nb = item - 1;
      if (lane == 1) want = (fl & T_RIGHT) != 0, nb = item + 1;
      if (lane == 2) want = (fl & T_TOP) != 0, nb = item - res.tx;
      if (lane == 3) want = (fl & T_BOTTOM) != 0, nb = item + res.tx;
      if (lane == 4) want = (fl & T_CAPPED) != 0;
      if (want && !wq::activate<DevOps>(q, A.act_cur, nb, Q.wait_ticks)) atomicOr(err, E_QPUSH);
      if (lane == 0) {
        if (fl & (T_ANY | T_CAPPED)) atomicAdd(shard_of(A, SH_CHANGED), 1u);
        // after every lane's activations (finish's order() waits for them)
        if (!wq::finish<DevOps>(q, A.act_cur, item, Q.wait_ticks)) atomicOr(err, E_QPUSH);
        wq::done<DevOps>(q);
      }
    }
    __syncthreads();
  }
}

// The persistent work-queue form on the wave form's tile visits: one
// workgroup (one wavefront, two for 5-8 layers) per visit slot, taking tiles
// from the device queues like relax_queue.  Visits of one tile never overlap
// (the queue protocol), so the write-back stays plain device-scope stores;
// every label is read with a device-scope load.  Lean fields only (no
// halo-owner planes or per-net extras); SP as in relax_wave.
template <int TW, int TH, int LM, bool PK, bool SP>
__global__ void __launch_bounds__(64 * wave_count(LM, TH), (WaveMin<TW, LM, PK>::waves))
    relax_wave_queue(RelaxArgs A, QueueArgs Q) {
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t cmd[2];
  __shared__ uint64_t start;
  constexpr uint32_t kExit = 0xFFFFFFFFu;
  if (threadIdx.x == 0) start = DevOps::now();
  // The queue's words are formed where they are used from the kernel
  // arguments (nothing of the protocol stays live through a visit).
  auto queue = [&]() {
    const uint32_t qi = blockIdx.x % kQueues;
    return wq::Queue{Q.slots + Q.slot0[qi], Q.cap[qi], Q.ctl + (3 * qi + 0) * kCtrStride,
                     Q.ctl + (3 * qi + 1) * kCtrStride, Q.ctl + (3 * qi + 2) * kCtrStride};
  };
  for (;;) {
    if (threadIdx.x == 0) {
      const wq::Queue q = queue();
      uint32_t* err = A.counters + C_ERROR * kCtrStride;
      uint32_t d = kExit, first = 0;
      if (DevOps::now() - start > Q.run_ticks) {
        atomicOr(err, E_QRUN);
      } else {
        uint32_t it = 0;
        const uint32_t r = wq::wait_item<DevOps>(q, wq::take_ticket<DevOps>(q), Q.wait_ticks, it);
        if (r == wq::GOT) {
          if (!wq::take<DevOps>(A.act_cur, it)) atomicOr(err, E_QTAKE);
          d = it;
          // the tile's first visit is the one that marks it touched
          first = (__hip_atomic_fetch_or(&A.state[it], S_TOUCHED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                   S_TOUCHED) ? 0u : 1u;
        } else if (r == wq::TIMEOUT) {
          atomicOr(err, E_QWAIT);
        }  // STOP: nothing pending, no item can come
      }
      cmd[0] = d;
      cmd[1] = first;
    }
    __syncthreads();
    const uint32_t item = uniform_u32(cmd[0]);
    const bool first = uniform_u32(cmd[1]) != 0u;
    __syncthreads();
    if (item == kExit) break;
    const uint32_t f = const_load(A.item_field + item);
    const FieldDesc F = const_load(A.fields + f);
    const BoardView B = const_load(A.boards + F.board);
    // SP: the boards' step price planes (the host picks it only for boards
    // that carry them), else prices formed from the planes
    TileResult res;
    if constexpr (SP)
      res = wave_tile<TW, TH, LM, ADD_NONE, PK, false, false, true, true>(A, item, f, !first, F, B, smem, 0u);
    else if (B.add16)
      res = wave_tile<TW, TH, LM, ADD_U16, PK, false, false, true>(A, item, f, !first, F, B, smem, 0u);
    else if (B.add)
      res = wave_tile<TW, TH, LM, ADD_U32, PK, false, false, true>(A, item, f, !first, F, B, smem, 0u);
    else
      res = wave_tile<TW, TH, LM, ADD_NONE, PK, false, false, true>(A, item, f, !first, F, B, smem, 0u);
    if (threadIdx.x < 64) {
      // lanes 0..3: left, right, top, bottom neighbour; lane 4: the tile
      // itself when its visit hit the in-LDS round cap
      const uint32_t lane = threadIdx.x, fl = res.flags;
      bool want = false;
      uint32_t nb = item;
      if (lane == 0) want = (fl & T_LEFT) != 0, nb = item - 1;
      if (lane == 1) want = (fl & T_RIGHT) != 0, nb = item + 1;
      if (lane == 2) want = (fl & T_TOP) != 0, nb = item - res.tx;
      if (lane == 3) want = (fl & T_BOTTOM) != 0, nb = item + res.tx;
      if (lane == 4) want = (fl & T_CAPPED) != 0;
      const wq::Queue q = queue();
      uint32_t* err = A.counters + C_ERROR * kCtrStride;
      if (want && !wq::activate<DevOps>(q, A.act_cur, nb, Q.wait_ticks)) atomicOr(err, E_QPUSH);
      if (lane == 0) {
        if (fl & (T_ANY | T_CAPPED)) atomicAdd(shard_of(A, SH_CHANGED), 1u);
        if (!wq::finish<DevOps>(q, A.act_cur, item, Q.wait_ticks)) atomicOr(err, E_QPUSH);
        wq::done<DevOps>(q);
      }
    }
    __syncthreads();
  }
}

// Single-thread round step between two relax_list launches.
// Largest of n per-item maxima into counters[C_PMAX]: a grid-stride pass,
// reduced per workgroup (256 threads), one atomic per workgroup.
__global__ void __launch_bounds__(256) reduce_item_max(const uint32_t* m, size_t n, uint32_t* out) {
  __shared__ uint32_t wmax[4];
  uint32_t v = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    v = max(v, m[i]);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
    if (v) atomicMax(out, v);
  }
}

__global__ void advance_round(uint32_t* ctr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t n = ctr[C_NEXTN * kCtrStride];
  if (ctr[C_CURN * kCtrStride]) ctr[C_PROUNDS * kCtrStride] += 1;
  ctr[C_CURN * kCtrStride] = n;
  ctr[C_NEXTN * kCtrStride] = 0;
  ctr[C_GROUND * kCtrStride] += 1;
}

// item_field of a batch from its field descriptors: block f writes field f's
// tile items [item0, item0 + tx * ty).
__global__ void fill_item_field(const FieldDesc* fields, uint32_t n_fields, uint32_t* item_field) {
  for (uint32_t f = blockIdx.x; f < n_fields; f += gridDim.x) {
    const FieldDesc F = fields[f];
    const uint32_t nt = F.tx * F.ty;
    for (uint32_t k = threadIdx.x; k < nt; k += blockDim.x) item_field[F.item0 + k] = f;
  }
}

// Every touched item state to round 0 (the round driver restarting its
// round index).
__global__ void restamp_states(uint32_t* state, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride)
    if (state[i]) state[i] = S_TOUCHED;
}

__global__ void fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------------------
// Board upload on the device.  board_scan reduces the bounds the host needs
// to size the planes (largest and smallest enter weight of a passable cell,
// smallest additive price of a passable cell, largest additive price, whether
// any cell is passable) into st[0..5), minima as maxima of the complement so
// that every word starts at 0.  board_build then writes one cell per thread:
// the directed planar-edge bits and via bits (mr-core lib.rs:466-496: a
// planar step is forbidden when either end's constraint bit for that
// direction is set; a via needs both ends free of BOARD_VIA_NODE and
// via_forbidden and the layer pair's via allowed), the trace and obstacle
// bits, the raw via flags, and the obstacle bitset (one ballot per wave).
__global__ void __launch_bounds__(256) board_scan(const uint32_t* cost, const uint32_t* add, size_t n,
                                                  uint32_t* st) {
  __shared__ uint32_t red[4][5];
  uint32_t v[5] = {0u, 0u, 0u, 0u, 0u};  // max_w, max_add, passable, ~min_w, ~min_a
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint32_t w = cost[i];
    const uint32_t a = add ? add[i] : 0u;
    v[1] = max(v[1], a);
    if (w != kMax) {
      v[0] = max(v[0], w);
      v[2] = 1u;
      v[3] = max(v[3], ~w);
      v[4] = max(v[4], ~a);
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    for (int o = 32; o >= 1; o >>= 1) v[k] = max(v[k], (uint32_t)__shfl_xor((int)v[k], o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int k = threadIdx.x;
    const uint32_t m = max(max(red[0][k], red[1][k]), max(red[2][k], red[3][k]));
    if (m) atomicMax(st + k, m);
  }
}

__global__ void __launch_bounds__(256) board_build(const uint32_t* __restrict__ cost,
                                                   const uint8_t* __restrict__ board,
                                                   const uint8_t* __restrict__ vf, uint32_t w, uint32_t h,
                                                   uint32_t layers, uint32_t va_bits, uint8_t* __restrict__ bm,
                                                   uint8_t* __restrict__ raw, uint64_t* __restrict__ obits) {
  const size_t n = (size_t)w * h * layers;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ob = false;
  if (i < n) {
    const size_t plane = (size_t)w * h;
    const uint32_t x = (uint32_t)(i % w), y = (uint32_t)((i / w) % h), l = (uint32_t)(i / plane);
    const uint8_t bc = board ? board[i] : 0u;
    ob = cost[i] == kMax;
    uint8_t m = 0;
    if (bc & 0x01) m |= M_TRACE;
    if (ob) m |= M_OBST;
    auto open = [&](size_t v, uint8_t sb, uint8_t db) { return !board || !((bc & sb) || (board[v] & db)); };
    if (y > 0 && open(i - w, 0x04, 0x20)) m |= M_NY;
    if (x > 0 && open(i - 1, 0x08, 0x10)) m |= M_NX;
    if (x + 1 < w && open(i + 1, 0x10, 0x08)) m |= M_PX;
    if (y + 1 < h && open(i + w, 0x20, 0x04)) m |= M_PY;
    auto free_of_via_block = [&](size_t q) { return !((board && (board[q] & 0x02)) || (vf && vf[q])); };
    const bool here = free_of_via_block(i);
    if (l > 0 && ((va_bits >> (l - 1)) & 1u) && here && free_of_via_block(i - plane)) m |= M_DN;
    if (l + 1 < layers && ((va_bits >> l) & 1u) && here && free_of_via_block(i + plane)) m |= M_UP;
    bm[i] = m;
    raw[i] = (uint8_t)(((bc & 0x02) ? R_VIANODE : 0) | ((vf && vf[i]) ? R_VF : 0));
  }
  const uint64_t bits = __ballot(ob);
  if ((threadIdx.x & 63) == 0 && i < n) obits[i >> 6] = bits;
}

__global__ void to_u16(const uint32_t* a, uint16_t* o, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) o[i] = (uint16_t)a[i];
}

// The lean wave visit's step prices of every cell, as the visit would form
// them from the masks, the additive plane and the gap prices (wave_step):
// x plane (from x-1 | from x+1), y plane (from y-1 | from y+1), via plane
// (from the layer below | above).  Values past the narrow limit are never
// read (such boards do not run the wave form).
__global__ void __launch_bounds__(256) board_steps(const uint8_t* __restrict__ bm, const uint16_t* __restrict__ add16,
                                                   const uint32_t* __restrict__ add, const uint32_t* __restrict__ xe,
                                                   const uint32_t* __restrict__ ye, const uint32_t* __restrict__ ve,
                                                   uint32_t w, uint32_t h, uint32_t layers, uint32_t* __restrict__ st) {
  const size_t n = (size_t)w * h * layers, plane = (size_t)w * h;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t l = (uint32_t)(i / plane);
    const uint32_t r = (uint32_t)(i - (size_t)l * plane), y = r / w, x = r - y * w;
    const uint8_t m = bm[i];
    const uint32_t a = add16 ? add16[i] : add ? add[i] : 0u;
    const uint32_t xl = x >= 1 ? xe[x - 1] : 0u, xr = x + 1 < w ? xe[x] : 0u;
    const uint32_t yl = y >= 1 ? ye[y - 1] : 0u, yr = y + 1 < h ? ye[y] : 0u;
    const uint32_t vl = l >= 1 ? ve[l - 1] : 0u, vh = l + 1 < layers ? ve[l] : 0u;
    st[i] = pack_steps(wave_step(m, M_NX, true, xl, a), wave_step(m, M_PX, true, xr, a));
    st[n + i] = pack_steps(wave_step(m, M_NY, true, yl, a), wave_step(m, M_PY, true, yr, a));
    st[2 * n + i] = pack_steps(wave_step(m, M_DN, l > 0, vl, a), wave_step(m, M_UP, l + 1 < layers, vh, a));
  }
}

// (Re)build a board's step price planes on stream s, if it has them.
static void build_steps(hipStream_t s, const DevBoard& b) {
  const size_t n = (size_t)b.w * b.h * b.layers;
  if (!b.steps || !n) return;
  const uint32_t blocks = (uint32_t)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(board_steps, dim3(blocks), dim3(256), 0, s, b.bmask, b.add16, b.add, b.xe, b.ye, b.ve, b.w, b.h,
                     b.layers, b.steps);
  HIP_OK(hipGetLastError());
}

// Activate items idx[0..n) for the first round: flags and list.
__global__ void seed_list(uint32_t* act, uint32_t* list, const uint32_t* idx, uint32_t n, uint32_t* key) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    act[idx[i]] = 1u;
    list[i] = idx[i];
    if (key) key[idx[i]] = 0u;  // sources start at search cost 0
  }
}

// Activate items idx[0..n) for a plain-form round: flags, and each item
// appended to its sub-list (count set `set`) of the sharded list.
__global__ void seed_list_sharded(uint32_t* act, uint32_t* list, uint32_t lcap, uint32_t* shards, uint32_t set,
                                  const uint32_t* idx, uint32_t n, uint32_t* key) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint32_t it = idx[i], s = it % kShards;
    act[it] = 1u;
    const uint32_t pos = atomicAdd(shards + ((size_t)(SH_L0 + set) * kShards + s) * kCtrStride, 1u);
    list[(size_t)s * lcap + pos] = it;
    if (key) key[it] = 0u;  // sources start at search cost 0
  }
}

// The list table of count set `set` (after seeds were appended to it): one
// workgroup of 64 threads.
__global__ void list_table(uint32_t* shards, uint32_t set) {
  uint32_t run = shards[((size_t)(SH_L0 + set) * kShards + threadIdx.x) * kCtrStride];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t up = (uint32_t)__shfl_up((int)run, o);
    if (threadIdx.x >= (uint32_t)o) run += up;
  }
  shards[kShardCounterWords + threadIdx.x] = run;
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
  uint32_t packed;  // labels are u32 (dist << 16) | hops (wave form), read back as u64
};

// Label of global cell g in field F (kInf outside the window or untouched tile).
__device__ uint64_t field_label(const FieldRead& R, const BoardView& B, const FieldDesc& F, uint32_t g) {
  uint32_t l = g / B.plane, r = g - l * B.plane, gy = r / B.w, gx = r - gy * B.w;
  if (gx < F.x0 || gy < F.y0) return kInf;
  uint32_t wx = gx - F.x0, wy = gy - F.y0;
  if (wx >= F.ww || wy >= F.wh) return kInf;
  uint32_t txi = wx / R.tw, tyi = wy / R.th, t = tyi * F.tx + txi;
  uint32_t x = wx - txi * R.tw, y = wy - tyi * R.th;
  // state and label are loaded together (the label slot exists for every
  // tile of the field); the label counts only if the tile was visited
  const uint32_t st = R.state[F.item0 + t];
  const uint64_t at = F.lab_off + t * tile_block_words(B.layers, R.tw, R.th) + (l * R.th + y) * R.tw + x;
  uint64_t lab;
  if (R.packed) {
    const uint32_t v = ((const uint32_t*)R.labels)[at];
    lab = v >= 0xFFFF0000u ? kUnreached : ((uint64_t)(v >> 16) << 32) | (v & 0xFFFFu);
  } else {
    lab = R.labels[at];
  }
  return (st & S_TOUCHED) ? lab : kInf;
}

// Label of the board cell (layer l, column gx, row gy) in field F, with no
// division: tws / ths are log2 of the label layout's tile width and height
// (every instantiated tile side is a power of two).
__device__ __forceinline__ uint64_t field_label_at(const FieldRead& R, const BoardView& B, const FieldDesc& F,
                                                   uint32_t l, uint32_t gx, uint32_t gy, uint32_t tws, uint32_t ths) {
  const uint32_t wx = gx - F.x0, wy = gy - F.y0;  // wrap below the window: outside
  if (wx >= F.ww || wy >= F.wh) return kInf;
  const uint32_t txi = wx >> tws, tyi = wy >> ths, t = tyi * F.tx + txi;
  const uint32_t x = wx - (txi << tws), y = wy - (tyi << ths);
  const uint32_t st = R.state[F.item0 + t];
  const uint64_t at = F.lab_off + t * tile_block_words(B.layers, R.tw, R.th) + ((l << ths) + y) * R.tw + x;
  uint64_t lab;
  if (R.packed) {
    const uint32_t v = ((const uint32_t*)R.labels)[at];
    lab = v >= 0xFFFF0000u ? kUnreached : ((uint64_t)(v >> 16) << 32) | (v & 0xFFFFu);
  } else {
    lab = R.labels[at];
  }
  return (st & S_TOUCHED) ? lab : kInf;
}

struct WalkArgs {
  FieldRead R;
  const uint32_t* pool;
  const uint64_t* poff;
  const uint32_t* plen;
  uint32_t* cells;
  uint32_t* counters;
  uint32_t n;
  uint64_t cap;  // cells the buffer holds; a path past it sets error bit 64 and is not walked
};

// Path lengths (hops + 1 of the dst's label, 0 when unreachable), search
// costs and exclusive offsets of n fields; poff[n] = total.  One workgroup
// of 1024 threads, 1024 fields per pass.
__global__ void __launch_bounds__(1024) path_scan(FieldRead R, uint32_t n, uint32_t* plen, uint32_t* pcost,
                                                  uint64_t* poff) {
  __shared__ uint64_t wsum[16];
  __shared__ uint64_t base_s;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
  if (tid == 0) base_s = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < n; c0 += 1024) {
    const uint32_t f = c0 + tid;
    uint32_t len = 0, cost = kMax;
    if (f < n) {
      const FieldDesc F = R.fields[f];
      const BoardView B = R.boards[F.board];
      if (F.dst != kMax) {
        const uint64_t lab = field_label(R, B, F, F.dst);
        if ((uint32_t)(lab >> 32) != kMax) {
          len = (uint32_t)lab + 1;
          cost = (uint32_t)(lab >> 32);
        }
      }
      plen[f] = len;
      pcost[f] = cost;
    }
    // inclusive sums within the wave, then across the waves
    uint64_t v = len;
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t up = __shfl_up(v, o);
      if (lane >= (uint32_t)o) v += up;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    uint64_t before = base_s;
    for (uint32_t w = 0; w < wv; ++w) before += wsum[w];
    if (f < n) poff[f] = before + v - len;
    __syncthreads();
    if (tid == 1023) base_s = before + v;
    __syncthreads();
  }
  if (tid == 0) poff[n] = base_s;
}

// Cells within Manhattan distance 3 of a walk's centre: 63 offsets, listed
// in ascending (layer, y, x) order, one per lane (lane 63 idle).  On a board
// every such offset order is the order of the cells' CellIdx, so the lowest
// lane that passes a test is the lowest-CellIdx cell that passes it.
constexpr int kWalkRadius = 3;
constexpr int kWalkCells = 63;

struct WalkOffsets {
  int8_t d[64][3];  // (layer, y, x) offset of lane k
};

constexpr WalkOffsets make_walk_offsets() {
  WalkOffsets t{};
  int k = 0;
  for (int a = -kWalkRadius; a <= kWalkRadius; ++a)
    for (int b = -kWalkRadius; b <= kWalkRadius; ++b)
      for (int c = -kWalkRadius; c <= kWalkRadius; ++c) {
        const int m = (a < 0 ? -a : a) + (b < 0 ? -b : b) + (c < 0 ? -c : c);
        if (m > kWalkRadius) continue;
        t.d[k][0] = (int8_t)a;
        t.d[k][1] = (int8_t)b;
        t.d[k][2] = (int8_t)c;
        ++k;
      }
  return t;
}

constexpr WalkOffsets kWalkHost = make_walk_offsets();
// the centre (0, 0, 0) sits in the middle of the symmetric list
constexpr int kWalkCentre = kWalkCells / 2;
static_assert(kWalkHost.d[kWalkCentre][0] == 0 && kWalkHost.d[kWalkCentre][1] == 0 &&
                  kWalkHost.d[kWalkCentre][2] == 0 && kWalkHost.d[kWalkCells - 1][0] == kWalkRadius,
              "walk offsets");
__constant__ WalkOffsets kWalkOff = make_walk_offsets();

// One wavefront per field.  Each memory round trip loads, for every cell of
// the distance-3 neighbourhood of the current cell, its label, its cell view
// and the edge prices towards +x, +y and the layer above; the walk then takes
// up to three steps inside the neighbourhood from registers.  A step from
// cell c tests c's six neighbours (all inside the neighbourhood while c is at
// most two steps from the centre): neighbour p is c's predecessor when c may
// be entered from p's side and label(p) extended by the price of entering c
// equals label(c); the lowest-CellIdx such p is taken (mr-metal
// lib.rs:308-396, gpu.rs:569-714).
__global__ void walk_paths(WalkArgs W) {
  const uint32_t f = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  if (f >= W.n || W.plen[f] == 0) return;
  if (W.poff[f] + W.plen[f] > W.cap) {
    if (lane == 0) atomicOr(&W.counters[(C_ERROR) * kCtrStride], 64u);  // the buffer is too small
    return;
  }
  const FieldDesc F = W.R.fields[f];
  const BoardView B = W.R.boards[F.board];
  uint32_t* out = W.cells + W.poff[f];
  const bool in_set = lane < (uint32_t)kWalkCells;
  const int dl = kWalkOff.d[lane][0], dy = kWalkOff.d[lane][1], dx = kWalkOff.d[lane][2];
  const uint32_t tws = __ffs(W.R.tw) - 1, ths = __ffs(W.R.th) - 1;
  uint32_t cur = F.dst;
  uint64_t lab = field_label(W.R, B, F, cur);
  uint32_t steps = 0;
  const uint32_t limit = W.plen[f];
  // the centre's coordinates, kept along the walk (no division per round trip)
  uint32_t cl = cur / B.plane, cy = (cur - cl * B.plane) / B.w, cx = cur - cl * B.plane - cy * B.w;
  for (;;) {
    const int pl = (int)cl + dl, py = (int)cy + dy, px = (int)cx + dx;
    const bool on = in_set && pl >= 0 && py >= 0 && px >= 0 && pl < (int)B.layers && py < (int)B.h && px < (int)B.w;
    const uint32_t g = on ? (uint32_t)(((uint32_t)pl * B.h + (uint32_t)py) * B.w + (uint32_t)px) : cur;
    // every load below depends only on the centre: one round trip
    const uint64_t plab = on ? field_label_at(W.R, B, F, (uint32_t)pl, (uint32_t)px, (uint32_t)py, tws, ths) : kInf;
    CellInfo ci;
    ci.mask = 0;
    ci.w = 0;
    ci.a = 0;
    if (on) ci = cell_info(B, F, W.pool, g);
    const uint32_t ex = on && px + 1 < (int)B.w ? B.xe[px] : 0u;
    const uint32_t ey = on && py + 1 < (int)B.h ? B.ye[py] : 0u;
    const uint32_t ev = on && pl + 1 < (int)B.layers ? B.ve[pl] : 0u;
    int c = kWalkCentre;
    for (int k = 0; k < kWalkRadius; ++k) {
      const uint32_t hops = (uint32_t)lab;
      const uint32_t cg = (uint32_t)__builtin_amdgcn_readlane((int)g, c);
      if (lane == 0) out[hops] = cg;
      if (hops == 0) return;
      if (++steps > limit) {
        if (lane == 0) atomicOr(&W.counters[(C_ERROR) * kCtrStride], 32u);
        return;
      }
      const int cdl = __builtin_amdgcn_readlane(dl, c), cdy = __builtin_amdgcn_readlane(dy, c),
                cdx = __builtin_amdgcn_readlane(dx, c);
      const uint32_t cmask = (uint32_t)__builtin_amdgcn_readlane((int)ci.mask, c);
      const uint32_t cw = (uint32_t)__builtin_amdgcn_readlane((int)ci.w, c);
      const uint32_t ca = (uint32_t)__builtin_amdgcn_readlane((int)ci.a, c);
      const uint32_t cex = (uint32_t)__builtin_amdgcn_readlane((int)ex, c);
      const uint32_t cey = (uint32_t)__builtin_amdgcn_readlane((int)ey, c);
      const uint32_t cev = (uint32_t)__builtin_amdgcn_readlane((int)ev, c);
      const int ol = dl - cdl, oy = dy - cdy, ox = dx - cdx;
      uint32_t bit = 0, e = 0;
      if (oy == 0 && ox == 0 && ol == -1) {
        bit = K_DN;
        e = ev;  // edge between this layer and the one above: its own
      } else if (oy == 0 && ox == 0 && ol == 1) {
        bit = K_UP;
        e = cev;
      } else if (ol == 0 && ox == 0 && oy == -1) {
        bit = K_NY;
        e = ey;
      } else if (ol == 0 && ox == 0 && oy == 1) {
        bit = K_PY;
        e = cey;
      } else if (ol == 0 && oy == 0 && ox == -1) {
        bit = K_NX;
        e = ex;
      } else if (ol == 0 && oy == 0 && ox == 1) {
        bit = K_PX;
        e = cex;
      }
      const bool ok = on && bit && (cmask & bit) && (uint32_t)(plab >> 32) != kMax &&
                      extend(plab, step_price(e, cw, ca)) == lab;
      const uint64_t mask = __ballot(ok);
      if (mask == 0) {
        if (lane == 0) atomicOr(&W.counters[(C_ERROR) * kCtrStride], 16u);
        return;
      }
      c = __ffsll((long long)mask) - 1;
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)plab, c);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(plab >> 32), c);
      lab = ((uint64_t)hi << 32) | lo;
    }
    cur = (uint32_t)__builtin_amdgcn_readlane((int)g, c);
    cl = (uint32_t)__builtin_amdgcn_readlane(pl, c);
    cy = (uint32_t)__builtin_amdgcn_readlane(py, c);
    cx = (uint32_t)__builtin_amdgcn_readlane(px, c);
  }
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

// ---------------------------------------------------------------------------
// Allocation-free calls: the library's scratch array.  One copy lives on
// every device the library's code object is loaded on; host code never uses
// the array's name as an address: scratch_ptr() returns the device address
// hipGetSymbolAddress gave for that device, resolved once, with the device
// current, when the first context opens on it.  Calls that overlap on one
// device share the array and race: the caller serialises them.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};
}  // namespace

uint8_t* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return g_scratch[device];
}

size_t scratch_bytes() { return kScratchBytes; }

// Once per device, with that device current.
static void resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(1, "the scratch array has no device address");
  g_scratch[device] = (uint8_t*)p;
}

void arena_begin(Ctx& c) {
  uint8_t* p = scratch_ptr(c.device);
  if (!p) throw Error(1, "no scratch array on this device");
  Workspace& A = c.ws_arena;
  A = Workspace();
  // the context's own buffers, taken from its workspace (never from W(),
  // which is the arena workspace once `arena` is set)
  const Workspace& own = c.ws;
  if (!own.counters || !own.pinned || !own.hostctr || !own.hostctr_dev)
    throw Error(1, "context buffers missing for an allocation-free call");
  A.counters = own.counters;
  A.pinned = own.pinned;
  A.hostctr = own.hostctr;
  A.hostctr_dev = own.hostctr_dev;
  A.ev0 = own.ev0;
  A.ev1 = own.ev1;
  A.ev2 = own.ev2;
  A.ek0 = own.ek0;
  A.ek1 = own.ek1;
  A.stamp = own.stamp;
  c.arena = p;
  c.arena_cap = kScratchBytes;
  c.arena_used = 0;
}

void arena_end(Ctx& c) {
  if (!c.arena) return;
  c.ws.stamp = c.ws_arena.stamp;  // launch stamps keep increasing across calls
  c.ws_arena = Workspace();
  c.arena = nullptr;
  c.arena_cap = c.arena_used = 0;
}

// A device buffer of at least `need` elements: grown with hipMalloc in the
// context's own workspace, or, in an allocation-free call, carved from the
// scratch array (256-byte aligned; too little room is an error).
template <typename T>
static void take(Ctx& c, T*& p, size_t& cap, size_t need) {
  if (!c.arena) {
    grow(p, cap, need);
    return;
  }
  const size_t bytes = (need * sizeof(T) + 255) & ~size_t(255);
  if (bytes > c.arena_cap - c.arena_used)
    throw Error(1, "the call needs more than the " + std::to_string(c.arena_cap >> 20) +
                       " MiB scratch array of the allocation-free entry point");
  p = (T*)(c.arena + c.arena_used);
  c.arena_used += bytes;
  cap = need;
}

// In an allocation-free call: every buffer a kernel of the solve will take is
// set, and each carved one lies inside the scratch array.
static void check_arena(const Ctx& c, std::initializer_list<std::pair<const void*, size_t>> carved) {
  if (!c.arena) return;
  if (!c.ws_arena.counters || !c.ws_arena.hostctr_dev) throw Error(1, "allocation-free call without counters");
  for (const auto& b : carved) {
    const uint8_t* q = (const uint8_t*)b.first;
    if (!q || q < c.arena || q + b.second > c.arena + c.arena_cap)
      throw Error(1, "allocation-free buffer outside the scratch array");
  }
}

template <typename T>
static T* dev_copy(const T* h, size_t n, hipStream_t s) {
  T* d = nullptr;
  HIP_OK(hipMalloc((void**)&d, std::max<size_t>(n, 1) * sizeof(T) + 16));
  if (n) HIP_OK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

// Instantiated tile shapes (width x height), layer bounds, step widths and
// lean variants.
struct KernelSet {
  uint32_t tw, th, lm;
  bool narrow, lean;
  const void* round;
  const void* list;
  const void* queue;
  // relax_wave forming its prices, for this shape and layer bound (narrow
  // sets only, else nullptr), by wave_index(packed labels, halo-owner
  // planes, per-net extras); packed labels never run with halo-owner planes
  const void* wave[8];
  // relax_wave_queue, [2 * step planes + packed labels] (narrow sets up to
  // 8 layers at 16 x 16, else nullptr)
  const void* wave_queue[4];
  // relax_wave on the boards' step price planes (lean launches), u64 and
  // packed labels
  const void* wave_sp[2];
};

static constexpr int wave_index(bool pk, bool ho, bool ex) { return (pk ? 4 : 0) + (ho ? 2 : 0) + (ex ? 1 : 0); }

template <int TW, int TH, int LM, bool NARROW, bool LEAN>
static KernelSet make_set() {
  KernelSet k{TW, TH, LM, NARROW, LEAN, (const void*)relax_round<TW, TH, LM, NARROW, LEAN>,
              (const void*)relax_list<TW, TH, LM, NARROW, LEAN>, (const void*)relax_queue<TW, TH, LM, NARROW, LEAN>,
              {}, {}, {}};
  if constexpr (NARROW && TW == 16 && TH == 16 && LM <= 8) {
    k.wave_sp[0] = (const void*)relax_wave<TW, TH, LM, false, false, false, true>;
    k.wave_sp[1] = (const void*)relax_wave<TW, TH, LM, true, false, false, true>;
    k.wave_queue[0] = (const void*)relax_wave_queue<TW, TH, LM, false, false>;
    k.wave_queue[1] = (const void*)relax_wave_queue<TW, TH, LM, true, false>;
    k.wave_queue[2] = (const void*)relax_wave_queue<TW, TH, LM, false, true>;
    k.wave_queue[3] = (const void*)relax_wave_queue<TW, TH, LM, true, true>;
    k.wave[wave_index(false, false, false)] = (const void*)relax_wave<TW, TH, LM, false, false, false>;
    k.wave[wave_index(false, false, true)] = (const void*)relax_wave<TW, TH, LM, false, false, true>;
    k.wave[wave_index(false, true, false)] = (const void*)relax_wave<TW, TH, LM, false, true, false>;
    k.wave[wave_index(false, true, true)] = (const void*)relax_wave<TW, TH, LM, false, true, true>;
    k.wave[wave_index(true, false, false)] = (const void*)relax_wave<TW, TH, LM, true, false, false>;
    k.wave[wave_index(true, false, true)] = (const void*)relax_wave<TW, TH, LM, true, false, true>;
  }
  return k;
}

// Wide step planes for every layer bound; narrow ones, full and lean, up to
// 8 layers.
template <int TW, int TH, bool WITH_NARROW>
static void add_shape(std::vector<KernelSet>& v) {
  v.push_back(make_set<TW, TH, 2, false, false>());
  v.push_back(make_set<TW, TH, 4, false, false>());
  v.push_back(make_set<TW, TH, 8, false, false>());
  v.push_back(make_set<TW, TH, 16, false, false>());
  if (WITH_NARROW) {
    v.push_back(make_set<TW, TH, 2, WITH_NARROW, false>());
    v.push_back(make_set<TW, TH, 4, WITH_NARROW, false>());
    v.push_back(make_set<TW, TH, 8, WITH_NARROW, false>());
    v.push_back(make_set<TW, TH, 2, WITH_NARROW, WITH_NARROW>());
    v.push_back(make_set<TW, TH, 4, WITH_NARROW, WITH_NARROW>());
    v.push_back(make_set<TW, TH, 8, WITH_NARROW, WITH_NARROW>());
  }
}

static const std::vector<KernelSet>& kernel_sets() {
  static const std::vector<KernelSet> v = [] {
    std::vector<KernelSet> k;
    add_shape<32, 32, true>(k);
    add_shape<32, 16, true>(k);
    add_shape<16, 16, true>(k);
    add_shape<16, 8, true>(k);
    add_shape<8, 4, true>(k);
    add_shape<4, 1, false>(k);
    return k;
  }();
  return v;
}

size_t lds_bytes_for(uint32_t tw, uint32_t th, uint32_t layers) { return lds_bytes(tw, th, layers, false); }

bool tile_shape_supported(uint32_t tw, uint32_t th) {
  for (const KernelSet& k : kernel_sets())
    if (k.tw == tw && k.th == th) return true;
  return false;
}

// The narrow variant when asked for and instantiated (lean when asked for
// and instantiated), else the wide one.
static const KernelSet& pick_kernels(uint32_t tw, uint32_t th, uint32_t layers, bool narrow, bool lean) {
  const uint32_t lm = layers <= 2 ? 2 : layers <= 4 ? 4 : layers <= 8 ? 8 : 16;
  const KernelSet *wide = nullptr, *full = nullptr;
  for (const KernelSet& k : kernel_sets()) {
    if (k.tw != tw || k.th != th || k.lm != lm) continue;
    if (narrow && k.narrow && k.lean == lean) return k;
    if (k.narrow && !k.lean) full = &k;
    if (!k.narrow) wide = &k;
  }
  if (narrow && full) return *full;
  if (wide) return *wide;
  throw Error(1, "tile shape " + std::to_string(tw) + "x" + std::to_string(th) + " is not instantiated");
}

// The round-form kernel of a launch: the wave form (one wavefront per tile
// up to four layers, two up to eight) when the shape asks for 64 threads and
// the launch qualifies -- a narrow set with a wave kernel, fields the wave
// form covers (field_wave) on boards without a weight plane, no delta
// stepping -- else relax_round with the shape's workgroup size.
// 16 x 16 tiles at 64 threads select the wave form; launches it does not
// cover (and the graph and queue forms) run the workgroup kernels at 256.
static bool wave_shape(const TileShape& S) { return S.threads == 64 && S.tw == 16 && S.th == 16; }
static uint32_t workgroup_threads(const TileShape& S) { return wave_shape(S) ? 256u : S.threads; }
// `pack`: the batch may use packed u32 labels (every open step of its boards
// costs at least 1); the wave form then runs them.  `halo`: some board of the
// launch has a halo-owner plane (never packed).  `ex`: some field carries own
// pads, via-exempt pads or additive overrides.
struct RoundLaunch {
  const void* fn;
  uint32_t threads;
  size_t lds;
  bool wave, packed;
};
static RoundLaunch round_launch(const KernelSet& K, const TileShape& S, uint32_t Lmax, bool wave_ok, bool keyed,
                                bool pack, bool halo, bool ex, bool sp) {
  if (halo) pack = false;
  // lean launches whose boards all carry step price planes load their
  // prices, the others form them
  const void* w = (sp && !halo && !ex) ? K.wave_sp[pack ? 1 : 0] : K.wave[wave_index(pack, halo, ex)];
  if (wave_shape(S) && w && K.narrow && wave_ok && !keyed && Lmax <= K.lm) {
    const uint32_t nt = 64u * wave_count(K.lm, S.th);
    return RoundLaunch{w, nt, wave_lds_bytes(S.tw, S.th, K.lm, pack, ex), true, pack};
  }
  return RoundLaunch{K.round, workgroup_threads(S), lds_bytes(S.tw, S.th, Lmax, K.narrow), false, false};
}

// Whether a lean wave visit on this board loads its step prices from the
// board's step price planes rather than forming them.  Measured: loading wins
// where the board has an additive plane (the formed visit loads it and adds
// it into every price; C2, C3) or where the planes (12 B per cell) stay in an
// XCD's L2 (C1: 13 % less kernel time); on a board without an additive plane
// whose planes do not fit L2, forming from the 1 B mask is cheaper (C2 board
// through the provider seam: 1.31 ms of relaxation formed, 1.51 ms loaded).
static bool use_steps(const BoardView& v) {
  return v.steps && (v.add || v.add16 || (uint64_t)v.plane * v.layers * 12 <= (2ull << 20));
}

// Whether a field and its board need none of what the lean kernels leave
// out: halo-owner and legal-search planes, additive overrides, own pads and
// via-exempt pads.
static bool field_lean(const BoardView& v, const FieldReq& q) {
  return !v.halo && !v.owner && !v.lhalo && q.ovr_cells.empty() && q.pads.empty() && q.via_pads.empty();
}

// Whether the wave form covers a field: no legal-search planes, and its
// board has no weight plane (halo-owner planes, own pads, via-exempt pads and
// overrides it applies per tile).
static bool field_wave(const BoardView& v, const FieldReq& q) { return !v.cost && !v.owner && !v.lhalo; }

// Whether a field carries what the wave form applies per tile: own pads,
// via-exempt pads, additive overrides.
static bool field_extras(const FieldReq& q) {
  return !q.pads.empty() || !q.via_pads.empty() || !q.ovr_cells.empty();
}

// Own pads and via-exempt pads of a field by tile, for the wave form:
// own-pad starts per tile [tx*ty+1], via-pad starts [tx*ty+1], then each
// kind's entries ordered by tile, as the cell's index in its tile block
// ((layer * TH + y) * TW + x).  Cells outside the window are dropped.
static void build_pad_tables(const FieldReq& q, const DevBoard& b, const TileShape& S, uint32_t tx, uint32_t ty,
                             std::vector<uint32_t>& seg) {
  seg.clear();
  if (q.pads.empty() && q.via_pads.empty()) return;
  const uint32_t plane = b.w * b.h, nt = tx * ty;
  std::vector<uint32_t> starts[2], ent[2];
  const std::vector<uint32_t>* src[2] = {&q.pads, &q.via_pads};
  for (int kind = 0; kind < 2; ++kind) {
    std::vector<std::pair<uint32_t, uint32_t>> v;  // (tile, index in the tile block)
    for (uint32_t g : *src[kind]) {
      if ((size_t)g >= (size_t)plane * b.layers) continue;
      const uint32_t l = g / plane, r = g % plane, gy = r / b.w, gx = r % b.w;
      if (gx < q.x0 || gx > q.x1 || gy < q.y0 || gy > q.y1) continue;
      const uint32_t wx = gx - q.x0, wy = gy - q.y0;
      v.push_back({(wy / S.th) * tx + wx / S.tw, (l * S.th + wy % S.th) * S.tw + wx % S.tw});
    }
    std::sort(v.begin(), v.end());
    starts[kind].assign(nt + 1, 0);
    for (const auto& e : v) starts[kind][e.first + 1]++;
    for (uint32_t t = 0; t < nt; ++t) starts[kind][t + 1] += starts[kind][t];
    for (const auto& e : v) ent[kind].push_back(e.second);
  }
  for (int kind = 0; kind < 2; ++kind) seg.insert(seg.end(), starts[kind].begin(), starts[kind].end());
  for (int kind = 0; kind < 2; ++kind) seg.insert(seg.end(), ent[kind].begin(), ent[kind].end());
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
  HIP_OK(hipEventCreate(&c.ws.ev0));
  HIP_OK(hipEventCreate(&c.ws.ev1));
  HIP_OK(hipEventCreate(&c.ws.ek0));
  HIP_OK(hipEventCreate(&c.ws.ek1));
  HIP_OK(hipEventCreate(&c.ws.ev2));
  size_t total = prop.totalGlobalMem;
  c.label_budget = std::min<size_t>(c.label_budget, total / 2);
  c.label_budget_default = c.label_budget;
  HIP_OK(hipMalloc((void**)&c.ws.counters, (kCtrWords + kShardWords) * sizeof(uint32_t)));
  HIP_OK(hipMalloc((void**)&c.upstat, 64));
  HIP_OK(hipHostMalloc((void**)&c.ws.pinned, (kCtrWords + kShardWords) * sizeof(uint32_t), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&c.ws.hostctr, 64, hipHostMallocMapped | hipHostMallocCoherent));
  memset(c.ws.hostctr, 0, 64);
  HIP_OK(hipHostGetDevicePointer((void**)&c.ws.hostctr_dev, c.ws.hostctr, 0));
  resolve_scratch(device);
  for (const KernelSet& k : kernel_sets()) {
    HIP_OK(hipFuncSetAttribute(k.round, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64));
    HIP_OK(hipFuncSetAttribute(k.list, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64));
    HIP_OK(hipFuncSetAttribute(k.queue, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 128));
    for (const void* w : k.wave)
      if (w) HIP_OK(hipFuncSetAttribute(w, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64));
    for (const void* w : k.wave_queue)
      if (w) HIP_OK(hipFuncSetAttribute(w, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 128));
    for (const void* w : k.wave_sp)
      if (w) HIP_OK(hipFuncSetAttribute(w, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64));
  }
}

void ctx_free(Ctx& c) {
  Workspace& w = c.ws;
  void* ptrs[] = {w.labels, w.state, w.act, w.key, w.list, w.fields, w.item_field, w.pool, w.counters,
                  w.plen, w.poff, w.pcost, w.pcells, w.fout, w.boards, w.qslots, w.qctl, w.fmax};
  (void)hipSetDevice(c.device);
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (void* p : {(void*)c.stage, (void*)c.tblock, (void*)c.upstat})
    if (p) (void)hipFree(p);
  c.stage = c.tblock = nullptr;
  c.upstat = nullptr;
  c.stage_cap = c.tblock_cap = 0;
  if (w.pinned) (void)hipHostFree(w.pinned);
  if (w.hpath) (void)hipHostFree(w.hpath);
  if (w.hostctr) (void)hipHostFree(w.hostctr);
  if (c.graph.exec) (void)hipGraphExecDestroy(c.graph.exec);
  c.graph = GraphCache();
  if (w.ev0) (void)hipEventDestroy(w.ev0);
  if (w.ev1) (void)hipEventDestroy(w.ev1);
  if (w.ek0) (void)hipEventDestroy(w.ek0);
  if (w.ek1) (void)hipEventDestroy(w.ek1);
  if (w.ev2) (void)hipEventDestroy(w.ev2);
  for (hipEvent_t e : w.round_ev) (void)hipEventDestroy(e);
  w = Workspace();
  if (c.stream) (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

static void launch_fill32(hipStream_t s, uint32_t* p, uint32_t v, size_t n) {
  if (!n) return;
  uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_u32, dim3(blocks), dim3(256), 0, s, p, v, n);
}

void board_free(DevBoard& b) {
  // cost, bmask, raw and the edge prices live in the block
  void* ptrs[] = {b.add, b.add16, b.halo, b.owner, b.lhalo};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b.lease)
    *b.lease = false;
  else if (b.block)
    (void)hipFree(b.block);
  b = DevBoard();
}

void board_upload(Ctx& c, DevBoard& b, uint32_t w, uint32_t h, uint32_t layers, const uint32_t* cost,
                  const uint8_t* via_forbidden, const uint8_t* board, const uint32_t* xe,
                  const uint32_t* ye, const uint32_t* ve, const uint8_t* via_allowed,
                  const uint32_t* add, const int32_t* halo, uint32_t halo_pen, bool transient) {
  if (layers > 16) throw Error(1, "at most 16 layers are supported");
  HIP_OK(hipSetDevice(c.device));
  board_free(b);
  hipStream_t s = c.stream;
  const size_t n = (size_t)w * h * layers;
  // 1. the caller's planes into the staging buffer (16 B of slack after each
  // piece: kernels read planes with aligned dword loads)
  size_t off = 0;
  auto carve = [](size_t& o, size_t bytes) {
    const size_t at = o;
    o += (bytes + 16 + 255) & ~size_t(255);
    return at;
  };
  const size_t s_cost = carve(off, n * 4);
  const size_t s_board = board ? carve(off, n) : 0;
  const size_t s_vf = via_forbidden ? carve(off, n) : 0;
  const size_t s_add = add ? carve(off, n * 4) : 0;
  grow(c.stage, c.stage_cap, off);
  const uint32_t* d_cost = (const uint32_t*)(c.stage + s_cost);
  const uint8_t* d_board = board ? c.stage + s_board : nullptr;
  const uint8_t* d_vf = via_forbidden ? c.stage + s_vf : nullptr;
  const uint32_t* d_add = add ? (const uint32_t*)(c.stage + s_add) : nullptr;
  HIP_OK(hipMemcpyAsync(c.stage + s_cost, cost, n * 4, hipMemcpyHostToDevice, s));
  if (board) HIP_OK(hipMemcpyAsync(c.stage + s_board, board, n, hipMemcpyHostToDevice, s));
  if (via_forbidden) HIP_OK(hipMemcpyAsync(c.stage + s_vf, via_forbidden, n, hipMemcpyHostToDevice, s));
  if (add) HIP_OK(hipMemcpyAsync(c.stage + s_add, add, n * 4, hipMemcpyHostToDevice, s));
  // 2. weight and price bounds
  launch_fill32(s, c.upstat, 0u, 5);
  hipLaunchKernelGGL(board_scan, dim3((uint32_t)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 1024))),
                     dim3(256), 0, s, d_cost, d_add, n, c.upstat);
  uint32_t* st = c.ws.pinned;
  HIP_OK(hipMemcpyAsync(st, c.upstat, 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  const bool passable = st[2] != 0;
  const uint64_t max_w = std::max<uint64_t>(1, st[0]), max_plane = st[1];
  const uint64_t min_w = passable ? (uint64_t)(uint32_t)~st[3] : UINT64_MAX;
  const uint64_t min_a = passable ? (uint64_t)(uint32_t)~st[4] : UINT64_MAX;
  // every passable cell weighs 1: no weight plane on the device
  const bool unit_w = !passable || (st[0] == 1u && min_w == 1u);
  const uint64_t max_a = n ? max_plane + (halo ? halo_pen : 0u) : 0u;
  // Bound on every step price min(E*w + a, MAX-1): the largest edge price
  // times the largest enter weight (obstacles count as weight 1, the price of
  // an own pad) plus the largest additive price.  Boards under the narrow
  // limit relax with 16-bit step planes in LDS.
  uint64_t max_e = 0, min_e = UINT64_MAX;
  for (uint32_t k = 0; k + 1 < w; ++k) max_e = std::max<uint64_t>(max_e, xe[k]), min_e = std::min<uint64_t>(min_e, xe[k]);
  for (uint32_t k = 0; k + 1 < h; ++k) max_e = std::max<uint64_t>(max_e, ye[k]), min_e = std::min<uint64_t>(min_e, ye[k]);
  for (uint32_t k = 0; k + 1 < layers; ++k)
    if (via_allowed[k] && ve) max_e = std::max<uint64_t>(max_e, ve[k]), min_e = std::min<uint64_t>(min_e, ve[k]);
  b.max_ew = max_e * max_w;
  b.max_step = std::min<uint64_t>(b.max_ew + max_a, kMax - 1);
  // no edge or no passable cell: no step exists, any bound holds
  b.min_ew = (min_e == UINT64_MAX || min_w == UINT64_MAX) ? 0 : min_e * min_w;
  b.min_step = b.min_ew + (min_a == UINT64_MAX ? 0 : min_a);
  b.w = w;
  b.h = h;
  b.layers = layers;
  // 3. the static planes in one block
  size_t boff = 0;
  const size_t o_bm = carve(boff, n), o_raw = carve(boff, n);
  const size_t words = (n + 63) / 64;
  const size_t o_bits = carve(boff, words * 8);
  const size_t o_edges = carve(boff, ((size_t)w + h + 16) * 4);
  const size_t o_cost = unit_w ? 0 : carve(boff, n * 4);
  // step price planes for the lean wave visit: boards without a weight or a
  // halo-owner plane, where the visit would load them (use_steps: an
  // additive plane, or planes that fit an XCD's L2)
  const bool with_steps = unit_w && !halo && n && (add || n * 12 <= (2ull << 20));
  const size_t o_steps = with_steps ? carve(boff, n * 12) : 0;
  if (transient && !c.tblock_busy) {
    grow(c.tblock, c.tblock_cap, boff);
    b.block = c.tblock;
    b.lease = &c.tblock_busy;
    c.tblock_busy = true;
  } else {
    HIP_OK(hipMalloc((void**)&b.block, boff));
  }
  b.bmask = b.block + o_bm;
  b.raw = b.block + o_raw;
  uint64_t* d_bits = (uint64_t*)(b.block + o_bits);
  b.xe = (uint32_t*)(b.block + o_edges);
  b.ye = b.xe + w;
  b.ve = b.ye + h;
  uint32_t va_bits = 0;
  b.via_allowed.assign(16, 0);
  std::vector<uint32_t> edges((size_t)w + h + 16, 0u);
  for (uint32_t k = 0; k + 1 < w; ++k) edges[k] = xe[k];
  for (uint32_t k = 0; k + 1 < h; ++k) edges[(size_t)w + k] = ye[k];
  for (uint32_t k = 0; k + 1 < layers; ++k) {
    edges[(size_t)w + h + k] = ve ? ve[k] : 0;
    b.via_allowed[k] = via_allowed[k] ? 1 : 0;
    if (via_allowed[k]) va_bits |= 1u << k;
  }
  HIP_OK(hipMemcpyAsync(b.xe, edges.data(), edges.size() * 4, hipMemcpyHostToDevice, s));
  if (n)
    hipLaunchKernelGGL(board_build, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, d_cost, d_board, d_vf, w, h,
                       layers, va_bits, b.bmask, b.raw, d_bits);
  if (!unit_w) {
    b.cost = (uint32_t*)(b.block + o_cost);
    HIP_OK(hipMemcpyAsync(b.cost, d_cost, n * 4, hipMemcpyDeviceToDevice, s));
  }
  b.bytes = n * ((unit_w ? 0 : 4) + 1 + 1 + (with_steps ? 12 : 0));
  const uint32_t cblocks = (uint32_t)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096));
  if (add && max_plane <= 0xFFFFu) {
    // additive prices that fit 16 bits travel as u16 (2 B per cell visit)
    HIP_OK(hipMalloc((void**)&b.add16, n * 2 + 16));
    hipLaunchKernelGGL(to_u16, dim3(cblocks), dim3(256), 0, s, d_add, b.add16, n);
    b.bytes += n * 2;
  } else if (add) {
    HIP_OK(hipMalloc((void**)&b.add, n * 4 + 16));
    HIP_OK(hipMemcpyAsync(b.add, d_add, n * 4, hipMemcpyDeviceToDevice, s));
    b.bytes += n * 4;
  }
  if (halo) {
    b.halo = dev_copy(halo, n, s);
    b.bytes += n * 4;
  }
  b.halo_pen = halo_pen;
  if (with_steps) {
    b.steps = (uint32_t*)(b.block + o_steps);
    build_steps(s, b);
  }
  b.obst_bits.resize(words);
  if (words) HIP_OK(hipMemcpyAsync(b.obst_bits.data(), d_bits, words * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(s));
}

void board_set_legal(Ctx& c, DevBoard& b, const int32_t* owner, const int32_t* halo) {
  HIP_OK(hipSetDevice(c.device));
  const size_t n = (size_t)b.w * b.h * b.layers;
  auto put = [&](int32_t*& dst, const int32_t* src) {
    if (!src) return;
    if (!dst) {
      HIP_OK(hipMalloc((void**)&dst, n * sizeof(int32_t)));
      b.bytes += n * 4;
    }
    HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  };
  put(b.owner, owner);
  put(b.lhalo, halo);
  HIP_OK(hipStreamSynchronize(c.stream));
  b.legal_owner = owner != nullptr;
  b.legal_halo = halo != nullptr;
  b.legal = true;
}

void board_clear_legal(DevBoard& b) { b.legal = false; }

void board_clear_additive(Ctx& c, DevBoard& b) {
  HIP_OK(hipSetDevice(c.device));
  if (b.add || b.add16) HIP_OK(hipStreamSynchronize(c.stream));  // no solve reads the plane any more
  if (b.add) {
    (void)hipFree(b.add);
    b.add = nullptr;
    b.bytes -= (size_t)b.w * b.h * b.layers * 4;
  }
  if (b.add16) {
    (void)hipFree(b.add16);
    b.add16 = nullptr;
    b.bytes -= (size_t)b.w * b.h * b.layers * 2;
  }
  b.max_step = std::min<uint64_t>(b.max_ew + (b.halo ? b.halo_pen : 0u), kMax - 1);
  b.min_step = b.min_ew;
  build_steps(c.stream, b);
  HIP_OK(hipStreamSynchronize(c.stream));
}

void board_set_additive(Ctx& c, DevBoard& b, const uint32_t* add, uint64_t max_add) {
  HIP_OK(hipSetDevice(c.device));
  const size_t n = (size_t)b.w * b.h * b.layers;
  uint64_t max_a = max_add != UINT64_MAX ? max_add : *std::max_element(add, add + n);
  if (b.halo) max_a += b.halo_pen;
  if (b.add16) {
    // a replaced plane is written as u32 (negotiation prices grow past 16 bits)
    HIP_OK(hipStreamSynchronize(c.stream));
    (void)hipFree(b.add16);
    b.add16 = nullptr;
    b.bytes -= n * 2;
  }
  if (!b.add) {
    b.add = dev_copy(add, n, c.stream);
    b.bytes += n * 4;
  } else {
    HIP_OK(hipMemcpyAsync(b.add, add, n * sizeof(uint32_t), hipMemcpyHostToDevice, c.stream));
  }
  b.max_step = std::min<uint64_t>(b.max_ew + max_a, kMax - 1);
  b.min_step = b.min_ew + *std::min_element(add, add + n);
  build_steps(c.stream, b);
  HIP_OK(hipStreamSynchronize(c.stream));
}

size_t field_label_words(const TileShape& s, uint32_t ww, uint32_t wh, uint32_t layers) {
  size_t tx = (ww + s.tw - 1) / s.tw, ty = (wh + s.th - 1) / s.th;
  return tx * ty * tile_block_words(layers, s.tw, s.th);
}

static RelaxArgs make_args(Ctx& c, uint32_t n_items) {
  RelaxArgs A;
  memset(&A, 0, sizeof(A));
  Workspace& W = c.W();
  A.boards = W.boards;
  A.fields = W.fields;
  A.item_field = W.item_field;
  A.pool = W.pool;
  A.labels = W.labels;
  A.state = W.state;
  A.act_cur = W.act;
  A.act_next = W.act + n_items;
  A.list_cur = W.list;
  A.list_next = W.list + n_items;
  A.counters = W.counters;
  A.shards = W.counters + kCtrWords;
  A.max_local = c.shape.max_local;
  A.next_ctr = C_NEXTN;
  A.clear_ctr = C_NEXTN;
  A.host_out = nullptr;
  A.key_cur = A.key_next = nullptr;
  A.key_ctr = A.key_clear = C_KEY0;
  A.delta = kMax;
  return A;
}

// Capture (or reuse) `graph_rounds` rounds of relax_list + advance_round.
static hipGraphExec_t graph_for(Ctx& c, const KernelSet& K, uint32_t grid, size_t lds, const RelaxArgs& A) {
  GraphCache& g = c.graph;
  std::vector<uint8_t> key(sizeof(RelaxArgs) + 32);
  memcpy(key.data(), &A, sizeof(RelaxArgs));
  uint64_t extra[4] = {(uint64_t)(uintptr_t)K.list, grid, lds,
                       (uint64_t)workgroup_threads(c.shape) * 1000 + c.graph_rounds};
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
    HIP_OK(hipLaunchKernel(K.list, dim3(grid), dim3(workgroup_threads(c.shape)), args, lds, s));
    hipLaunchKernelGGL(advance_round, dim3(1), dim3(64), 0, s, c.W().counters);
  }
  HIP_OK(hipStreamEndCapture(s, &graph));
  HIP_OK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
  HIP_OK(hipGraphDestroy(graph));
  g.key = key;
  return g.exec;
}

// Lay out a field's additive overrides inside its window in the pool (see
// FieldDesc::ovr_off): by cell for point lookups, and by tile for the loader.
// Override table of one field, independent of where it lands in the pool:
// the n overrides inside the window (cells, then prices, in request order),
// the per-tile offsets (tx * ty + 1 words), then cells and prices grouped by
// tile.
static void build_overrides(const FieldReq& q, const DevBoard& b, const TileShape& S, uint32_t tx, uint32_t ty,
                            std::vector<uint32_t>& seg, uint32_t& n_out) {
  seg.clear();
  n_out = 0;
  if (q.ovr_cells.empty()) return;
  if (q.ovr_add.size() != q.ovr_cells.size()) throw Error(1, "override cells and prices differ in length");
  const uint32_t plane = b.w * b.h, nt = tx * ty;
  std::vector<uint32_t> cells, adds, tile;
  for (size_t k = 0; k < q.ovr_cells.size(); ++k) {
    const uint32_t g = q.ovr_cells[k];
    if (k && g <= q.ovr_cells[k - 1]) throw Error(1, "override cells must be strictly ascending");
    if ((size_t)g >= (size_t)plane * b.layers) throw Error(1, "override cell outside the board");
    const uint32_t r = g % plane, gy = r / b.w, gx = r % b.w;
    if (gx < q.x0 || gx > q.x1 || gy < q.y0 || gy > q.y1) continue;
    cells.push_back(g);
    adds.push_back(q.ovr_add[k]);
    tile.push_back(((gy - q.y0) / S.th) * tx + (gx - q.x0) / S.tw);
  }
  const uint32_t n = (uint32_t)cells.size();
  n_out = n;
  if (!n) return;
  std::vector<uint32_t> order(n);
  for (uint32_t k = 0; k < n; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return tile[a] < tile[c]; });
  seg.reserve((size_t)4 * n + nt + 1);
  seg.insert(seg.end(), cells.begin(), cells.end());
  seg.insert(seg.end(), adds.begin(), adds.end());
  std::vector<uint32_t> tab(nt + 1, 0);
  for (uint32_t k = 0; k < n; ++k) tab[tile[k] + 1]++;
  for (uint32_t t = 0; t < nt; ++t) tab[t + 1] += tab[t];
  seg.insert(seg.end(), tab.begin(), tab.end());
  for (uint32_t k : order) seg.push_back(cells[k]);
  for (uint32_t k : order) seg.push_back(adds[k]);
}

// Override tables of every field of a batch, on up to 16 host threads when the
// batch carries many overrides (negotiation batches: a few hundred cells per
// net); the first error of any field is rethrown here.
static void build_all_overrides(const std::vector<const DevBoard*>& boards, const std::vector<FieldReq>& reqs,
                                const TileShape& S, std::vector<std::vector<uint32_t>>& segs,
                                std::vector<uint32_t>& ns) {
  segs.assign(reqs.size(), {});
  ns.assign(reqs.size(), 0);
  size_t total = 0;
  for (const FieldReq& q : reqs) total += q.ovr_cells.size();
  if (!total) return;
  auto work = [&](size_t a, size_t e) {
    for (size_t i = a; i < e; ++i) {
      const FieldReq& q = reqs[i];
      const uint32_t tx = (q.x1 - q.x0 + S.tw) / S.tw, ty = (q.y1 - q.y0 + S.th) / S.th;
      build_overrides(q, *boards[q.board], S, tx, ty, segs[i], ns[i]);
    }
  };
  // one chunk per 16 K override cells on average, on the shared host workers
  host::parallel_for(reqs.size(), std::max<size_t>(1, reqs.size() * 16384 / total), work);
}

// Whether a field's step prices all fit the narrow step planes: its board's
// bound, and no forced endpoint on an obstacle (priced at the obstacle's
// weight unless it is an own pad).
static bool field_narrow(const DevBoard& b, const FieldReq& q) {
  if (b.max_step > kNarrowMaxStep) return false;
  for (uint32_t a : q.ovr_add)
    if (b.max_ew + a + (b.halo ? b.halo_pen : 0u) > kNarrowMaxStep) return false;
  for (uint32_t g : q.forced)
    if (b.obstacle(g) && !std::binary_search(q.pads.begin(), q.pads.end(), g)) return false;
  return true;
}

void solve_fields(Ctx& c, const std::vector<const DevBoard*>& boards, const std::vector<FieldReq>& reqs,
                  SolveStats* st) {
  HIP_OK(hipSetDevice(c.device));
  auto t0 = std::chrono::steady_clock::now();
  const TileShape& S = c.shape;
  uint32_t Lmax = 1;
  c.views.clear();
  for (const DevBoard* b : boards) {
    Lmax = std::max(Lmax, b->layers);
    c.views.push_back(b->view());
  }
  if ((size_t)Lmax * S.th * S.tw > 16384 || S.tw < 4 || S.th < 1)
    throw Error(1, "tile shape too large for LDS");
  bool narrow = true, lean = true, wave_ok = true, halo = false, ex = false, ovr = false, sp = true;
  for (const FieldReq& q : reqs) {
    if (q.board >= boards.size()) throw Error(1, "field names a board outside the batch");
    sp = sp && use_steps(c.views[q.board]);
    narrow = narrow && field_narrow(*boards[q.board], q);
    lean = lean && field_lean(c.views[q.board], q);
    wave_ok = wave_ok && field_wave(c.views[q.board], q);
    halo = halo || c.views[q.board].halo != nullptr;
    ex = ex || field_extras(q);
    ovr = ovr || !q.ovr_cells.empty();
  }
  const KernelSet& K = pick_kernels(S.tw, S.th, Lmax, narrow, lean);
  size_t lds = lds_bytes(S.tw, S.th, Lmax, K.narrow);
  if (lds > 160 * 1024 - 64) throw Error(1, "tile shape exceeds 160 KiB of LDS");

  std::vector<std::vector<uint32_t>> ovr_seg;
  std::vector<uint32_t> ovr_n;
  build_all_overrides(boards, reqs, S, ovr_seg, ovr_n);
  std::vector<FieldDesc>& layout = c.layout;
  layout.clear();
  std::vector<uint32_t> pool, seeds, seed_field;
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
    F.own_group = q.own_group;
    F.flags = q.flags;
    F.board = q.board;
    uint32_t nt = F.tx * F.ty;
    F.ovr_off = (uint32_t)pool.size();
    F.n_ovr = ovr_n[i];
    pool.insert(pool.end(), ovr_seg[i].begin(), ovr_seg[i].end());
    std::vector<uint32_t> xtab;
    build_pad_tables(q, b, S, F.tx, F.ty, xtab);
    F.xtab_off = (uint32_t)pool.size();
    pool.insert(pool.end(), xtab.begin(), xtab.end());
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
    seed_field.insert(seed_field.end(), st_items.size(), (uint32_t)i);
    item += nt;
    lab_off += (uint64_t)nt * tile_block_words(L, S.tw, S.th);
    field_cells += (uint64_t)F.ww * F.wh * L;
    layout.push_back(F);
  }
  // A batch stays within the label budget; a single field larger than the
  // budget runs alone, up to the default budget (half of device memory).
  if ((reqs.size() > 1 && lab_off * 8 > c.label_budget) || lab_off * 8 > c.label_budget_default)
    throw Error(1, "label working set exceeds the device budget");

  if (c.arena && c.mode != MODE_PLAIN) throw Error(1, "allocation-free calls run the plain form");
  if (c.arena) c.arena_used = 0;  // each solve carves the scratch array afresh
  Workspace& W = c.W();
  take(c, W.labels, W.labels_cap, std::max<size_t>(lab_off, 1));
  take(c, W.state, W.state_cap, std::max<size_t>(item, 1));
  take(c, W.act, W.act_cap, std::max<size_t>((size_t)item * 2, 2));
  const bool keyed = c.mode == MODE_PLAIN && c.delta != kMax;
  if (keyed) take(c, W.key, W.key_cap, std::max<size_t>((size_t)item * 2, 2));
  // two lists (round form: sharded, lshard_words each; graph form: item each)
  const size_t list_words = std::max<size_t>(2 * std::max<uint64_t>(lshard_words(item), item), 2);
  take(c, W.list, W.list_cap, list_words);
  take(c, W.fields, W.fields_cap, std::max<size_t>(reqs.size(), 1));
  take(c, W.item_field, W.item_field_cap, std::max<size_t>(item, 1));
  take(c, W.pool, W.pool_cap, std::max<size_t>(pool.size() + seeds.size(), 1));
  take(c, W.boards, W.boards_cap, std::max<size_t>(c.views.size(), 1));
  check_arena(c, {{W.labels, std::max<size_t>(lab_off, 1) * 8},
                  {W.state, std::max<size_t>(item, 1) * 4},
                  {W.act, std::max<size_t>((size_t)item * 2, 2) * 4},
                  {W.list, list_words * 4},
                  {W.fields, std::max<size_t>(reqs.size(), 1) * sizeof(FieldDesc)},
                  {W.item_field, std::max<size_t>(item, 1) * 4},
                  {W.pool, std::max<size_t>(pool.size() + seeds.size(), 1) * 4},
                  {W.boards, std::max<size_t>(c.views.size(), 1) * sizeof(BoardView)}});
  if (keyed) check_arena(c, {{W.key, std::max<size_t>((size_t)item * 2, 2) * 4}});
  hipStream_t s = c.stream;
  HIP_OK(hipMemcpyAsync(W.boards, c.views.data(), c.views.size() * sizeof(BoardView), hipMemcpyHostToDevice, s));
  if (!layout.empty())
    HIP_OK(hipMemcpyAsync(W.fields, layout.data(), layout.size() * sizeof(FieldDesc),
                          hipMemcpyHostToDevice, s));
  if (item && !layout.empty())
    hipLaunchKernelGGL(fill_item_field, dim3((uint32_t)std::min<size_t>(layout.size(), 4096)), dim3(256), 0, s,
                       W.fields, (uint32_t)layout.size(), W.item_field);
  // the seed items follow the pool (this call synchronises before it returns)
  const size_t pool_n = pool.size();
  pool.insert(pool.end(), seeds.begin(), seeds.end());
  if (!pool.empty())
    HIP_OK(hipMemcpyAsync(W.pool, pool.data(), pool.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  const uint32_t* d_seeds = W.pool + pool_n;
  launch_fill32(s, W.state, 0u, item);
  launch_fill32(s, W.counters, 0u, kCtrWords + kShardWords);

  // Packed u32 labels when every open step of the batch's boards costs at
  // least 1 (hops <= dist then) and the wave form runs; checked after the
  // solve, re-solved in wide labels when a label came within a step of the
  // packed dist limit.
  // With overrides a step costs at least the edge price times the weight
  // (the net's own prices may undercut the plane's) and at most the narrow
  // limit (field_narrow checked the override prices).
  // The queue form runs the wave form's visits (relax_wave_queue) for lean
  // batches of the wave shape; other batches run relax_queue.
  const bool wave_queue = c.mode == MODE_QUEUE && wave_shape(S) && K.wave_queue[0] && K.narrow && wave_ok &&
                          !halo && !ex && Lmax <= K.lm;
  bool pack = !c.no_pack && (c.mode == MODE_PLAIN || wave_queue) && !halo;
  uint64_t pk_step = ovr ? kNarrowMaxStep : 0;
  for (const DevBoard* b : boards) {
    pack = pack && (ovr ? b->min_ew : b->min_step) >= 1 && b->max_step < 0xFFFFu && !b->pack_failed;
    pk_step = std::max<uint64_t>(pk_step, b->max_step);
  }
  RelaxArgs A = make_args(c, item);
  const RoundLaunch RL =
      round_launch(K, S, Lmax, wave_ok, c.mode == MODE_PLAIN && c.delta != kMax, pack, halo, ex, sp);
  // labels of this solve are packed u32
  const bool packed_run = c.mode == MODE_PLAIN ? RL.packed : (wave_queue && pack);
  if (packed_run) {
    take(c, W.fmax, W.fmax_cap, std::max<size_t>(item, 1));
    check_arena(c, {{W.fmax, std::max<size_t>(item, 1) * 4}});
    launch_fill32(s, W.fmax, 0u, item);
    A.item_max = W.fmax;
  }
  SolveStats local;
  local.field_cells = field_cells;
  bool timed = false;
  size_t pairs = 0;  // per-round event pairs (plain form) recorded and not yet read
  auto harvest = [&](bool sync) {
    if (sync) HIP_OK(hipStreamSynchronize(s));
    for (size_t k = 0; k < pairs; ++k) {
      float kms = 0;
      HIP_OK(hipEventElapsedTime(&kms, W.round_ev[2 * k], W.round_ev[2 * k + 1]));
      local.kernel_ms += kms;
    }
    pairs = 0;
  };
  if (item && !seeds.empty()) {
    uint32_t* pinned = W.pinned;
    void* args[] = {&A};
    // Rounds of relax_round over the active-item list until no item is
    // activated.  Every round starts at a kernel boundary, so it reads
    // coherent labels; an empty next list certifies the fixpoint.
    // Per round: one launch, then the host waits for the launch's stamp in
    // host-mapped memory.  The activation count alternates between two
    // counters (the kernel clears the one the host has read), and the
    // launch's last workgroup publishes the counts and then the stamp, so no
    // fill or copy is queued between rounds and the next launch follows as
    // soon as the stamp lands.  A stamp that has not landed within
    // c.stamp_spin_us (20 ms unless set) hands over to a stream
    // synchronisation, which reports a failed launch; after it the stamp must
    // be there.  Per-round kernel times are read from
    // an event pair per round once the last round is done.
    auto plain_rounds = [&](uint32_t& rounds, uint32_t n_active) {
      A.host_out = W.hostctr_dev;
      while (n_active) {
        A.lset_next = rounds % 3;
        A.lset_cur = (rounds + 2) % 3;
        A.lset_clear = (rounds + 1) % 3;
        A.next_ctr = C_NEXT0 + (rounds & 1);
        A.clear_ctr = C_NEXT0 + ((rounds + 1) & 1);
        A.key_ctr = C_KEY0 + (rounds & 1);
        A.key_clear = C_KEY0 + ((rounds + 1) & 1);
        if (++W.stamp == 0) W.stamp = 1;
        A.stamp = W.stamp;
        A.round = rounds;
        // per-round kernel times: in the calls the context times (every
        // call unless set otherwise), never in allocation-free calls, which
        // create no events
        const bool per_round = !c.arena && c.time_rounds;
        if (per_round && pairs == 512) harvest(true);  // bounds the event pool
        const size_t ei = 2 * pairs;
        if (per_round) {
          ++pairs;
          while (W.round_ev.size() < ei + 2) {
            hipEvent_t e = nullptr;
            HIP_OK(hipEventCreate(&e));
            W.round_ev.push_back(e);
          }
          HIP_OK(hipEventRecord(W.round_ev[ei], s));
        }
        HIP_OK(hipLaunchKernel(RL.fn, dim3(n_active), dim3(RL.threads), args, RL.lds, s));
        if (per_round) HIP_OK(hipEventRecord(W.round_ev[ei + 1], s));
        volatile uint32_t* hc = W.hostctr;
        const auto spin0 = std::chrono::steady_clock::now();
        while (hc[2] != A.stamp) {
          if (std::chrono::steady_clock::now() - spin0 >= std::chrono::microseconds(c.stamp_spin_us)) {
            HIP_OK(hipStreamSynchronize(s));
            if (hc[2] != A.stamp) throw Error(1, "relaxation round did not publish its counts");
          }
        }
        std::atomic_thread_fence(std::memory_order_acquire);
        ++local.launches;
        if (RL.wave) ++local.wave_launches;
        if (per_round) ++local.timed_launches;
        ++rounds;
        if (hc[1]) break;
        n_active = hc[0];
        std::swap(A.act_cur, A.act_next);
        std::swap(A.key_cur, A.key_next);
        uint32_t* lc = const_cast<uint32_t*>(A.list_cur);
        A.list_cur = A.list_next;
        A.list_next = lc;
        if (rounds > 1000000) throw Error(1, "relaxation did not converge");
      }
      // the event pairs are read after the solve's closing synchronisation
    };
    uint32_t rounds = 0;
    launch_fill32(s, W.act, 0u, (size_t)item * 2);
    if (c.mode == MODE_PLAIN) {
      if (keyed) {
        // every item starts unkeyed; seeds at key 0, so the first threshold is delta
        launch_fill32(s, W.key, kMax, (size_t)item * 2);
        launch_fill32(s, W.counters + C_KEY0 * kCtrStride, kMax, 1);
        launch_fill32(s, W.counters + (C_KEY0 + 1) * kCtrStride, kMax, 1);
        launch_fill32(s, W.counters + C_THRESH * kCtrStride, c.delta, 1);
        A.key_cur = W.key;
        A.key_next = W.key + item;
        A.delta = c.delta;
      }
      // sharded lists: the seeds fill round 0's list, counted in set 2
      A.lcap = lshard_cap(item);
      A.list_next = W.list + lshard_words(item);
      hipLaunchKernelGGL(seed_list_sharded, dim3(((uint32_t)seeds.size() + 255) / 256), dim3(256), 0, s, W.act,
                         W.list, A.lcap, A.shards, 2u, d_seeds, (uint32_t)seeds.size(),
                         keyed ? W.key : (uint32_t*)nullptr);
      hipLaunchKernelGGL(list_table, dim3(1), dim3(kShards), 0, s, A.shards, 2u);
      HIP_OK(hipEventRecord(W.ev0, s));
      timed = true;
      plain_rounds(rounds, (uint32_t)seeds.size());
    } else if (c.mode == MODE_GRAPH) {
      hipLaunchKernelGGL(seed_list, dim3(((uint32_t)seeds.size() + 255) / 256), dim3(256), 0, s, W.act, W.list,
                         d_seeds, (uint32_t)seeds.size(), (uint32_t*)nullptr);
      launch_fill32(s, W.counters + C_CURN * kCtrStride, (uint32_t)seeds.size(), 1);
      int occ = 0;
      HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, K.list, workgroup_threads(S), lds));
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
      local.launches = rounds;
    } else if (c.mode == MODE_QUEUE) {
      // fields over the queues, longest first onto the least loaded queue
      std::vector<uint32_t> order(layout.size()), fq(layout.size());
      std::iota(order.begin(), order.end(), 0u);
      std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return (uint64_t)layout[a].tx * layout[a].ty > (uint64_t)layout[b].tx * layout[b].ty;
      });
      uint64_t qload[kQueues] = {};
      for (uint32_t i : order) {
        uint32_t k = 0;
        for (uint32_t j = 1; j < kQueues; ++j)
          if (qload[j] < qload[k]) k = j;
        fq[i] = k;
        qload[k] += (uint64_t)layout[i].tx * layout[i].ty;
      }
      QueueArgs Q;
      memset(&Q, 0, sizeof(Q));
      uint64_t nslots = 0;
      for (uint32_t k = 0; k < kQueues; ++k) {
        // an item is queued at most once at a time
        Q.slot0[k] = (uint32_t)nslots;
        Q.cap[k] = (uint32_t)(qload[k] + 64);
        nslots += Q.cap[k];
      }
      if (nslots >= (1ull << 32)) throw Error(1, "too many tiles for the queue form");
      std::vector<uint64_t> slots(nslots, 0ull);
      std::vector<uint32_t> ctl((size_t)3 * kQueues * kCtrStride, 0u);
      for (size_t j = 0; j < seeds.size(); ++j) {
        const uint32_t k = fq[seed_field[j]];
        uint32_t& tail = ctl[(3 * k + 1) * kCtrStride];
        slots[Q.slot0[k] + tail] = wq::entry(tail, Q.cap[k], seeds[j] + 1u);  // ticket `tail`
        ++tail;
        ctl[(3 * k + 2) * kCtrStride] = tail;  // pending: one per queued seed
      }
      take(c, W.qslots, W.qslots_cap, nslots);
      take(c, W.qctl, W.qctl_cap, ctl.size());
      HIP_OK(hipMemcpyAsync(W.qslots, slots.data(), nslots * 8, hipMemcpyHostToDevice, s));
      HIP_OK(hipMemcpyAsync(W.qctl, ctl.data(), ctl.size() * 4, hipMemcpyHostToDevice, s));
      Q.slots = W.qslots;
      Q.ctl = W.qctl;
      Q.wait_ticks = (uint64_t)c.queue_wait_ms * 100000ull;  // s_memrealtime runs at 100 MHz
      Q.run_ticks = (uint64_t)c.queue_run_ms * 100000ull;
      // every label of the solve starts unreached: any label a visit reads is
      // an upper bound of its final value, whichever workgroup wrote it
      launch_fill32(s, (uint32_t*)W.labels, 0xFFFFFFFFu, (size_t)lab_off * 2);
      hipLaunchKernelGGL(seed_list, dim3(((uint32_t)seeds.size() + 255) / 256), dim3(256), 0, s, W.act, W.list,
                         d_seeds, (uint32_t)seeds.size(), (uint32_t*)nullptr);
      const void* qfn = wave_queue ? K.wave_queue[(sp ? 2 : 0) + (pack ? 1 : 0)] : K.queue;
      const uint32_t qthreads = wave_queue ? 64u * wave_count(K.lm, S.th) : workgroup_threads(S);
      const size_t qlds = wave_queue ? wave_lds_bytes(S.tw, S.th, K.lm, pack, false) : lds;
      int occ = 0;
      HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, qfn, qthreads, qlds));
      if (occ < 1) throw Error(1, "relaxation kernel does not fit on a CU");
      const uint32_t grid = std::max<uint32_t>(kQueues, ((uint32_t)occ * (uint32_t)c.cus) / kQueues * kQueues);
      void* qargs[] = {&A, &Q};
      HIP_OK(hipEventRecord(W.ev0, s));
      timed = true;
      HIP_OK(hipLaunchKernel(qfn, dim3(grid), dim3(qthreads), qargs, qlds, s));
      HIP_OK(hipGetLastError());
      rounds = 1;
      local.launches = 1;
      if (wave_queue) local.wave_launches = 1;
    } else {
      throw Error(1, "the persistent grid-barrier form is disabled (it did not return on MI355X); "
                     "use the plain, graph or queue form");
    }
    local.rounds = rounds;
  }
  if (timed) HIP_OK(hipEventRecord(W.ev1, s));
  if (packed_run && item)
    hipLaunchKernelGGL(reduce_item_max, dim3((uint32_t)std::min<size_t>((item + 255) / 256, 256)), dim3(256), 0, s,
                       W.fmax, (size_t)item, W.counters + C_PMAX * kCtrStride);
  uint32_t* cnt = W.pinned;
  HIP_OK(hipMemcpyAsync(cnt, W.counters, (kCtrWords + kShardWords) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  harvest(false);
  if (cnt[C_ERROR * kCtrStride]) throw Error(1, "device relaxation reported error flags " + std::to_string(cnt[C_ERROR * kCtrStride]));
  if (packed_run && (uint64_t)cnt[C_PMAX * kCtrStride] + ((pk_step << 16) | 1u) >= 0xFFFF0000ull) {
    // a path came within one step of the packed dist limit: the labels may
    // be clipped, so the batch is solved again in wide labels, and so are
    // later batches on these boards
    for (const DevBoard* b : boards) b->pack_failed = true;
    c.no_pack = true;
    try {
      solve_fields(c, boards, reqs, st);
    } catch (...) {
      c.no_pack = false;
      throw;
    }
    c.no_pack = false;
    return;
  }
  c.packed = packed_run;
  float ms = 0;
  if (timed) HIP_OK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
  local.relax_ms = ms;
  local.tile_visits = shard_sum(cnt + kCtrWords, SH_VISITS);
  local.changed_visits = shard_sum(cnt + kCtrWords, SH_CHANGED);
  local.label_bytes = packed_run ? 4u : 8u;
  local.local_rounds = shard_sum(cnt + kCtrWords, SH_ROUNDS);
  local.cells_visited = local.tile_visits * Lmax * S.th * S.tw;
  local.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (local.launches && local.kernel_ms == 0) local.kernel_ms = ms;  // graph form: replays only
  if (st) *st = local;
  c.last.add(local);
}

static FieldRead make_read(Ctx& c) {
  FieldRead R;
  Workspace& W = c.W();
  R.boards = W.boards;
  R.fields = W.fields;
  R.labels = W.labels;
  R.state = W.state;
  R.tw = c.shape.tw;
  R.th = c.shape.th;
  R.packed = c.packed ? 1u : 0u;
  return R;
}

// Canonical paths of fields R.fields[0..n) (pool: the cell pool their
// FieldDescs index); path buffers come from the context's workspace.
// Lengths, costs and offsets come from path_scan on the device, the walk
// follows at once, and one synchronisation reads everything back through
// page-locked memory: the counts and offsets, the error flags and the cells,
// as many as the last call's paths held (the rest, when the paths grew, in a
// second copy).  Allocation-free calls carve the cell buffer from the rest
// of the scratch array and read back into the output vectors directly.
static void reconstruct_with(Ctx& c, const FieldRead& R, const uint32_t* pool, size_t n,
                             std::vector<uint8_t>& routed, std::vector<uint32_t>& cost, std::vector<uint64_t>& off,
                             std::vector<uint32_t>& cells, SolveStats* st) {
  HIP_OK(hipSetDevice(c.device));
  Workspace& W = c.W();
  hipStream_t s = c.stream;
  routed.assign(n, 0);
  cost.assign(n, kMax);
  off.assign(n + 1, 0);
  cells.clear();
  if (n == 0) return;
  take(c, W.plen, W.plen_cap, n);
  take(c, W.pcost, W.pcost_cap, n);
  take(c, W.poff, W.poff_cap, n + 1);
  check_arena(c, {{W.plen, n * 4}, {W.pcost, n * 4}, {W.poff, (n + 1) * 8}});
  // the cell buffer: grown to the last call's total (at least 1 Mi cells)
  // outside allocation-free calls, the rest of the scratch array in them
  uint64_t cap;
  if (c.arena) {
    cap = (c.arena_cap - c.arena_used) / 4;
    if (cap == 0) throw Error(1, "the paths need more than the scratch array of the allocation-free entry point");
    W.pcells = (uint32_t*)(c.arena + c.arena_used);
    W.pcells_cap = cap;
  } else {
    grow(W.pcells, W.pcells_cap, std::max<size_t>(c.ws.path_guess, size_t(1) << 20));
    cap = W.pcells_cap;
  }
  // page-locked read-back: plen, cost, err, off, then the cells
  const size_t head = (2 * n + 2) * 4 + (n + 1) * 8;
  const size_t guess = c.arena ? 0 : std::min<size_t>(c.ws.path_guess, cap);
  uint8_t* hp = nullptr;
  if (!c.arena) {
    if (c.ws.hpath_cap < head + guess * 4) {
      if (c.ws.hpath) HIP_OK(hipHostFree(c.ws.hpath));
      c.ws.hpath = nullptr;
      c.ws.hpath_cap = 0;
      const size_t want = (head + guess * 4) * 5 / 4 + 4096;
      HIP_OK(hipHostMalloc((void**)&c.ws.hpath, want, hipHostMallocDefault));
      c.ws.hpath_cap = want;
    }
    hp = c.ws.hpath;
  }
  WalkArgs WA;
  memset(&WA, 0, sizeof(WA));
  WA.R = R;
  WA.pool = pool;
  WA.poff = W.poff;
  WA.plen = W.plen;
  WA.cells = W.pcells;
  WA.counters = W.counters;
  WA.n = (uint32_t)n;
  WA.cap = cap;
  HIP_OK(hipEventRecord(W.ev1, s));
  hipLaunchKernelGGL(path_scan, dim3(1), dim3(1024), 0, s, R, (uint32_t)n, W.plen, W.pcost, W.poff);
  launch_fill32(s, W.counters + C_ERROR * kCtrStride, 0u, 1);
  hipLaunchKernelGGL(walk_paths, dim3((uint32_t)n), dim3(64), 0, s, WA);
  HIP_OK(hipGetLastError());
  uint32_t* h_len = hp ? (uint32_t*)hp : nullptr;
  uint32_t* h_cost = hp ? h_len + n : nullptr;
  uint32_t* h_err = hp ? h_cost + n : nullptr;
  uint64_t* h_off = hp ? (uint64_t*)(hp + (2 * n + 2) * 4) : nullptr;
  uint32_t* h_cells = hp ? (uint32_t*)(hp + head) : nullptr;
  std::vector<uint32_t> plen_v;
  uint32_t err_v = 0;
  if (hp) {
    HIP_OK(hipMemcpyAsync(h_len, W.plen, n * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(h_cost, W.pcost, n * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(h_err, W.counters + C_ERROR * kCtrStride, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(h_off, W.poff, (n + 1) * 8, hipMemcpyDeviceToHost, s));
    if (guess) HIP_OK(hipMemcpyAsync(h_cells, W.pcells, guess * 4, hipMemcpyDeviceToHost, s));
  } else {
    plen_v.resize(n);
    HIP_OK(hipMemcpyAsync(plen_v.data(), W.plen, n * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(cost.data(), W.pcost, n * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&err_v, W.counters + C_ERROR * kCtrStride, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(off.data(), W.poff, (n + 1) * 8, hipMemcpyDeviceToHost, s));
  }
  HIP_OK(hipEventRecord(W.ev2, s));
  HIP_OK(hipStreamSynchronize(s));
  const uint64_t total = hp ? h_off[n] : off[n];
  uint32_t err = hp ? *h_err : err_v;
  bool rewalked = false;
  if (err & 64u) {
    // the paths outgrew the cell buffer (total > cap): walk again into one
    // that holds them
    if (c.arena) throw Error(1, "the paths need more than the scratch array of the allocation-free entry point");
    grow(W.pcells, W.pcells_cap, total);
    WA.cells = W.pcells;
    WA.cap = W.pcells_cap;
    launch_fill32(s, W.counters + C_ERROR * kCtrStride, 0u, 1);
    hipLaunchKernelGGL(walk_paths, dim3((uint32_t)n), dim3(64), 0, s, WA);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(&err_v, W.counters + C_ERROR * kCtrStride, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipEventRecord(W.ev2, s));
    HIP_OK(hipStreamSynchronize(s));
    err = err_v;
    rewalked = true;
  }
  if (err) throw Error(1, "device path reconstruction failed (flags " + std::to_string(err) + ")");
  cells.resize(total);
  if (hp) {
    for (size_t i = 0; i < n; ++i) {
      routed[i] = h_len[i] ? 1 : 0;
      cost[i] = h_cost[i];
    }
    memcpy(off.data(), h_off, (n + 1) * 8);
    // the page-locked copy holds the first `guess` cells of the first walk
    const size_t copied = rewalked ? 0 : std::min<size_t>(guess, total);
    if (copied) memcpy(cells.data(), h_cells, copied * 4);
    if (total > copied)
      HIP_OK(hipMemcpy(cells.data() + copied, W.pcells + copied, (total - copied) * 4, hipMemcpyDeviceToHost));
    c.ws.path_guess = total;
  } else {
    for (size_t i = 0; i < n; ++i) routed[i] = plen_v[i] ? 1 : 0;
    if (total) HIP_OK(hipMemcpy(cells.data(), W.pcells, total * 4, hipMemcpyDeviceToHost));
  }
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, W.ev1, W.ev2));
  if (st) st->path_ms = ms;
  c.last.path_ms += ms;
}

void reconstruct_paths(Ctx& c, size_t n, std::vector<uint8_t>& routed, std::vector<uint32_t>& cost,
                       std::vector<uint64_t>& off, std::vector<uint32_t>& cells, SolveStats* st) {
  reconstruct_with(c, make_read(c), c.W().pool, n, routed, cost, off, cells, st);
}

void read_field(Ctx& c, size_t f, uint32_t* dist, uint32_t* hops) {
  HIP_OK(hipSetDevice(c.device));
  if (f >= c.layout.size()) throw Error(1, "field index out of range");
  const FieldDesc& F = c.layout[f];
  const BoardView& b = c.views[F.board];
  size_t n = (size_t)F.ww * F.wh * b.layers;
  Workspace& W = c.W();
  take(c, W.fout, W.fout_cap, 2 * n);
  check_arena(c, {{W.fout, 2 * n * 4}});
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

// ---------------------------------------------------------------------------
// round driver (serving loop)
// ---------------------------------------------------------------------------

void RangeAlloc::reset(uint64_t cap) {
  cap_ = cap;
  used_ = 0;
  free_.clear();
  if (cap) free_[0] = cap;
}

bool RangeAlloc::alloc(uint64_t n, uint64_t& off) {
  if (n == 0) {
    off = 0;
    return true;
  }
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < n) continue;
    off = it->first;
    const uint64_t rest = it->second - n;
    free_.erase(it);
    if (rest) free_[off + n] = rest;
    used_ += n;
    return true;
  }
  return false;
}

void RangeAlloc::release(uint64_t off, uint64_t n) {
  if (n == 0) return;
  used_ -= n;
  auto it = free_.emplace(off, n).first;
  auto nx = std::next(it);
  if (nx != free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    free_.erase(nx);
  }
  if (it != free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      free_.erase(it);
    }
  }
}

// Flag the groups that own an item of the next round's list.
namespace {
uint64_t driver_items(const RoundDriver& d) { return d.ia.cap(); }

template <typename T>
void dmalloc(T*& p, uint64_t n) {
  HIP_OK(hipMalloc((void**)&p, std::max<uint64_t>(n, 1) * sizeof(T)));
}

template <typename T>
void stage_copy(RoundDriver& d, T* dst, const std::vector<T>& src) {
  if (src.empty()) return;
  d.staging.emplace_back((const uint8_t*)src.data(), (const uint8_t*)src.data() + src.size() * sizeof(T));
  HIP_OK(hipMemcpyAsync(dst, d.staging.back().data(), src.size() * sizeof(T), hipMemcpyHostToDevice, d.c->stream));
}
}  // namespace

void driver_init(RoundDriver& d, Ctx& c, size_t label_bytes, uint32_t max_boards, uint32_t max_layers,
                 uint32_t max_groups) {
  HIP_OK(hipSetDevice(c.device));
  const TileShape& S = c.shape;
  if (!tile_shape_supported(S.tw, S.th)) throw Error(1, "unsupported tile shape");
  if (max_layers < 1 || max_layers > 16) throw Error(1, "max_layers must be in 1..16");
  if (lds_bytes(S.tw, S.th, max_layers, false) > 160 * 1024 - 64) throw Error(1, "tile shape too large for max_layers");
  if (max_boards < 1 || max_groups < 1) throw Error(1, "server needs at least one board and one group slot");
  d.c = &c;
  d.max_layers = max_layers;
  d.max_boards = max_boards;
  d.max_groups = max_groups;
  const uint64_t words = std::max<uint64_t>(label_bytes / 8, 1);
  const uint64_t items = std::min<uint64_t>(words / tile_block_words(1, S.tw, S.th) + 1, 1ull << 26);
  const uint64_t nfields = std::min<uint64_t>(items, 1ull << 22);
  const uint64_t pool_words = 1ull << 24;
  dmalloc(d.labels, words);
  dmalloc(d.state, items);
  dmalloc(d.act, 2 * items);
  dmalloc(d.list, 2 * lshard_words(items));
  dmalloc(d.item_field, items);
  dmalloc(d.item_group, items);
  dmalloc(d.seeds, items);
  dmalloc(d.fields, nfields);
  dmalloc(d.pool, pool_words);
  dmalloc(d.boards, max_boards);
  dmalloc(d.ctr, kCtrWords + max_groups);
  dmalloc(d.shards, kShardWords);
  dmalloc(d.imax, items);
  dmalloc(d.pmax, 16);
  HIP_OK(hipHostMalloc((void**)&d.pinned, (kCtrWords + max_groups + kShardWords) * sizeof(uint32_t),
                       hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&d.hostctr, (kDriverFlagsAt + max_groups) * sizeof(uint32_t),
                       hipHostMallocMapped | hipHostMallocCoherent));
  memset(d.hostctr, 0, (kDriverFlagsAt + max_groups) * sizeof(uint32_t));
  HIP_OK(hipHostGetDevicePointer((void**)&d.hostctr_dev, d.hostctr, 0));
  launch_fill32(c.stream, d.act, 0u, 2 * items);
  launch_fill32(c.stream, d.ctr, 0u, kCtrWords + max_groups);
  launch_fill32(c.stream, d.shards, 0u, kShardWords);
  HIP_OK(hipStreamSynchronize(c.stream));
  d.fa.reset(nfields);
  d.ia.reset(items);
  d.la.reset(words);
  d.pa.reset(pool_words);
  d.views.assign(max_boards, BoardView{});
  d.devs.assign(max_boards, nullptr);
  d.board_used.assign(max_boards, 0);
  d.groups.assign(max_groups, Group{});
  d.pending_seeds.clear();
  d.n_cur = 0;
  d.cur = 0;
}

void driver_free(RoundDriver& d) {
  if (!d.c) return;
  (void)hipSetDevice(d.c->device);
  void* ptrs[] = {d.labels, d.state,  d.act,   d.list,   d.item_field,  d.item_group,
                  d.seeds,  d.fields, d.pool,  d.boards, d.ctr, d.shards, d.imax, d.pmax};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (d.pinned) (void)hipHostFree(d.pinned);
  if (d.hostctr) (void)hipHostFree(d.hostctr);
  for (hipEvent_t e : d.ev_free) (void)hipEventDestroy(e);
  for (auto& pr : d.ev_pending) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  d = RoundDriver();
}

size_t driver_device_bytes(const RoundDriver& d) {
  const uint64_t items = d.ia.cap();
  return d.la.cap() * 8 + items * 4 * 8 + d.fa.cap() * sizeof(FieldDesc) + d.pa.cap() * 4 +
         (size_t)d.max_boards * sizeof(BoardView) + (size_t)d.max_groups * 4;
}

int driver_add_board(RoundDriver& d, const DevBoard& b) {
  if (b.layers > d.max_layers) throw Error(1, "board has more layers than the server was configured for");
  for (uint32_t k = 0; k < d.max_boards; ++k) {
    if (d.board_used[k]) continue;
    d.views[k] = b.view();
    d.devs[k] = &b;
    d.board_used[k] = 1;
    HIP_OK(hipMemcpyAsync(d.boards + k, &d.views[k], sizeof(BoardView), hipMemcpyHostToDevice, d.c->stream));
    HIP_OK(hipStreamSynchronize(d.c->stream));
    return (int)k;
  }
  return -1;
}

void driver_remove_board(RoundDriver& d, int slot) {
  if (slot >= 0 && (uint32_t)slot < d.max_boards) {
    d.board_used[slot] = 0;
    d.devs[slot] = nullptr;
  }
}

uint64_t driver_group_words(const RoundDriver& d, const std::vector<FieldReq>& reqs) {
  uint64_t w = 0;
  for (const FieldReq& q : reqs) {
    if (q.board >= d.max_boards || !d.board_used[q.board]) throw Error(1, "field names an unknown board");
    w += field_label_words(d.c->shape, q.x1 - q.x0 + 1, q.y1 - q.y0 + 1, d.views[q.board].layers);
  }
  return w;
}

bool driver_admit(RoundDriver& d, const std::vector<FieldReq>& reqs, int& gid, bool allow_pack) {
  const TileShape& S = d.c->shape;
  int g = -1;
  for (uint32_t k = 0; k < d.max_groups; ++k)
    if (!d.groups[k].live) {
      g = (int)k;
      break;
    }
  if (g < 0) return false;
  uint64_t nitems = 0, words = 0, npool = 0;
  Group G;
  G.narrow = true;
  G.lean = true;
  G.wave = true;
  G.halo = false;
  G.steps = true;
  bool steps_ok = true;  // every open step of its boards costs 1 .. 0xFFFE
  for (const FieldReq& q : reqs) {
    if (q.board >= d.max_boards || !d.board_used[q.board]) throw Error(1, "field names an unknown board");
    if (!q.ovr_cells.empty()) throw Error(1, "the round driver takes no additive overrides");
    const uint32_t ww = q.x1 - q.x0 + 1, wh = q.y1 - q.y0 + 1;
    nitems += (uint64_t)((ww + S.tw - 1) / S.tw) * ((wh + S.th - 1) / S.th);
    words += field_label_words(S, ww, wh, d.views[q.board].layers);
    npool += q.sources.size() + q.forced.size() + q.pads.size() + q.via_pads.size();
    if (!q.pads.empty() || !q.via_pads.empty())  // the pad tables of build_pad_tables
      npool += 2ull * (((ww + S.tw - 1) / S.tw) * ((wh + S.th - 1) / S.th) + 1) + q.pads.size() + q.via_pads.size();
    const DevBoard& b = *d.devs[q.board];
    G.layers = std::max(G.layers, d.views[q.board].layers);
    G.narrow = G.narrow && field_narrow(b, q);
    G.lean = G.lean && field_lean(d.views[q.board], q);
    G.wave = G.wave && field_wave(d.views[q.board], q);
    G.halo = G.halo || d.views[q.board].halo != nullptr;
    G.steps = G.steps && use_steps(d.views[q.board]);
    G.ex = G.ex || field_extras(q);
    steps_ok = steps_ok && b.min_step >= 1 && b.max_step < 0xFFFFu && !b.pack_failed;
    G.pk_step = std::max<uint64_t>(G.pk_step, b.max_step);
  }
  // packed labels only where the wave form runs them: the wave shape, a
  // narrow wave-form group of at most eight layers, no halo-owner plane
  G.pack = allow_pack && steps_ok && wave_shape(S) && G.narrow && G.wave && !G.halo && G.layers <= 8;
  for (uint32_t k = 0; k < d.max_groups; ++k)
    if (d.groups[k].live && d.groups[k].pack != G.pack) return false;  // the other label width is running
  // packed: u32 labels, two to a label word of the allocator
  const uint64_t alloc_words = G.pack ? (words + 1) / 2 : words;
  if (!d.fa.alloc(reqs.size(), G.field0)) return false;
  if (!d.ia.alloc(nitems, G.item0)) {
    d.fa.release(G.field0, reqs.size());
    return false;
  }
  if (!d.la.alloc(alloc_words, G.lab0)) {
    d.fa.release(G.field0, reqs.size());
    d.ia.release(G.item0, nitems);
    return false;
  }
  if (!d.pa.alloc(npool, G.pool0)) {
    d.fa.release(G.field0, reqs.size());
    d.ia.release(G.item0, nitems);
    d.la.release(G.lab0, alloc_words);
    return false;
  }
  G.nfields = reqs.size();
  G.nitems = nitems;
  G.labwords = alloc_words;
  G.npool = npool;
  std::vector<FieldDesc> descs(reqs.size());
  std::vector<uint32_t> pool, item_field;
  pool.reserve(npool);
  item_field.reserve(nitems);
  // label offsets in labels: u64 words, or u32 halves of the group's words
  uint64_t item = G.item0, lab = G.pack ? 2 * G.lab0 : G.lab0;
  std::vector<uint32_t> seeds;
  for (size_t i = 0; i < reqs.size(); ++i) {
    const FieldReq& q = reqs[i];
    const BoardView& b = d.views[q.board];
    FieldDesc& F = descs[i];
    memset(&F, 0, sizeof(F));
    F.x0 = q.x0;
    F.y0 = q.y0;
    F.ww = q.x1 - q.x0 + 1;
    F.wh = q.y1 - q.y0 + 1;
    F.tx = (F.ww + S.tw - 1) / S.tw;
    F.ty = (F.wh + S.th - 1) / S.th;
    F.item0 = (uint32_t)item;
    F.lab_off = lab;
    auto put = [&](const std::vector<uint32_t>& v, uint32_t& off, uint32_t& n) {
      off = (uint32_t)(G.pool0 + pool.size());
      n = (uint32_t)v.size();
      pool.insert(pool.end(), v.begin(), v.end());
    };
    put(q.sources, F.src_off, F.n_src);
    put(q.forced, F.forced_off, F.n_forced);
    put(q.pads, F.pad_off, F.n_pad);
    put(q.via_pads, F.vpad_off, F.n_vpad);
    std::vector<uint32_t> xtab;
    build_pad_tables(q, *d.devs[q.board], S, F.tx, F.ty, xtab);
    uint32_t n_xtab = 0;
    put(xtab, F.xtab_off, n_xtab);
    F.dst = q.dst;
    F.own_id = q.own_id;
    F.own_group = q.own_group;
    F.flags = q.flags;
    F.board = q.board;
    const uint32_t nt = F.tx * F.ty;
    for (uint32_t k = 0; k < nt; ++k) item_field.push_back((uint32_t)(G.field0 + i));
    std::vector<uint32_t> st;
    for (uint32_t src : q.sources) {
      const uint32_t r = src % b.plane, gy = r / b.w, gx = r % b.w;
      if ((size_t)src >= (size_t)b.plane * b.layers || gx < F.x0 || gy < F.y0 || gx > q.x1 || gy > q.y1) continue;
      st.push_back((uint32_t)item + ((gy - F.y0) / S.th) * F.tx + (gx - F.x0) / S.tw);
    }
    std::sort(st.begin(), st.end());
    st.erase(std::unique(st.begin(), st.end()), st.end());
    seeds.insert(seeds.end(), st.begin(), st.end());
    item += nt;
    lab += (uint64_t)nt * tile_block_words(b.layers, S.tw, S.th);
  }
  hipStream_t s = d.c->stream;
  stage_copy(d, d.fields + G.field0, descs);
  stage_copy(d, d.pool + G.pool0, pool);
  stage_copy(d, d.item_field + G.item0, item_field);
  launch_fill32(s, d.item_group + G.item0, (uint32_t)g, nitems);
  launch_fill32(s, d.state + G.item0, 0u, nitems);
  if (G.pack) launch_fill32(s, d.imax + G.item0, 0u, nitems);
  G.live = true;
  G.active = !seeds.empty();
  d.pending_seeds.insert(d.pending_seeds.end(), seeds.begin(), seeds.end());
  d.groups[g] = G;
  gid = g;
  return true;
}

uint32_t driver_step(RoundDriver& d) {
  Ctx& c = *d.c;
  HIP_OK(hipSetDevice(c.device));
  hipStream_t s = c.stream;
  const TileShape& S = c.shape;
  const uint64_t items = driver_items(d);
  uint32_t* act_cur = d.act + (uint64_t)d.cur * items;
  uint32_t* act_next = d.act + (uint64_t)(1 - d.cur) * items;
  const uint32_t lcap = lshard_cap(items);
  uint32_t* list_cur = d.list + (uint64_t)d.cur * lshard_words(items);
  uint32_t* list_next = d.list + (uint64_t)(1 - d.cur) * lshard_words(items);
  if (!d.pending_seeds.empty()) {
    const uint32_t k = (uint32_t)d.pending_seeds.size();
    stage_copy(d, d.seeds, d.pending_seeds);
    // appended to the round's sharded list (count set (rounds + 2) % 3); the
    // grid covers the longest sub-list it can reach
    hipLaunchKernelGGL(seed_list_sharded, dim3((k + 255) / 256), dim3(256), 0, s, act_cur, list_cur, lcap, d.shards,
                       (uint32_t)((d.rounds + 2) % 3), d.seeds, k, (uint32_t*)nullptr);
    hipLaunchKernelGGL(list_table, dim3(1), dim3(kShards), 0, s, d.shards, (uint32_t)((d.rounds + 2) % 3));
    d.n_cur += k;
    d.pending_seeds.clear();
  }
  if (d.n_cur == 0) {
    HIP_OK(hipStreamSynchronize(s));
    d.staging.clear();
    return 0;
  }
  // With one live group its fields are active exactly when the next list is
  // not empty: no group flags are set or read.  Otherwise a visit stamps its
  // group's flag with the round (flags need no clearing) and the launch's
  // last workgroup copies the flags to host-mapped memory.  As in the plain
  // form, the next-list count alternates between two counters (the kernel
  // clears the other one), the launch publishes count, error flags and stamp,
  // and the host spins on the stamp: no fill, copy or stream synchronisation
  // per round.  Visit and in-LDS round counters accumulate on the device and
  // are read by driver_sync_stats.
  uint32_t live = 0;
  for (uint32_t g = 0; g < d.max_groups; ++g) live += d.groups[g].live ? 1u : 0u;
  RelaxArgs A;
  memset(&A, 0, sizeof(A));
  A.boards = d.boards;
  A.fields = d.fields;
  A.item_field = d.item_field;
  A.pool = d.pool;
  A.labels = d.labels;
  A.state = d.state;
  A.act_cur = act_cur;
  A.act_next = act_next;
  A.list_cur = list_cur;
  A.list_next = list_next;
  A.counters = d.ctr;
  A.group_active = live > 1 ? d.ctr + kCtrWords : nullptr;
  A.shards = d.shards;
  A.item_group = d.item_group;
  A.host_flags = live > 1 ? d.hostctr_dev + kDriverFlagsAt : nullptr;
  A.n_groups = d.max_groups;
  A.max_local = S.max_local;
  A.next_ctr = C_NEXT0 + (uint32_t)(d.rounds & 1);
  A.clear_ctr = C_NEXT0 + (uint32_t)((d.rounds + 1) & 1);
  A.host_out = d.hostctr_dev;
  A.key_cur = A.key_next = nullptr;
  A.key_ctr = A.key_clear = C_KEY0;
  A.delta = kMax;
  // The tile kernel and its LDS are sized for the most layers of any live
  // group's board, not for the server's maximum.
  uint32_t Lmax = 1;
  bool narrow = true, lean = true, wave_ok = true, halo = false, pack = false, ex = false, sp = true;
  for (uint32_t g = 0; g < d.max_groups; ++g)
    if (d.groups[g].live) {
      sp = sp && d.groups[g].steps;
      Lmax = std::max(Lmax, d.groups[g].layers);
      narrow = narrow && d.groups[g].narrow;
      lean = lean && d.groups[g].lean;
      wave_ok = wave_ok && d.groups[g].wave;
      halo = halo || d.groups[g].halo;
      ex = ex || d.groups[g].ex;
      pack = d.groups[g].pack;  // the same for every live group
    }
  const KernelSet& K = pick_kernels(S.tw, S.th, Lmax, narrow, lean);
  const RoundLaunch RL = round_launch(K, S, Lmax, wave_ok, false, pack, halo, ex, sp);
  if (RL.packed != pack) throw Error(1, "packed groups without the packed kernel");
  A.item_max = d.imax;
  void* args[] = {&A};
  if (d.stamp_round >= (1u << 30)) {
    // states and flags stay below 2^31: every touched item now reads as
    // first visited in round 0, before every later round, and no group flag
    // holds a later round than the next one
    hipLaunchKernelGGL(restamp_states, dim3(1024), dim3(256), 0, s, d.state, (size_t)items);
    launch_fill32(s, d.ctr + kCtrWords, 0u, d.max_groups);
    d.stamp_round = 1;
  }
  A.round = d.stamp_round++;
  if (++d.stamp == 0) d.stamp = 1;
  A.stamp = d.stamp;
  // kernel time from an event pair per round, read once the pair has passed,
  // when the context times rounds (a server's context does not)
  const bool timed = c.timing_every != 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (timed) {
    for (hipEvent_t* e : {&e0, &e1}) {
      if (!d.ev_free.empty()) {
        *e = d.ev_free.back();
        d.ev_free.pop_back();
      } else {
        HIP_OK(hipEventCreate(e));
      }
    }
    HIP_OK(hipEventRecord(e0, s));
  }
  A.lcap = lcap;
  A.lset_next = (uint32_t)(d.rounds % 3);
  A.lset_cur = (uint32_t)((d.rounds + 2) % 3);
  A.lset_clear = (uint32_t)((d.rounds + 1) % 3);
  HIP_OK(hipLaunchKernel(RL.fn, dim3(d.n_cur), dim3(RL.threads), args, RL.lds, s));
  HIP_OK(hipGetLastError());
  if (timed) {
    HIP_OK(hipEventRecord(e1, s));
    d.ev_pending.push_back({e0, e1});
  }
  volatile uint32_t* hc = d.hostctr;
  const auto spin0 = std::chrono::steady_clock::now();
  while (hc[2] != A.stamp) {
    if (std::chrono::steady_clock::now() - spin0 >= std::chrono::microseconds(c.stamp_spin_us)) {
      HIP_OK(hipStreamSynchronize(s));
      if (hc[2] != A.stamp) throw Error(1, "relaxation round did not publish its counts");
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  d.staging.clear();  // the launch's stamp follows the staged copies in the stream
  if (hc[1]) throw Error(1, "device relaxation reported error flags " + std::to_string(hc[1]));
  while (!d.ev_pending.empty() && hipEventQuery(d.ev_pending.front().second) == hipSuccess) {
    float kms = 0;
    HIP_OK(hipEventElapsedTime(&kms, d.ev_pending.front().first, d.ev_pending.front().second));
    d.kernel_ms += kms;
    d.ev_free.push_back(d.ev_pending.front().first);
    d.ev_free.push_back(d.ev_pending.front().second);
    d.ev_pending.pop_front();
  }
  d.rounds += 1;
  d.n_cur = hc[0];
  d.cur = 1 - d.cur;
  const volatile uint32_t* hf = d.hostctr + kDriverFlagsAt;
  for (uint32_t g = 0; g < d.max_groups; ++g)
    if (d.groups[g].live) d.groups[g].active = live > 1 ? hf[g] == A.round + 1u : d.n_cur != 0;
  return d.n_cur;
}

void driver_sync_stats(RoundDriver& d) {
  Ctx& c = *d.c;
  HIP_OK(hipSetDevice(c.device));
  HIP_OK(hipMemcpyAsync(d.pinned, d.shards, kShardWords * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  const uint32_t v = (uint32_t)shard_sum(d.pinned, SH_VISITS), r = (uint32_t)shard_sum(d.pinned, SH_ROUNDS);
  d.tile_visits += (uint32_t)(v - d.seen_visits);  // the device counters wrap at 2^32
  d.local_rounds += (uint32_t)(r - d.seen_lrounds);
  d.seen_visits = v;
  d.seen_lrounds = r;
  while (!d.ev_pending.empty()) {
    float kms = 0;
    HIP_OK(hipEventElapsedTime(&kms, d.ev_pending.front().first, d.ev_pending.front().second));
    d.kernel_ms += kms;
    d.ev_free.push_back(d.ev_pending.front().first);
    d.ev_free.push_back(d.ev_pending.front().second);
    d.ev_pending.pop_front();
  }
}

void driver_finished(RoundDriver& d, std::vector<int>& out) {
  out.clear();
  for (uint32_t g = 0; g < d.max_groups; ++g)
    if (d.groups[g].live && !d.groups[g].active) out.push_back((int)g);
}

bool driver_paths(RoundDriver& d, int g, std::vector<uint8_t>& routed, std::vector<uint32_t>& cost,
                  std::vector<uint64_t>& off, std::vector<uint32_t>& cells) {
  const Group& G = d.groups.at(g);
  hipStream_t s = d.c->stream;
  if (G.pack) {
    // the group's certificate: its largest stored label, read back with the
    // paths (reconstruct_with synchronises the stream)
    launch_fill32(s, d.pmax, 0u, 1);
    hipLaunchKernelGGL(reduce_item_max, dim3((uint32_t)std::min<uint64_t>((G.nitems + 255) / 256, 256)), dim3(256),
                       0, s, d.imax + G.item0, (size_t)G.nitems, d.pmax);
    HIP_OK(hipMemcpyAsync(d.pinned, d.pmax, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  }
  FieldRead R;
  R.boards = d.boards;
  R.fields = d.fields + G.field0;
  R.labels = d.labels;
  R.state = d.state;
  R.packed = G.pack ? 1 : 0;
  R.tw = d.c->shape.tw;
  R.th = d.c->shape.th;
  reconstruct_with(*d.c, R, d.pool, G.nfields, routed, cost, off, cells, nullptr);
  if (G.pack && (uint64_t)d.pinned[0] + ((G.pk_step << 16) | 1u) >= 0xFFFF0000ull) {
    routed.clear();
    cost.clear();
    off.clear();
    cells.clear();
    return false;
  }
  return true;
}

void driver_release(RoundDriver& d, int g) {
  Group& G = d.groups.at(g);
  if (!G.live) return;
  d.fa.release(G.field0, G.nfields);
  d.ia.release(G.item0, G.nitems);
  d.la.release(G.lab0, G.labwords);
  d.pa.release(G.pool0, G.npool);
  G = Group{};
}

}  // namespace mrf
