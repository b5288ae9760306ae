// Internal engine interface shared by the kernels (engine.hip), the C ABI
// (capi.cpp) and the serving loop (serve.cpp).  Host-only C++; no torch types.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace mrf {

constexpr uint32_t kMax = 0xFFFFFFFFu;  // mr-core OBSTACLE / unreachable sentinel

// Board mask bits (one byte per cell, built once per resident board).
enum : uint8_t {
  M_TRACE = 0x01,  // BOARD_TRACE_NODE: never passable (mr-core lib.rs:445-452)
  M_OBST = 0x02,   // cost == OBSTACLE
  M_NY = 0x04,     // planar edge to y-1 open (is_board_planar_step_forbidden false)
  M_NX = 0x08,     // planar edge to x-1 open
  M_PX = 0x10,     // planar edge to x+1 open
  M_PY = 0x20,     // planar edge to y+1 open
  M_DN = 0x40,     // via to layer-1 legal without per-net exemptions
  M_UP = 0x80,     // via to layer+1 legal without per-net exemptions
};
// Raw via bits kept for the per-net via-exemption path.
enum : uint8_t { R_VIANODE = 0x01, R_VF = 0x02 };

// Per-field flags.
enum : uint32_t {
  F_FORCE = 0x1,  // sources and dst stay passable unless trace-node (negotiated.rs:2946-2955)
};

struct BoardView {
  uint32_t w, h, layers, plane;
  const uint32_t* cost;     // enter weight; kMax = obstacle; nullptr when every
                            // passable cell weighs 1 (obstacles: M_OBST)
  const uint32_t* add;      // additive price plane (u32) or nullptr
  const uint16_t* add16;    // additive price plane (u16) or nullptr; at most one of add / add16
  const int32_t* halo;      // halo owner plane (-1 free) or nullptr
  uint32_t halo_pen;
  const uint8_t* bmask;     // M_* bits
  const uint8_t* raw;       // R_* bits
  const uint32_t* xe;       // planar x gap prices [w-1] (+1 pad)
  const uint32_t* ye;       // planar y gap prices [h-1] (+1 pad)
  const uint32_t* ve;       // via prices [layers-1] (+1 pad)
  uint32_t va;              // bit k: vias allowed between layers k and k+1
  uint32_t pad_;
  // Legal searches: committed copper owner and clearance-halo planes (-1
  // free, halo -2 mixed, else a group), or nullptr.
  const int32_t* owner;
  const int32_t* lhalo;
  // Step prices of the lean wave visit, precomputed per cell: three planes of
  // n u32 (x, y, via), each a pair of narrow prices (lo: entering from the
  // lower neighbour, hi: from the upper; 0xFFFF closed).  nullptr when the
  // board has a weight or halo-owner plane.
  const uint32_t* steps;
};

// One complete field: a window (all layers), its sources, overrides.
struct FieldDesc {
  uint64_t lab_off;    // offset of the field's first tile block in the label buffer (u64 units)
  uint32_t x0, y0;     // window origin (board coords)
  uint32_t ww, wh;     // window size
  uint32_t tx, ty;     // tiles along x and y
  uint32_t item0;      // first work item (tile) of the field
  uint32_t src_off, n_src;    // sources in the cell pool
  uint32_t forced_off, n_forced;  // forced-passable cells in the cell pool
  uint32_t pad_off, n_pad;    // sorted own pads
  uint32_t vpad_off, n_vpad;  // sorted via-exempt pads
  uint32_t dst;        // target cell or kMax
  int32_t own_id;      // halo owner id of this net
  uint32_t flags;
  uint32_t board;      // index into the launch's board table
  // Additive overrides of the net (negotiated batches), n_ovr entries in the
  // pool at ovr_off: cells ascending [n], their additive prices [n], per-tile
  // start indices [tx*ty+1], then the same entries ordered by tile: cells [n],
  // prices [n].
  uint32_t ovr_off, n_ovr;
  int32_t own_group;  // legal searches: the net's group
  // Own pads and via-exempt pads by tile (the wave form), in the pool at
  // xtab_off when the field has either: own-pad starts [tx*ty+1], via-pad
  // starts [tx*ty+1], then the entries of each kind ordered by tile, as the
  // cell's index in its tile block ((layer * TH + y) * TW + x).
  uint32_t xtab_off;
};

struct TileShape {
  uint32_t tw = 16, th = 16;  // tile width / height in cells (all layers per tile)
  // Workgroup size.  64 with 16 x 16 tiles: one wavefront per tile visit
  // (the wave form) where the launch qualifies, 256 for the other launches.
  uint32_t threads = 64;
  uint32_t max_local = 32;    // in-LDS relaxation rounds per tile visit
};

struct DevBoard {
  uint32_t w = 0, h = 0, layers = 0;
  uint32_t* cost = nullptr;     // nullptr: every passable cell weighs 1
  uint32_t* add = nullptr;      // additive prices as u32 ...
  uint16_t* add16 = nullptr;    // ... or as u16 when all fit 16 bits
  int32_t* halo = nullptr;
  uint32_t halo_pen = 0;
  uint8_t* bmask = nullptr;
  uint8_t* raw = nullptr;
  uint32_t* xe = nullptr;
  uint32_t* ye = nullptr;
  uint32_t* ve = nullptr;
  std::vector<uint8_t> via_allowed;  // host copy
  // A packed-label solve on this board failed its certificate: later solves
  // on it go straight to wide labels.
  mutable bool pack_failed = false;
  // Host copy of the obstacle bits (cost == kMax), one bit per cell, built on
  // the device with the masks.
  std::vector<uint64_t> obst_bits;
  // Step price pairs for the lean wave visit (3 x n u32, in the block), or
  // nullptr (a weight or halo-owner plane); rebuilt when the additive plane
  // changes.
  uint32_t* steps = nullptr;
  bool obstacle(size_t g) const { return (g >> 6) < obst_bits.size() && ((obst_bits[g >> 6] >> (g & 63)) & 1u); }
  // The static planes (bmask, raw, obstacle bits, edge prices, weights) share
  // one allocation: the board's own, or, for a board that lives for one call,
  // the context's borrowed block (lease points at its busy flag).
  uint8_t* block = nullptr;
  bool* lease = nullptr;
  uint64_t max_step = 0;  // bound on any step price (obstacles priced at weight 1, as own pads)
  uint64_t max_ew = 0;    // largest edge price x largest enter weight (max_step without the additive part)
  // Lower bound on every open step price: the smallest edge price (planar
  // gaps, allowed vias) x the smallest enter weight of a passable cell plus
  // the smallest additive price (min_step without the additive part: min_ew).
  uint64_t min_step = 0, min_ew = 0;
  int32_t* owner = nullptr;  // legal-search planes (allocated on first use)
  int32_t* lhalo = nullptr;
  bool legal = false;        // the owner / halo planes apply to the current call
  bool legal_owner = false, legal_halo = false;  // which of them the call passed
  size_t bytes = 0;
  BoardView view() const;
};

// Host-side description of one field to solve.
struct FieldReq {
  uint32_t x0, y0, x1, y1;  // inclusive window
  std::vector<uint32_t> sources;
  std::vector<uint32_t> forced;
  std::vector<uint32_t> pads;      // sorted unique
  std::vector<uint32_t> via_pads;  // sorted unique
  uint32_t dst = kMax;
  int32_t own_id = -1;
  uint32_t flags = 0;
  uint32_t board = 0;  // index into the boards passed to solve_fields
  // Per-net additive prices replacing the board plane's on these cells
  // (ascending, unique); cells outside the window are ignored.
  std::vector<uint32_t> ovr_cells, ovr_add;
  int32_t own_group = -1;  // legal searches: the net's group
};

struct SolveStats {
  double relax_ms = 0, path_ms = 0, total_ms = 0;
  uint64_t tile_visits = 0;    // device counter of tile loads (HBM sweep visits)
  uint64_t changed_visits = 0; // visits that improved at least one label
  uint64_t local_rounds = 0;   // in-LDS relaxation rounds
  uint64_t cells_visited = 0;  // tile_visits x tile cells
  uint32_t rounds = 0;         // host rounds (plain form) or 1 (persistent)
  uint64_t field_cells = 0;    // cells in all fields (window x layers)
  double persistent_ms = 0;    // persistent kernel alone (persistent form)
  uint32_t replays = 0;        // hipGraph replays (graph form)
  double kernel_ms = 0;        // relaxation kernels alone: sum of per-launch event times (plain form)
  double validate_ms = 0;      // host re-validation of device paths
  uint32_t launches = 0;       // relaxation kernel launches
  uint32_t wave_launches = 0;  // of which ran the one-wavefront-per-tile kernel
  uint32_t timed_launches = 0; // of which kernel_ms holds the per-launch event times
  uint32_t label_bytes = 8;    // label width of the solve (4: packed u32 labels)
  void add(const SolveStats& o) {
    relax_ms += o.relax_ms;
    path_ms += o.path_ms;
    total_ms += o.total_ms;
    tile_visits += o.tile_visits;
    changed_visits += o.changed_visits;
    local_rounds += o.local_rounds;
    cells_visited += o.cells_visited;
    rounds += o.rounds;
    field_cells += o.field_cells;
    persistent_ms += o.persistent_ms;
    replays += o.replays;
    kernel_ms += o.kernel_ms;
    validate_ms += o.validate_ms;
    launches += o.launches;
    wave_launches += o.wave_launches;
    timed_launches += o.timed_launches;
    label_bytes = o.label_bytes;
  }
};

// MODE_PLAIN: one launch per round, host reads the activation count each round.
// MODE_PERSISTENT: resident-grid kernel with a device grid barrier between
//   rounds.  Disabled: it did not return on MI355X (a device grid barrier that never completed); requesting
//   it raises an error.
// MODE_GRAPH: grid-stride kernel over a device work list per round, rounds
//   separated by kernel boundaries, captured in a hipGraph and replayed; the
//   host reads the activation count once per replay.
// MODE_QUEUE: one persistent launch; workgroups pop tiles from device work
//   queues (one per XCD) and push the neighbours their visits improve, with
//   no rounds and no barrier between workgroups.
enum SolveMode : int { MODE_PLAIN = 0, MODE_PERSISTENT = 1, MODE_GRAPH = 2, MODE_QUEUE = 3 };

class Error : public std::exception {
 public:
  Error(int code, std::string msg, uint32_t net = 0) : code_(code), net_(net), msg_(std::move(msg)) {}
  const char* what() const noexcept override { return msg_.c_str(); }
  int code() const { return code_; }
  uint32_t net() const { return net_; }

 private:
  int code_;
  uint32_t net_;
  std::string msg_;
};

// Device workspace for one batch of fields (grow-only buffers).
struct Workspace {
  uint64_t* labels = nullptr;
  size_t labels_cap = 0;
  uint32_t* state = nullptr;  // per-item state bits
  size_t state_cap = 0;
  uint32_t* act = nullptr;    // 2 x items activation flags
  size_t act_cap = 0;
  uint32_t* key = nullptr;    // 2 x items activation keys (delta stepping)
  size_t key_cap = 0;
  uint32_t* list = nullptr;   // 2 x items active-item lists (round form)
  size_t list_cap = 0;
  FieldDesc* fields = nullptr;
  size_t fields_cap = 0;
  uint32_t* item_field = nullptr;
  size_t item_field_cap = 0;
  uint32_t* pool = nullptr;  // cell pool (sources, forced, pads, seed items)
  size_t pool_cap = 0;
  uint32_t* counters = nullptr;  // device counters
  uint32_t* plen = nullptr;      // per-field path length
  size_t plen_cap = 0;
  uint64_t* poff = nullptr;      // per-field path offsets
  size_t poff_cap = 0;
  uint32_t* pcost = nullptr;     // per-field search cost
  size_t pcost_cap = 0;
  uint32_t* pcells = nullptr;    // path cells
  size_t pcells_cap = 0;
  // page-locked read-back of path counts, offsets and cells, and the last
  // call's total cells (how many to read back with the counts)
  uint8_t* hpath = nullptr;
  size_t hpath_cap = 0;
  size_t path_guess = 0;
  uint64_t* qslots = nullptr;    // queue form: every queue's ring
  size_t qslots_cap = 0;
  uint32_t* qctl = nullptr;      // queue form: head, tail, pending of each queue
  size_t qctl_cap = 0;
  uint32_t* fout = nullptr;      // extracted field output staging
  size_t fout_cap = 0;
  uint32_t* fmax = nullptr;      // packed labels: per-field running maximum of the stored labels
  size_t fmax_cap = 0;
  uint32_t* pinned = nullptr;    // host-pinned counter readback
  uint32_t* hostctr = nullptr;   // host-mapped words the plain form's kernel publishes to
  uint32_t* hostctr_dev = nullptr;
  BoardView* boards = nullptr;   // device board table of the current batch
  size_t boards_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ek0 = nullptr, ek1 = nullptr;
  // Plain form: one (start, end) event pair per round of a solve, read after
  // the last round; and the stamp of the last launch whose counts the kernel
  // published to hostctr (never 0).
  std::vector<hipEvent_t> round_ev;
  uint32_t stamp = 0;
};

struct GraphCache {
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> key;  // launch arguments the graph was captured with
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex lane;  // single GPU lane (mr-cli lib.rs:499 METAL_ISOLATED_MUTEX)
  Workspace ws;
  GraphCache graph;
  TileShape shape;
  int mode = MODE_PLAIN;
  uint32_t graph_rounds = 8;  // rounds captured per hipGraph replay
  uint32_t delta = kMax;      // plain form: delta-stepping window in search cost (kMax = off)
  size_t label_budget = size_t(48) << 30;  // bytes of labels per chunk (set: mrf_set_label_budget)
  size_t label_budget_default = size_t(48) << 30;  // min(48 GiB, half of device memory)
  int cus = 256;
  SolveStats last;
  std::vector<FieldDesc> layout;  // fields of the last solve (for paths / reads)
  bool packed = false;             // the last solve's labels are packed u32 (wave form)
  bool no_pack = false;            // solve in wide labels (set for the re-solve of a packed batch
                                   // whose certificate failed)
  std::vector<BoardView> views;   // boards of the last solve
  // Fault injection for fail-closed tests (0 none, 1 backend error, 2 corrupt path).
  int inject = 0;
  // Plain form: how long the host spins on a launch's stamp before it hands
  // over to a stream synchronisation (microseconds).
  uint32_t stamp_spin_us = 20000;
  // Plain-form rounds bracketed by HIP events: in every `timing_every`-th
  // solving call (1: every call, 0: none, the default: the event pairs cost
  // host turnaround every round); `time_rounds` is this call's decision,
  // `timing_calls` counts the calls.
  uint32_t timing_every = 0;
  uint64_t timing_calls = 0;
  bool time_rounds = false;
  // Queue form: the longest a workgroup waits for work or a queue slot, and
  // the longest one runs, before the launch stops with an error (ms).
  uint32_t queue_wait_ms = 2000, queue_run_ms = 20000;
  // Allocation-free calls (arena set, between arena_begin and arena_end):
  // every per-solve device buffer is carved from `arena`, the library's
  // scratch array on this device, instead of grown with hipMalloc; the
  // buffers made when the context opened (device counters, host-mapped
  // words, events) and the launch stamp are the context's own.
  uint8_t* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  Workspace ws_arena;
  Workspace& W() { return arena ? ws_arena : ws; }
  // Board uploads: the caller's planes are staged in `stage` (grown on
  // demand) and the masks built from them on the device; a board that lives
  // for one call borrows `tblock` for its static planes while it is free.
  uint8_t* stage = nullptr;
  size_t stage_cap = 0;
  uint8_t* tblock = nullptr;
  size_t tblock_cap = 0;
  bool tblock_busy = false;
  uint32_t* upstat = nullptr;  // device words of the upload's weight / price bounds
};

// ---- engine.hip -------------------------------------------------------------
// The library's scratch array (`scratch` in engine.hip): its device address on
// `device` (resolved when the first context opens on it; nullptr before) and
// its size in bytes.
uint8_t* scratch_ptr(int device);
size_t scratch_bytes();
// Allocation-free mode of a context for the duration of one call; arena_begin
// throws when the device has no scratch address.
void arena_begin(Ctx& c);
void arena_end(Ctx& c);
void board_upload(Ctx& c, DevBoard& b, uint32_t w, uint32_t h, uint32_t layers,
                  const uint32_t* cost, const uint8_t* via_forbidden, const uint8_t* board,
                  const uint32_t* xe, const uint32_t* ye, const uint32_t* ve,
                  const uint8_t* via_allowed, const uint32_t* add, const int32_t* halo,
                  uint32_t halo_pen, bool transient = false);
// transient: the board is freed before the call that uploaded it returns, so
// its static planes may borrow the context's block.
void board_free(DevBoard& b);
// Replace (or add) the board's additive price plane; host array of w*h*layers.
// max_add: the plane's largest value when the caller has it (UINT64_MAX:
// computed here).
void board_set_additive(Ctx& c, DevBoard& b, const uint32_t* add, uint64_t max_add = UINT64_MAX);
// Upload the legal-search owner / halo planes (either may be null: none) and
// make them apply to the board's solves until board_clear_legal.
void board_set_legal(Ctx& c, DevBoard& b, const int32_t* owner, const int32_t* halo);
void board_clear_legal(DevBoard& b);
// Drop the board's additive price plane.
void board_clear_additive(Ctx& c, DevBoard& b);

// Solve complete fields for `reqs` (FieldReq::board indexes `boards`) in one
// launch.  Labels stay on the device in the workspace until the next solve.
void solve_fields(Ctx& c, const std::vector<const DevBoard*>& boards, const std::vector<FieldReq>& reqs,
                  SolveStats* st);
inline void solve_fields(Ctx& c, const DevBoard& b, const std::vector<FieldReq>& reqs, SolveStats* st) {
  solve_fields(c, std::vector<const DevBoard*>{&b}, reqs, st);
}
// After solve_fields: reconstruct canonical paths for fields with a dst.
// routed[i]=0 when the dst is unreachable.
void reconstruct_paths(Ctx& c, size_t n_fields, std::vector<uint8_t>& routed, std::vector<uint32_t>& cost,
                       std::vector<uint64_t>& off, std::vector<uint32_t>& cells, SolveStats* st);
inline void reconstruct_paths(Ctx& c, const DevBoard&, size_t n, std::vector<uint8_t>& routed,
                              std::vector<uint32_t>& cost, std::vector<uint64_t>& off,
                              std::vector<uint32_t>& cells, SolveStats* st) {
  reconstruct_paths(c, n, routed, cost, off, cells, st);
}
// After solve_fields: copy field i's dist (and hops) out in its board's full
// layout (cells outside the window are kMax).
void read_field(Ctx& c, size_t field, uint32_t* dist, uint32_t* hops);
inline void read_field(Ctx& c, const DevBoard&, size_t field, uint32_t* dist, uint32_t* hops) {
  read_field(c, field, dist, hops);
}
// Device pointer variant used by tests of the persistent kernel: copy all
// fields as window-local row-major (dist) into a host buffer.
void ctx_init(Ctx& c, int device);
void ctx_free(Ctx& c);
size_t field_label_words(const TileShape& s, uint32_t ww, uint32_t wh, uint32_t layers);
bool tile_shape_supported(uint32_t tw, uint32_t th);
size_t lds_bytes_for(uint32_t tw, uint32_t th, uint32_t layers);  // LDS of one tile visit

// ---- round driver: fields join and leave a running relaxation -------------
//
// The serving loop keeps one relaxation running.  A group (the fields of one
// request) is admitted at a round boundary: it gets contiguous ranges of
// field descriptors, work items, label words and pool words, and its seed
// tiles join the next round's active list.  After every round the driver
// reports the groups with no active item left; their labels are final (the
// fixpoint of a field does not depend on the other fields), so their paths
// are reconstructed and their ranges freed while the other groups keep
// relaxing.

// First-fit allocator of ranges in [0, cap).
class RangeAlloc {
 public:
  void reset(uint64_t cap);
  bool alloc(uint64_t n, uint64_t& off);
  void release(uint64_t off, uint64_t n);
  uint64_t cap() const { return cap_; }
  uint64_t used() const { return used_; }

 private:
  uint64_t cap_ = 0, used_ = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length
};

struct Group {
  bool live = false;
  bool narrow = false;  // every step price fits the narrow step planes
  bool lean = false;    // no halo-owner / legal planes, overrides, pads or via pads
  bool wave = false;    // the wave form covers every field (no weight plane, legal planes, overrides, pads)
  bool halo = false;    // some board of its fields has a halo-owner plane
  bool ex = false;      // some field has own pads or via-exempt pads
  bool steps = false;   // every board of its fields has step price planes
  // Packed u32 labels (the wave form on boards whose every open step costs
  // at least 1 and below 0xFFFF); live groups all share one label width.
  bool pack = false;
  uint64_t pk_step = 0;  // largest step price of its boards (the certificate)
  uint64_t field0 = 0, nfields = 0;
  uint64_t item0 = 0, nitems = 0;
  uint64_t lab0 = 0, labwords = 0;
  uint64_t pool0 = 0, npool = 0;
  uint32_t layers = 0;  // most layers of any of its boards
  bool active = false;  // has an item on the current active list
};

struct RoundDriver {
  Ctx* c = nullptr;
  uint32_t max_layers = 0, max_boards = 0, max_groups = 0;
  // device arrays
  uint64_t* labels = nullptr;
  uint32_t* state = nullptr;
  uint32_t* act = nullptr;   // 2 x max_items
  uint32_t* list = nullptr;  // 2 x max_items
  uint32_t* item_field = nullptr;
  uint32_t* item_group = nullptr;
  uint32_t* seeds = nullptr;  // staging for seed items of newly admitted groups
  FieldDesc* fields = nullptr;
  uint32_t* pool = nullptr;
  BoardView* boards = nullptr;
  // relaxation counters (kCtrWords) followed by one activity flag per group
  // (the last round in which one of its visits activated a tile, plus one)
  uint32_t* ctr = nullptr;
  uint32_t* shards = nullptr;  // sharded visit / round / done counters (engine.hip kShardWords)
  uint32_t* imax = nullptr;    // per item: the largest packed label its visits stored
  uint32_t* pmax = nullptr;    // a packed group's largest label (one word)
  uint32_t* pinned = nullptr;  // readback of ctr (statistics)
  // host-mapped words the relaxation publishes to: next count, error flags,
  // stamp, then the group flags from kDriverFlagsAt
  uint32_t* hostctr = nullptr;
  uint32_t* hostctr_dev = nullptr;
  uint32_t stamp = 0;
  uint32_t seen_visits = 0, seen_lrounds = 0;
  std::vector<hipEvent_t> ev_free;
  std::deque<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
  RangeAlloc fa, ia, la, pa;   // fields, items, label words, pool words
  std::vector<BoardView> views;
  std::vector<const DevBoard*> devs;  // caller-owned boards, by slot
  std::vector<uint8_t> board_used;
  std::vector<Group> groups;
  std::vector<uint32_t> pending_seeds;
  std::vector<std::vector<uint8_t>> staging;  // host buffers of uploads in flight
  uint32_t n_cur = 0;  // the round's list length
  uint32_t cur = 0;  // which half of act/list is the current round's
  uint64_t rounds = 0, tile_visits = 0, local_rounds = 0;
  // Round index stored in item states at first visits (below 2^30; restarted
  // at 1 after every item state is rewritten to round 0).
  uint32_t stamp_round = 1;
  double kernel_ms = 0;
};

// label_bytes: HBM for labels of admitted groups (the admission budget).
void driver_init(RoundDriver& d, Ctx& c, size_t label_bytes, uint32_t max_boards, uint32_t max_layers,
                 uint32_t max_groups);
void driver_free(RoundDriver& d);
size_t driver_device_bytes(const RoundDriver& d);
// The board must outlive its slot (driver_remove_board).
int driver_add_board(RoundDriver& d, const DevBoard& b);  // slot, or -1 when the table is full
void driver_remove_board(RoundDriver& d, int slot);
// Label words a group of fields needs (FieldReq::board is a driver slot).
uint64_t driver_group_words(const RoundDriver& d, const std::vector<FieldReq>& reqs);
// Admit a group at the next round boundary; false when it does not fit now
// (memory, slots, or live groups of the other label width).  allow_pack:
// the group may run on packed labels where its boards allow.
bool driver_admit(RoundDriver& d, const std::vector<FieldReq>& reqs, int& group, bool allow_pack = true);
// Tiles the next round will visit: the active list plus admitted seeds.
inline uint32_t driver_next_tiles(const RoundDriver& d) { return d.n_cur + (uint32_t)d.pending_seeds.size(); }
// One relaxation round over every active item (new groups' seeds included).
// Returns the number of items active for the next round.
uint32_t driver_step(RoundDriver& d);
// Reads the device's visit and in-LDS round counters and the kernel times of
// finished rounds into d's statistics (one copy and stream synchronisation).
void driver_sync_stats(RoundDriver& d);
// Live groups with no item on the active list: their fields are complete.
void driver_finished(RoundDriver& d, std::vector<int>& out);
// Paths of a finished group.  False (and no paths) when the group ran on
// packed labels and its certificate failed: a label came within one step of
// the packed limit, so the group must be solved again on wide labels
// (driver_admit with allow_pack false).
bool driver_paths(RoundDriver& d, int group, std::vector<uint8_t>& routed, std::vector<uint32_t>& cost,
                  std::vector<uint64_t>& off, std::vector<uint32_t>& cells);
void driver_release(RoundDriver& d, int group);

}  // namespace mrf
