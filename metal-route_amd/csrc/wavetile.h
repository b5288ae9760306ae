// Wave form of the tile visit: one wavefront per (field, tile) item.
// Included by engine.hip after relax_round (it uses RelaxArgs, TileResult,
// round_tail and the device helpers defined there).
//
// The workgroup form (process_tile) runs a tile on four wavefronts; in its
// row and column phases a 16 x 16 x 4 tile has 64 lines, so one wavefront
// sweeps while three wait at the barrier, and the tile's 25.8 KB of LDS
// (labels with a halo ring plus six step planes) lets six visits share a CU.
// Here one wavefront owns a 16 x 16 tile of up to four layers (two
// wavefronts up to eight layers: NW below):
//   * labels live in an LDS buffer [layer][y][x] (row stride TW + 1, odd, so
//     a lane per row and a lane per column both hit distinct banks); each
//     phase reads the lines it sweeps into registers in its own layout --
//     a thread per row (layer, y), a thread per column (layer, x), or
//     TW*TH/(64 NW) via columns per thread -- sweeps them forward and
//     backward and writes a changed line back.  With L*TH = L*TW = 64 NW
//     every thread sweeps in every phase.  Two wavefronts meet only at the
//     phase barriers (the phase's change flag is their joint ballot).
//   * the six step prices of every cell are computed once per visit (in the
//     row layout, where the board planes are loaded) and kept in registers,
//     two 16-bit prices per register, in the layout of the phase that reads
//     them: x prices in the row layout, y prices moved to the column layout
//     and via prices to the via layout through the label buffer before the
//     labels go in.
//   * the halo labels (the neighbours' border cells) and the prices of
//     stepping from the tile into them stay in registers of the row layout
//     (left, right) and the column layout (top, bottom).
// The LDS of a visit is the label buffer and the tile's border labels as
// loaded (to tell which border cells this visit improved): 10.5 KB for
// 16 x 16 x 4, so the register file, not the LDS, bounds the visits per CU.
//
// Labels come in two forms (template parameter PK):
//   * wide: u64 (dist << 32) | hops, the form every other kernel uses;
//   * packed: u32 (dist << 16) | hops, 4 B per cell in HBM and LDS.  The
//     step of price s is (s << 16) | 1, added with saturation, so the packed
//     minimum over paths is min(dist * 2^16 + hops).  That equals the
//     lexicographic (dist, hops) minimum at every cell whose dist is below
//     0xFFFF and whose hops are below 2^16; on a board whose every open step
//     costs at least 1, hops <= dist, so dist < 0xFFFF suffices.  A closed
//     step is 0xFFFF << 16 | 1, which saturates from any label with dist >= 1
//     and from a source (dist 0) gives a label with dist field 0xFFFF that
//     never propagates and never beats a real label: dist field 0xFFFF reads
//     as unreached.  Every visit stores the largest label it writes back
//     (excluding those) for its tile; the host accepts the packed result
//     only if the largest of them plus the largest packed step stays below
//     0xFFFF << 16 (then no path to any cell reaches dist 0xFFFF), and
//     solves the batch again in the wide form otherwise.
//
// Scope: fields without legal planes (halo-owner planes, own pads,
// via-exempt pads and additive overrides are applied per tile: HO, EX) on
// boards without a weight plane whose step prices fit the narrow (15-bit)
// form, in the round form without delta stepping.  A forced endpoint on an
// obstacle makes a field wide, so every passable cell here is a cell that is
// neither a trace node nor an obstacle and every open step price is e + a.
// The host picks this kernel only for such launches.

// Wavefronts per tile visit: one up to 64 rows (4 layers of 16x16), two up
// to 128 (8 layers).
__host__ __device__ constexpr uint32_t wave_count(uint32_t lm, uint32_t th) { return (lm * th + 63) / 64; }

template <int TW, int TH, int LM, bool PK = false>
struct WaveTile {
  static constexpr uint32_t NW = wave_count(LM, TH);  // wavefronts
  static constexpr uint32_t NT = 64 * NW;             // threads
  static_assert(LM * TH <= NT && LM * TW <= NT, "a lane per row and a lane per column");
  static_assert((TW * TH) % NT == 0 && NT % TW == 0, "whole via columns per lane; a fixed x per lane in block order");
  static_assert(TH == TW, "the column layout reads rows cl * TH + y of the NT-row buffer");
  static constexpr uint32_t RS = TW + 1;         // label buffer row stride
  static constexpr uint32_t PT = TW * TH;        // cells of one layer of the tile
  static constexpr uint32_t NV = PT / NT;        // via columns per lane
  static constexpr uint32_t NS = LM * PT / NT;   // cells per lane in block order
  // label slots: NT rows whatever the layer bound, so the row layout's
  // thread index is a row of the buffer for every thread
  static constexpr uint32_t LAB = NT * RS;
  // border labels as loaded: left [LM*TH], right [LM*TH], top [LM*TW], bottom [LM*TW]
  static constexpr uint32_t BORDER = 2 * LM * TH + 2 * LM * TW;
  // then 16 words for the wavefronts' flags and maxima (two waves)
  static constexpr size_t lds = (size_t)(LAB + BORDER) * (PK ? 4 : 8) + 64;
  // per-net extras (own pads, via-exempt pads, additive overrides): the
  // tile's raw via flags, four to a word, and two bitmaps of its cells
  static constexpr uint32_t EXW = LM * PT / 4 + 2 * (LM * PT / 32);
};

__host__ __device__ inline size_t wave_lds_bytes(uint32_t tw, uint32_t th, uint32_t lm, bool pk, bool ex) {
  return (64ull * wave_count(lm, th) * (tw + 1) + 2ull * lm * th + 2ull * lm * tw) * (pk ? 4 : 8) + 64 +
         (ex ? 4ull * (lm * tw * th / 4 + 2 * (lm * tw * th / 32)) : 0);
}

// Whether x holds in any thread of a visit's NW wavefronts; every wave has
// passed a barrier on return.  With two waves each writes its ballot to
// slot[wave] of the pair the caller alternates between consecutive calls, so
// a wave that runs ahead never overwrites a flag another has yet to read
// (a wave writing the pair again has passed the barrier of the call between,
// which every wave reaches only after its reads).
template <uint32_t NW>
__device__ __forceinline__ bool visit_any(bool x, volatile uint32_t* slot) {
  const uint64_t b = __ballot(x);
  if constexpr (NW == 1) {
    __syncthreads();
    return b != 0;
  } else {
    if ((threadIdx.x & 63u) == 0) slot[threadIdx.x >> 6] = b != 0 ? 1u : 0u;
    __syncthreads();
    uint32_t r = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) r |= slot[w];
    return r != 0;
  }
}

// 16-bit step prices, two to a register: lo = entering from the lower
// neighbour (x-1, y-1, layer-1), hi = from the upper one; -1 = closed.  The
// sign-extended value is the u32 price, kMax when closed.
__device__ __forceinline__ uint32_t step_lo(uint32_t w) { return (uint32_t)(int32_t)(int16_t)(uint16_t)w; }
__device__ __forceinline__ uint32_t step_hi(uint32_t w) { return (uint32_t)((int32_t)w >> 16); }
__device__ __forceinline__ uint32_t pack_steps(uint32_t lo, uint32_t hi) { return (lo & 0xFFFFu) | (hi << 16); }

// relax_to without the per-cell change bit: returns whether `cur` improved.
__device__ __forceinline__ bool relax_w(uint64_t from, uint32_t s, uint64_t& cur) {
  const uint32_t hi = __builtin_elementwise_add_sat((uint32_t)(from >> 32), s);
  const uint32_t lo = (uint32_t)from + 1u;
  const uint64_t cand = ((uint64_t)hi << 32) | lo;
  const bool better = cand < cur;
  cur = better ? cand : cur;
  return better;
}

// Candidate a border label offers the neighbour's adjacent cell through an
// edge of price p (kMax when closed saturates, never below any label).
__device__ __forceinline__ uint64_t offer(uint64_t v, uint32_t p) {
  return ((uint64_t)__builtin_elementwise_add_sat((uint32_t)(v >> 32), p) << 32) | (uint32_t)((uint32_t)v + 1u);
}

// The two label forms: type, unreached value, relaxation through the lower
// or upper step of a packed pair, the candidate a label offers through a
// price pair half, and the halo value taken from a neighbour.
template <bool PK>
struct WaveLab;

template <>
struct WaveLab<false> {
  using T = uint64_t;
  static constexpr uint64_t unr = kUnreached;
  __device__ static bool relax_lo(uint64_t from, uint32_t pair, uint64_t& cur) { return relax_w(from, step_lo(pair), cur); }
  __device__ static bool relax_hi(uint64_t from, uint32_t pair, uint64_t& cur) { return relax_w(from, step_hi(pair), cur); }
  __device__ static uint64_t offer_lo(uint64_t v, uint32_t pair) { return offer(v, step_lo(pair)); }
  __device__ static uint64_t offer_hi(uint64_t v, uint32_t pair) { return offer(v, step_hi(pair)); }
  __device__ static uint64_t halo(uint64_t v) { return lab_in(v); }
};

// Packed labels whose dist field is 0xFFFF read as unreached.
constexpr uint32_t kPackedUnreached = 0xFFFF0000u;

template <>
struct WaveLab<true> {
  using T = uint32_t;
  static constexpr uint32_t unr = 0xFFFFFFFFu;
  __device__ static uint32_t w_lo(uint32_t pair) { return (pair << 16) | 1u; }
  __device__ static uint32_t w_hi(uint32_t pair) { return (pair & 0xFFFF0000u) | 1u; }
  __device__ static bool relax(uint32_t from, uint32_t w, uint32_t& cur) {
    const uint32_t cand = __builtin_elementwise_add_sat(from, w);
    const bool better = cand < cur;
    cur = better ? cand : cur;
    return better;
  }
  __device__ static bool relax_lo(uint32_t from, uint32_t pair, uint32_t& cur) { return relax(from, w_lo(pair), cur); }
  __device__ static bool relax_hi(uint32_t from, uint32_t pair, uint32_t& cur) { return relax(from, w_hi(pair), cur); }
  __device__ static uint32_t offer_lo(uint32_t v, uint32_t pair) { return __builtin_elementwise_add_sat(v, w_lo(pair)); }
  __device__ static uint32_t offer_hi(uint32_t v, uint32_t pair) { return __builtin_elementwise_add_sat(v, w_hi(pair)); }
  __device__ static uint32_t halo(uint32_t v) { return v; }
};

// Narrow step price of entering a cell with board mask bm and additive
// price a through the edge whose open bit is `edge`, over gap price e; the
// extra condition `ok` closes it too (layer bounds).
__device__ __forceinline__ uint32_t wave_step(uint8_t bm, uint8_t edge, bool ok, uint32_t e, uint32_t a) {
  const bool open = ok && !(bm & (M_TRACE | M_OBST)) && (bm & edge);
  return open ? e + a : kMax;
}

// Additive plane of a board: none, u16 or u32 (a template parameter of the
// visit, so no load waits on a pointer test).
enum AddKind : int { ADD_NONE = 0, ADD_U16 = 1, ADD_U32 = 2 };

template <int AK>
__device__ __forceinline__ uint32_t add_k(const BoardView& B, uint32_t g) {
  if constexpr (AK == ADD_U16) return gbl(B.add16)[g];
  if constexpr (AK == ADD_U32) return gbl(B.add)[g];
  return 0u;
}

// HO: the board has a halo-owner plane; a cell whose owner is another net
// costs the board's halo penalty on top of its additive price (add_price).
// EX: the field may carry own pads (obstacle cells passable at weight 1),
// via-exempt pads (via_forbidden lifted) and additive overrides (the net's
// own price of a cell), applied per tile from the host's tables.
// QU: a visit of the persistent queue form: labels start unreached (filled
// before the launch) and other workgroups write them while this one runs, so
// every label is read with a device-scope load (another XCD's L2 may hold a
// stale copy) and every existing neighbour's border counts; `touched` is
// false on the tile's first visit (seeds) only.
// SP: the board's step price planes (B.steps) hold every cell's prices: the
// visit loads them in the layout of each phase instead of forming them from
// the masks and planes (lean fields only: no HO, no EX; AK unused).
template <int TW, int TH, int LM, int AK, bool PK, bool HO, bool EX, bool QU = false, bool SP = false>
__device__ __forceinline__ TileResult wave_tile(const RelaxArgs& A, uint32_t item, uint32_t f, bool touched, const FieldDesc& F,
                                const BoardView& B, char* smem, uint32_t round) {
  static_assert(!(SP && (HO || EX)), "step price planes are for lean fields");
  using WT = WaveTile<TW, TH, LM, PK>;
  using WL = WaveLab<PK>;
  using LT = typename WL::T;
  constexpr uint32_t RS = WT::RS, PT = WT::PT, NV = WT::NV, NS = WT::NS, LAB = WT::LAB, NT = WT::NT, NW = WT::NW;
  LT* lab = (LT*)smem;
  // step staging before the labels go in: y prices, then via prices (packed
  // labels: one after the other in the first LAB words; wide: side by side)
  uint32_t* lab32 = (uint32_t*)smem;
  constexpr uint32_t VOFF = PK ? 0u : LAB;
  LT* b0 = lab + LAB;
  // the wavefronts' flags (two alternating pairs) and results
  volatile uint32_t* xw = (volatile uint32_t*)(smem + (size_t)(LAB + WT::BORDER) * sizeof(LT));
  uint32_t par = 0;  // which flag pair the next visit_any uses
  // tid: the thread's place in the tile layouts; lane: its lane in its wave
  const uint32_t tid = threadIdx.x, lane = NW == 1 ? tid : tid & 63u;

  const uint32_t t = item - F.item0;
  const uint32_t txi = t % F.tx, tyi = t / F.tx;
  const uint32_t L = B.layers;
  const uint32_t ncell = L * PT;
  const uint32_t wx0 = txi * TW, wy0 = tyi * TH;
  const uint64_t TB = tile_block_words(L, TW, TH);
  const uint64_t blockbase = F.lab_off + (uint64_t)t * TB;
  const bool has_l = txi > 0, has_r = txi + 1 < F.tx, has_t = tyi > 0, has_b = tyi + 1 < F.ty;

  // ---- loads, all issued before the first use, in the order they are used ------
  // (vector loads complete in issue order: the step prices wait only for the
  // planes, the tile's labels are still in flight meanwhile)
  // row layout: lane (rl, ry) holds row ry of layer rl
  const uint32_t rl = tid / TH, ry = tid - rl * TH;
  const bool rrow = rl < L;
  const bool row_in = rrow && wy0 + ry < F.wh;
  const uint32_t rgy = F.y0 + min(wy0 + ry, F.wh - 1);
  const uint32_t rowg = ((rrow ? rl : 0u) * B.h + rgy) * B.w + F.x0;  // board cell of window column 0
  // The row's planes as packed words: mask bytes four to a word, u16
  // additive prices two to a word, u32 ones one to a word.  Rows that start
  // on a multiple of four cells (every row when the board width and the
  // tile's first column are) load whole words; others, and tiles past the
  // window's edge, load cell by cell (clamped to the window) and pack.
  constexpr uint32_t AW = AK == ADD_U32 ? TW : AK == ADD_U16 ? TW / 2 : 1;  // additive words
  uint32_t mw[TW / 4], aw[AW];
  uint32_t sx[TW];  // row layout: from x-1 | from x+1
  // SP: the row's x price pairs, one word per cell (clamped to the window;
  // cells past its edge are closed below)
  const uint64_t nb = (uint64_t)B.plane * L;  // cells of the board
  if constexpr (SP) {
    const auto* PX = gbl(B.steps);
    if (wx0 + TW <= F.ww) {
#pragma unroll
      for (int x = 0; x < TW; ++x) sx[x] = PX[rowg + wx0 + x];
    } else {
#pragma unroll
      for (int x = 0; x < TW; ++x) sx[x] = PX[rowg + min(wx0 + x, F.ww - 1)];
    }
  } else if ((B.w % 4) == 0 && ((F.x0 + wx0) % 4) == 0 && wx0 + TW <= F.ww) {
    const uint32_t g0 = rowg + wx0;
    const auto* mp = (const __attribute__((address_space(1))) uint32_t*)(B.bmask + g0);
#pragma unroll
    for (uint32_t q = 0; q < TW / 4; ++q) mw[q] = mp[q];
    if constexpr (AK == ADD_U16) {
      const auto* ap = (const __attribute__((address_space(1))) uint32_t*)(B.add16 + g0);
#pragma unroll
      for (uint32_t q = 0; q < AW; ++q) aw[q] = ap[q];
    }
    if constexpr (AK == ADD_U32) {
#pragma unroll
      for (uint32_t q = 0; q < AW; ++q) aw[q] = gbl(B.add)[g0 + q];
    }
  } else {
    uint32_t bq[TW], aq[TW];
#pragma unroll
    for (int x = 0; x < TW; ++x) {
      const uint32_t g = rowg + min(wx0 + x, F.ww - 1);
      bq[x] = gbl(B.bmask)[g];
      aq[x] = add_k<AK>(B, g);
    }
#pragma unroll
    for (uint32_t q = 0; q < TW / 4; ++q)
      mw[q] = bq[4 * q] | (bq[4 * q + 1] << 8) | (bq[4 * q + 2] << 16) | (bq[4 * q + 3] << 24);
    if constexpr (AK == ADD_U16) {
#pragma unroll
      for (uint32_t q = 0; q < AW; ++q) aw[q] = aq[2 * q] | (aq[2 * q + 1] << 16);
    }
    if constexpr (AK == ADD_U32) {
#pragma unroll
      for (uint32_t q = 0; q < AW; ++q) aw[q] = aq[q];
    }
  }
  if constexpr (AK == ADD_NONE) aw[0] = 0;
  // foreign-halo cells of the row, one bit per cell
  uint32_t hom = 0;
  if constexpr (HO) {
    const int32_t own = F.own_id;
#pragma unroll
    for (int x = 0; x < TW; ++x) {
      const int32_t o = gbl(B.halo)[rowg + min(wx0 + x, F.ww - 1)];
      hom |= (o != -1 && o != own) ? 1u << x : 0u;
    }
  }
  // ---- per-net extras of the tile ------------------------------------------------
  // Own-pad cells and override cells of each row as bit masks (padm, ovm),
  // override prices staged at the row's slots of the label buffer, and the
  // row's via bits made again where via-exempt pads lift via_forbidden.
  // (before the tile's labels are loaded: their registers are not held
  // through this pass's barriers)
  uint32_t padm = 0, ovm = 0, v0 = 0, v1 = 0;
  bool tile_pads = false;
  uint32_t* raww = (uint32_t*)(smem + WT::lds);  // [LM*PT/4]
  uint32_t* padb = raww + LM * PT / 4;           // [LM*PT/32]
  if constexpr (EX) {
    uint32_t* ovb = padb + LM * PT / 32;           // [LM*PT/32]
    const uint32_t ntile = F.tx * F.ty;
    const uint32_t* P = A.pool + F.xtab_off;
    uint32_t p0 = 0, p1 = 0, o0 = 0, o1 = 0;
    if (F.n_pad | F.n_vpad) {
      p0 = const_load(P + t);
      p1 = const_load(P + t + 1);
      v0 = const_load(P + ntile + 1 + t);
      v1 = const_load(P + ntile + 2 + t);
    }
    const uint32_t* OT = A.pool + F.ovr_off + 2 * F.n_ovr;  // per-tile starts, cells, prices
    if (F.n_ovr) {
      o0 = const_load(OT + t);
      o1 = const_load(OT + t + 1);
    }
    tile_pads = p1 > p0;
    if (p1 > p0 || v1 > v0 || o1 > o0) {
      for (uint32_t i = tid; i < 2 * (LM * PT / 32); i += NT) padb[i] = 0;
      if (v1 > v0 && rrow) {
#pragma unroll
        for (uint32_t q = 0; q < TW / 4; ++q) {
          uint32_t wd = 0;
#pragma unroll
          for (uint32_t b = 0; b < 4; ++b) wd |= (uint32_t)gbl(B.raw)[rowg + min(wx0 + 4 * q + b, F.ww - 1)] << (8 * b);
          raww[(rl * TH + ry) * (TW / 4) + q] = wd;
        }
      }
      __syncthreads();
      const uint32_t* PE = P + 2 * (ntile + 1);
      for (uint32_t k = p0 + tid; k < p1; k += NT) {
        const uint32_t loc = PE[k];
        atomicOr(&padb[loc >> 5], 1u << (loc & 31));
      }
      const uint32_t* VE = PE + const_load(P + ntile);
      for (uint32_t k = v0 + tid; k < v1; k += NT) {
        const uint32_t loc = VE[k];
        atomicAnd(&raww[loc >> 2], ~((uint32_t)R_VF << (8 * (loc & 3))));
      }
      const uint32_t* OC = OT + ntile + 1;
      const uint32_t* OA = OC + F.n_ovr;
      for (uint32_t k = o0 + tid; k < o1; k += NT) {
        const uint32_t g = OC[k];
        const uint32_t l = g / B.plane, r = g - l * B.plane, gy = r / B.w, gx = r - gy * B.w;
        const uint32_t x = gx - F.x0 - wx0, y = gy - F.y0 - wy0, loc = (l * TH + y) * TW + x;
        lab32[(l * TH + y) * RS + x] = OA[k];
        atomicOr(&ovb[loc >> 5], 1u << (loc & 31));
      }
      __syncthreads();
      const uint32_t rloc = (rl * TH + ry) * TW;  // a multiple of 16
      if (rrow) {
        padm = (padb[rloc >> 5] >> (rloc & 31)) & 0xFFFFu;
        ovm = (ovb[rloc >> 5] >> (rloc & 31)) & 0xFFFFu;
      }
    }
  }

  // column layout: thread (cl, cx) holds column cx of layer cl
  const uint32_t cl = tid / TW, cx = tid - cl * TW;
  const bool ccol = cl < L;
  const bool col_in = ccol && wx0 + cx < F.ww;
  // Board cells of the halo (the price of stepping from the tile into them)
  // and, further down, the halo labels (kept only if the neighbour's first
  // visit ran in an earlier round).  Every load is issued: where a neighbour
  // or the cell does not exist the address falls back to one inside this
  // tile, and the value is dropped.
  const bool el = row_in && has_l, er = row_in && has_r, et = col_in && has_t, eb = col_in && has_b;
  uint8_t hbm[4];
  uint32_t ha[4];
  uint32_t hsw[4];  // SP: the halo cells' x (left, right) and y (top, bottom) price pairs
  // SP: the column and via layouts' price pairs, issued with the rows'
  uint32_t sy[TH];      // column layout: from y-1 | from y+1
  uint32_t sv[NV][LM];  // via layout: from layer-1 | from layer+1
  constexpr uint32_t RPJ = NT / TW;  // rows per NT cells
  const uint32_t sx_ = tid % TW, sy_ = tid / TW;
  if constexpr (SP) {
    const uint32_t cgx = F.x0 + wx0 + cx;
    const auto* PX = gbl(B.steps);
    const auto* PY = gbl(B.steps + nb);
    const auto* PV = gbl(B.steps + 2 * nb);
    hsw[0] = PX[el ? rowg + wx0 - 1 : rowg];
    hsw[1] = PX[er ? rowg + wx0 + TW : rowg];
    hsw[2] = PY[et ? (cl * B.h + F.y0 + wy0 - 1) * B.w + cgx : rowg];
    hsw[3] = PY[eb ? (cl * B.h + F.y0 + wy0 + TH) * B.w + cgx : rowg];
    const uint32_t ccg = F.x0 + min(wx0 + cx, F.ww - 1);  // the column's board x, clamped
    const uint32_t cbase = ((ccol ? cl : 0u) * B.h + F.y0) * B.w + ccg;
#pragma unroll
    for (int y = 0; y < TH; ++y) sy[y] = PY[cbase + min(wy0 + y, F.wh - 1) * B.w];
    const uint32_t vg = F.x0 + min(wx0 + sx_, F.ww - 1);
#pragma unroll
    for (uint32_t k = 0; k < NV; ++k) {
      const uint32_t vy = F.y0 + min(wy0 + sy_ + RPJ * k, F.wh - 1);
#pragma unroll
      for (int l = 0; l < LM; ++l) sv[k][l] = PV[(min((uint32_t)l, L - 1) * B.h + vy) * B.w + vg];
    }
  } else {
    const uint32_t cgx = F.x0 + wx0 + cx;
    const uint32_t g[4] = {el ? rowg + wx0 - 1 : rowg, er ? rowg + wx0 + TW : rowg,
                           et ? (cl * B.h + F.y0 + wy0 - 1) * B.w + cgx : rowg,
                           eb ? (cl * B.h + F.y0 + wy0 + TH) * B.w + cgx : rowg};
    const bool e[4] = {el, er, et, eb};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint8_t m = gbl(B.bmask)[g[k]];
      ha[k] = add_k<AK>(B, g[k]);
      if constexpr (HO) {
        const int32_t o = gbl(B.halo)[g[k]];
        if (o != -1 && o != F.own_id) ha[k] += B.halo_pen;
      }
      hbm[k] = e[k] ? m : (uint8_t)M_TRACE;
    }
    if constexpr (EX) {
      // The halo prices only decide activations, so a lower bound is
      // enough: with own pads any obstacle may be one (passable), with
      // overrides any additive price may be lower (0).
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (F.n_pad) hbm[k] &= (uint8_t)~M_OBST;
        if (F.n_ovr) ha[k] = 0;
      }
    }
  }
  // gap prices: x gaps uniform (lane i <= TW holds the price of entering
  // tile column i from column i-1), y gaps and via prices per lane
  uint32_t xgv = 0, yg0 = 0, yg1 = 0, vlo = 0, vhi = 0, ytop = 0, ybot = 0;
  if constexpr (!SP) {
    const uint32_t gx = F.x0 + wx0 + lane;
    if (lane <= TW && gx >= 1 && gx < B.w) xgv = gbl(B.xe)[gx - 1];
    const uint32_t gyy = F.y0 + wy0 + ry;
    yg0 = (gyy >= 1 && gyy < B.h) ? gbl(B.ye)[gyy - 1] : 0u;
    yg1 = (gyy + 1 < B.h) ? gbl(B.ye)[gyy] : 0u;
    vlo = (rrow && rl > 0) ? gbl(B.ve)[rl - 1] : 0u;
    vhi = (rrow && rl + 1 < L) ? gbl(B.ve)[rl] : 0u;
    const uint32_t gyt = F.y0 + wy0, gyb = F.y0 + wy0 + TH;
    ytop = (gyt >= 1 && gyt < B.h) ? gbl(B.ye)[gyt - 1] : 0u;
    ybot = (gyb >= 1 && gyb < B.h) ? gbl(B.ye)[gyb - 1] : 0u;
  }
  // neighbour states (lanes 1..4): whether their border labels exist
  uint32_t stv = 0;
  if (!QU && lane >= 1 && lane < 5) {
    uint32_t nb = item;
    bool ok = true;
    if (lane == 1) ok = has_l, nb = item - 1;
    if (lane == 2) ok = has_r, nb = item + 1;
    if (lane == 3) ok = has_t, nb = item - F.tx;
    if (lane == 4) ok = has_b, nb = item + F.tx;
    stv = ok ? __hip_atomic_load(&A.state[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
  }
  // Halo labels.  Labels are read and written with plain 8-byte accesses
  // (never torn): every label this visit takes was stored in an earlier
  // launch, and a neighbour that is visited again in this one only lowers
  // its labels, so whichever value a load returns is an upper bound.
  const uint64_t own = blockbase + tid % TW;
  const uint64_t hla = el ? blockbase - TB + ncell + L * TH + rl * TH + ry : own;
  const uint64_t hra = er ? blockbase + TB + ncell + rl * TH + ry : own;
  const uint64_t hta = et ? blockbase - (uint64_t)F.tx * TB + cl * PT + (TH - 1) * TW + cx : own;
  const uint64_t hba = eb ? blockbase + (uint64_t)F.tx * TB + cl * PT + cx : own;
  const auto* glab = gbl((const LT*)A.labels);
  auto lab_at = [&](uint64_t at) -> LT {
    if constexpr (QU)
      return __hip_atomic_load((const LT*)A.labels + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      return glab[at];
  };
  LT hl = lab_at(hla), hr = lab_at(hra), ht = lab_at(hta), hb = lab_at(hba);
  // The tile's labels in block order: thread tid holds cells tid + NT j,
  // i.e. column x = tid % TW of row (NT / TW) * (j % NV) + tid / TW of layer
  // j / NV -- the same x for every j, so each cell's LDS slot is a constant
  // offset from one per-thread base (`sbase`, also the via layout's base).
  // Layers beyond the board's re-read its last layer (dropped below).
  const uint32_t sbase = sy_ * RS + sx_;
  LT sl[NS];
  if (touched || QU) {
#pragma unroll
    for (uint32_t j = 0; j < NS; ++j) sl[j] = lab_at(blockbase + tid + NT * (min(j / NV, L - 1) * NV + j % NV));
  } else {
#pragma unroll
    for (uint32_t j = 0; j < NS; ++j) sl[j] = WL::unr;
  }

  auto earlier = [&](uint32_t st) { return (st & S_TOUCHED) && (st >> 1) < round; };
  const uint32_t nbv = QU ? (has_l ? 1u : 0u) | (has_r ? 2u : 0u) | (has_t ? 4u : 0u) | (has_b ? 8u : 0u)
                          : (earlier(__builtin_amdgcn_readlane(stv, 1)) ? 1u : 0u) |
                                (earlier(__builtin_amdgcn_readlane(stv, 2)) ? 2u : 0u) |
                                (earlier(__builtin_amdgcn_readlane(stv, 3)) ? 4u : 0u) |
                                (earlier(__builtin_amdgcn_readlane(stv, 4)) ? 8u : 0u);
  hl = (nbv & 1u) ? WL::halo(hl) : WL::unr;
  hr = (nbv & 2u) ? WL::halo(hr) : WL::unr;
  ht = (nbv & 4u) ? WL::halo(ht) : WL::unr;
  hb = (nbv & 8u) ? WL::halo(hb) : WL::unr;
  // prices of stepping out of the tile into the halo cells: the cell's edge
  // facing the tile must be open and the cell passable (narrow: at most 0x7FFE)
  uint32_t p_lr, p_tb;
  if constexpr (SP) {
    // entering the left / top halo cell from the tile is its step from x+1 /
    // y+1 (hi half), the right / bottom one's from x-1 / y-1 (lo half)
    p_lr = (el ? hsw[0] >> 16 : 0xFFFFu) | ((er ? hsw[1] & 0xFFFFu : 0xFFFFu) << 16);
    p_tb = (et ? hsw[2] >> 16 : 0xFFFFu) | ((eb ? hsw[3] & 0xFFFFu : 0xFFFFu) << 16);
  } else {
    const uint32_t xleft = __builtin_amdgcn_readlane(xgv, 0), xright = __builtin_amdgcn_readlane(xgv, TW);
    p_lr = pack_steps(wave_step(hbm[0], M_PX, true, xleft, ha[0]), wave_step(hbm[1], M_NX, true, xright, ha[1]));
    p_tb = pack_steps(wave_step(hbm[2], M_PY, true, ytop, ha[2]), wave_step(hbm[3], M_NY, true, ybot, ha[3]));
  }
  // formed here, so the halo cells' planes are not kept through the sweeps
  asm("" : "+v"(p_lr), "+v"(p_tb));

  // via bits of rows holding via-exempt pads, made again without the
  // exempt cells' via_forbidden flags (lifted in the LDS copy above)
  if constexpr (EX) {
    if (v1 > v0 && rrow) {
      // a via between layers l and l+1 at (x, y) is open when the pair
      // allows vias and neither end is a via node or via-forbidden
#pragma unroll
      for (uint32_t q = 0; q < TW / 4; ++q) {
        const uint32_t cur = raww[(rl * TH + ry) * (TW / 4) + q];
        const uint32_t dn = rl > 0 ? raww[((rl - 1) * TH + ry) * (TW / 4) + q] : 0x03030303u;
        const uint32_t up = rl + 1 < L ? raww[((rl + 1) * TH + ry) * (TW / 4) + q] : 0x03030303u;
        uint32_t wd = mw[q];
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
          const bool okc = ((cur >> (8 * b)) & 3u) == 0;
          const bool okd = rl > 0 && ((B.va >> (rl - 1)) & 1u) && okc && ((dn >> (8 * b)) & 3u) == 0;
          const bool oku = rl + 1 < L && ((B.va >> rl) & 1u) && okc && ((up >> (8 * b)) & 3u) == 0;
          wd = (wd & ~((uint32_t)(M_DN | M_UP) << (8 * b))) |
               ((uint32_t)((okd ? M_DN : 0) | (oku ? M_UP : 0)) << (8 * b));
        }
        mw[q] = wd;
      }
    }
  }

  // ---- step prices of every cell (row layout) ---------------------------------
  if constexpr (SP) {
    // loaded above; cells outside the window (edge tiles) and rows, columns
    // and via cells beyond the board's layers are closed, as the formed
    // prices close them
    if (wx0 + TW > F.ww || wy0 + TH > F.wh || L < LM) {
      const bool cin = ccol && wx0 + cx < F.ww, vin = wx0 + sx_ < F.ww;
#pragma unroll
      for (int x = 0; x < TW; ++x) sx[x] = (row_in && wx0 + x < F.ww) ? sx[x] : 0xFFFFFFFFu;
#pragma unroll
      for (int y = 0; y < TH; ++y) sy[y] = (cin && wy0 + y < F.wh) ? sy[y] : 0xFFFFFFFFu;
#pragma unroll
      for (uint32_t k = 0; k < NV; ++k)
#pragma unroll
        for (int l = 0; l < LM; ++l)
          sv[k][l] = ((uint32_t)l < L && vin && wy0 + sy_ + RPJ * k < F.wh) ? sv[k][l] : 0xFFFFFFFFu;
    }
  } else {
  uint32_t sv_st[TW];  // via prices of the row, staged after the y prices
  {
    uint32_t xg[TW + 1];
#pragma unroll
    for (int i = 0; i <= TW; ++i) xg[i] = __builtin_amdgcn_readlane(xgv, i);
#pragma unroll
    for (int x = 0; x < TW; ++x) {
      uint8_t mb = (uint8_t)(mw[x / 4] >> (8 * (x % 4)));
      if constexpr (EX) {
        if ((padm >> x) & 1u) mb &= (uint8_t)~M_OBST;  // an own pad: passable, weight 1
      }
      const uint8_t m = (row_in && wx0 + x < F.ww) ? mb : (uint8_t)M_TRACE;  // outside the window: closed
      uint32_t a = AK == ADD_U32 ? aw[x % AW] : AK == ADD_U16 ? (aw[(x / 2) % AW] >> (16 * (x % 2))) & 0xFFFFu : 0u;
      if constexpr (EX) {
        if ((ovm >> x) & 1u) a = lab32[tid * RS + x];  // the net's own price, staged above
      }
      if constexpr (HO) a += ((hom >> x) & 1u) ? B.halo_pen : 0u;
      uint32_t px = pack_steps(wave_step(m, M_NX, true, xg[x], a), wave_step(m, M_PX, true, xg[x + 1], a));
      // keep the packed form: the sweeps read each half with a sign-extending
      // operand select, no per-step gate
      asm("" : "+v"(px));
      sx[x] = px;
      // staged for the column and via layouts; lanes without a row
      // (layers beyond the board's) stage closed prices in rows no lane reads
      const uint32_t at = tid * RS + x;  // tid = rl * TH + ry
      lab32[at] = pack_steps(wave_step(m, M_NY, true, yg0, a), wave_step(m, M_PY, true, yg1, a));
      sv_st[x] = pack_steps(wave_step(m, M_DN, rl > 0, vlo, a), wave_step(m, M_UP, rl + 1 < L, vhi, a));
      if (!PK) lab32[VOFF + at] = sv_st[x];
    }
  }
  __syncthreads();
#pragma unroll
  for (int y = 0; y < TH; ++y) sy[y] = ccol ? lab32[(cl * TH + y) * RS + cx] : kMax;
  if (PK) {
    __syncthreads();
#pragma unroll
    for (int x = 0; x < TW; ++x) lab32[tid * RS + x] = sv_st[x];
    __syncthreads();
  }
#pragma unroll
  for (uint32_t k = 0; k < NV; ++k) {
#pragma unroll
    for (int l = 0; l < LM; ++l) sv[k][l] = (uint32_t)l < L ? lab32[VOFF + sbase + (l * TH + RPJ * k) * RS] : kMax;
  }
  __syncthreads();
  }
  // ---- labels into the buffer -------------------------------------------------
  const bool sx_in = wx0 + sx_ < F.ww;
#pragma unroll
  for (uint32_t j = 0; j < NS; ++j) {
    const uint32_t l = j / NV, y = RPJ * (j % NV);  // + sy_
    const bool in = l < L && sx_in && wy0 + y + sy_ < F.wh;
    lab[sbase + (l * TH + y) * RS] = in ? sl[j] : WL::unr;
  }
  __syncthreads();
  // the border labels as loaded, before seeds and sweeps change them
  if (rrow) {
    b0[rl * TH + ry] = lab[(rl * TH + ry) * RS];
    b0[LM * TH + rl * TH + ry] = lab[(rl * TH + ry) * RS + TW - 1];
  }
  if (ccol) {
    b0[2 * LM * TH + cl * TW + cx] = lab[(cl * TH) * RS + cx];
    b0[2 * LM * TH + LM * TW + cl * TW + cx] = lab[(cl * TH + TH - 1) * RS + cx];
  }
  // seeds (first visit only): passable sources start at label (0, 0)
  bool seeded = false;
  if (!touched) {
    __syncthreads();
    const uint32_t plane = B.plane;
    bool put = false;
    for (uint32_t k = tid; k < F.n_src; k += NT) {
      const uint32_t g = A.pool[F.src_off + k];
      const uint32_t l = g / plane, r = g - l * plane, gy = r / B.w, gx = r - gy * B.w;
      if (l >= L || gx < F.x0 + wx0 || gy < F.y0 + wy0) continue;
      const uint32_t x = gx - F.x0 - wx0, y = gy - F.y0 - wy0;
      if (x >= TW || y >= TH || wx0 + x >= F.ww || wy0 + y >= F.wh) continue;
      const uint8_t sm = gbl(B.bmask)[g];
      bool pass = !(sm & (M_TRACE | M_OBST));
      if constexpr (EX) {
        const uint32_t loc = (l * TH + y) * TW + x;
        pass = pass || (!(sm & M_TRACE) && tile_pads && ((padb[loc >> 5] >> (loc & 31)) & 1u));
      }
      if (pass) {
        lab[(l * TH + y) * RS + x] = LT(0);
        put = true;
      }
    }
    seeded = visit_any<NW>(put, xw + 2 * (par++ & 1u));
  }
  __syncthreads();

  // ---- relaxation to a local fixpoint ------------------------------------------
  const uint32_t nphase = L > 1 ? 3u : 2u;
  uint32_t k = 0, quiet = 0, rounds = 0;
  bool any = seeded, capped = true;
  for (rounds = 0; rounds < A.max_local; ++rounds) {
    // rows
    {
      bool ch = false;
      // the price pairs stay packed across rounds: derived prices are formed
      // inside each sweep, not kept for every cell
#pragma unroll
      for (int x = 0; x < TW; ++x) asm("" : "+v"(sx[x]));
      if (rrow) {
        const uint32_t base = (rl * TH + ry) * RS;
        LT v[TW];
#pragma unroll
        for (int x = 0; x < TW; ++x) v[x] = lab[base + x];
        LT p = hl;
#pragma unroll
        for (int x = 0; x < TW; ++x) {
          ch |= WL::relax_lo(p, sx[x], v[x]);
          p = v[x];
        }
        p = hr;
#pragma unroll
        for (int x = TW - 1; x >= 0; --x) {
          ch |= WL::relax_hi(p, sx[x], v[x]);
          p = v[x];
        }
        if (ch) {
#pragma unroll
          for (int x = 0; x < TW; ++x) lab[base + x] = v[x];
        }
      }
      const bool c = visit_any<NW>(ch, xw + 2 * (par++ & 1u));
      any = any || c;
      quiet = c ? 0u : quiet + 1;
      if (++k >= nphase && quiet + 1 >= nphase) {
        capped = false;
        break;
      }
    }
    // columns
    {
      bool ch = false;
#pragma unroll
      for (int y = 0; y < TH; ++y) asm("" : "+v"(sy[y]));
      if (ccol) {
        const uint32_t base = (cl * TH) * RS + cx;
        LT v[TH];
#pragma unroll
        for (int y = 0; y < TH; ++y) v[y] = lab[base + y * RS];
        LT p = ht;
#pragma unroll
        for (int y = 0; y < TH; ++y) {
          ch |= WL::relax_lo(p, sy[y], v[y]);
          p = v[y];
        }
        p = hb;
#pragma unroll
        for (int y = TH - 1; y >= 0; --y) {
          ch |= WL::relax_hi(p, sy[y], v[y]);
          p = v[y];
        }
        if (ch) {
#pragma unroll
          for (int y = 0; y < TH; ++y) lab[base + y * RS] = v[y];
        }
      }
      const bool c = visit_any<NW>(ch, xw + 2 * (par++ & 1u));
      any = any || c;
      quiet = c ? 0u : quiet + 1;
      if (++k >= nphase && quiet + 1 >= nphase) {
        capped = false;
        break;
      }
    }
    // via columns
    if (L > 1) {
      bool ch = false;
#pragma unroll
      for (uint32_t kk = 0; kk < NV; ++kk)
#pragma unroll
        for (int l = 0; l < LM; ++l) asm("" : "+v"(sv[kk][l]));
#pragma unroll
      for (uint32_t kk = 0; kk < NV; ++kk) {
        const uint32_t base = sbase + RPJ * kk * RS;
        LT v[LM];
#pragma unroll
        for (int l = 0; l < LM; ++l) v[l] = (uint32_t)l < L ? lab[l * TH * RS + base] : WL::unr;
        bool cc = false;
#pragma unroll
        for (int l = 1; l < LM; ++l) cc |= WL::relax_lo(v[l - 1], sv[kk][l], v[l]);
#pragma unroll
        for (int l = LM - 2; l >= 0; --l) cc |= WL::relax_hi(v[l + 1], sv[kk][l], v[l]);
        if (cc) {
#pragma unroll
          for (int l = 0; l < LM; ++l)
            if ((uint32_t)l < L) lab[l * TH * RS + base] = v[l];
        }
        ch = ch || cc;
      }
      const bool c = visit_any<NW>(ch, xw + 2 * (par++ & 1u));
      any = any || c;
      quiet = c ? 0u : quiet + 1;
      if (++k >= nphase && quiet + 1 >= nphase) {
        capped = false;
        break;
      }
    }
  }

  // ---- activations: border cells this visit improved ---------------------------
  // A changed border cell wakes the neighbour across that side only if the
  // candidate it offers beats the neighbour's label there as loaded (which
  // only decreases since); the planar edges are symmetric, so a cell whose
  // entry from that side is closed cannot leave towards it either.
  uint32_t fl = 0;
  if (row_in) {
    const uint32_t base = (rl * TH + ry) * RS;
    const LT v0 = lab[base], v1 = lab[base + TW - 1];
    if (has_l && v0 < b0[rl * TH + ry] && step_lo(sx[0]) != kMax && WL::offer_lo(v0, p_lr) < hl) fl |= T_LEFT;
    if (has_r && v1 < b0[LM * TH + rl * TH + ry] && step_hi(sx[TW - 1]) != kMax && WL::offer_hi(v1, p_lr) < hr)
      fl |= T_RIGHT;
  }
  if (col_in) {
    const uint32_t base = (cl * TH) * RS + cx;
    const LT v0 = lab[base], v1 = lab[base + (TH - 1) * RS];
    if (has_t && v0 < b0[2 * LM * TH + cl * TW + cx] && step_lo(sy[0]) != kMax && WL::offer_lo(v0, p_tb) < ht)
      fl |= T_TOP;
    if (has_b && v1 < b0[2 * LM * TH + LM * TW + cl * TW + cx] && step_hi(sy[TH - 1]) != kMax &&
        WL::offer_hi(v1, p_tb) < hb)
      fl |= T_BOTTOM;
  }
  // ---- write back: every label of the tile, then the border-column copies ------
  // (device-scope stores: written through as they are issued, not left
  // dirty in the XCD's L2 for the end of the launch to write back)
  uint32_t vmax = 0;  // packed: largest label stored with a real dist
  {
    LT* lp = (LT*)A.labels + blockbase + tid;
#pragma unroll
    for (uint32_t j = 0; j < NS; ++j) {
      const uint32_t l = j / NV, y = RPJ * (j % NV);
      if (l < L) {
        const LT v = lab[sbase + (l * TH + y) * RS];
        __hip_atomic_store(lp + NT * j, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (PK) vmax = max(vmax, (uint32_t)v < kPackedUnreached ? (uint32_t)v : 0u);
      }
    }
  }
  if (rrow) {
    const uint32_t base = (rl * TH + ry) * RS;
    LT* bp = (LT*)A.labels + blockbase + ncell + rl * TH + ry;
    __hip_atomic_store(bp, lab[base], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(bp + L * TH, lab[base + TW - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (PK) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) vmax = max(vmax, (uint32_t)__shfl_xor((int)vmax, o));
  }
  uint32_t sides = (__ballot(fl & T_LEFT) ? T_LEFT : 0u) | (__ballot(fl & T_RIGHT) ? T_RIGHT : 0u) |
                   (__ballot(fl & T_TOP) ? T_TOP : 0u) | (__ballot(fl & T_BOTTOM) ? T_BOTTOM : 0u);
  if constexpr (NW > 1) {
    // the waves' side flags and maxima, combined by every wave
    volatile uint32_t* xr = xw + 4;
    if (lane == 0) {
      xr[tid >> 6] = sides;
      xr[NW + (tid >> 6)] = vmax;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
      sides |= xr[w];
      vmax = max(vmax, (uint32_t)xr[NW + w]);
    }
  }
  // the tile's largest stored label: its last visit's value is what the
  // host's certificate reads (labels only decrease)
  if (PK && tid == 0) A.item_max[item] = vmax;
  TileResult res;
  res.flags = sides | (any ? T_ANY : 0u) | (capped ? T_CAPPED : 0u);
  res.tx = F.tx;
  res.kl = res.kr = res.kt = res.kb = res.ka = kMax;  // keys: the wave form runs without delta stepping
  if (tid == 0) {
    if (!touched && !QU)
      __hip_atomic_fetch_or(&A.state[item], S_TOUCHED | (round << 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd(shard_of(A, SH_VISITS), 1u);
    atomicAdd(shard_of(A, SH_ROUNDS), rounds);
  }
  return res;
}

// Waves per SIMD the wave form is compiled for: the packed form fits five
// (at most 96 VGPRs), the wide form's u64 lines four.
template <int TW, int LM, bool PK>
struct WaveMin {
  static constexpr int waves = PK ? 5 : 4;
};

// HO: the launch has boards with a halo-owner plane; EX: fields with own
// pads, via-exempt pads or overrides (separate kernels, so the others do not
// carry that code's registers).
template <int TW, int TH, int LM, bool PK, bool HO, bool EX, int AK>
__device__ __forceinline__ TileResult wave_visit(const RelaxArgs& A, uint32_t item, uint32_t f, bool touched,
                                                 const FieldDesc& F, const BoardView& B, char* smem) {
  if (HO && B.halo) return wave_tile<TW, TH, LM, AK, PK, HO, EX>(A, item, f, touched, F, B, smem, A.round);
  return wave_tile<TW, TH, LM, AK, PK, false, EX>(A, item, f, touched, F, B, smem, A.round);
}

// SP: every board of the launch carries step price planes (lean launches).
template <int TW, int TH, int LM, bool PK, bool HO, bool EX, bool SP = false>
__global__ void __launch_bounds__(64 * wave_count(LM, TH), (WaveMin<TW, LM, PK>::waves)) relax_wave(RelaxArgs A) {
  extern __shared__ __align__(16) char smem[];
  const uint32_t item = round_item(A);
  const uint32_t grp = A.group_active ? const_load(A.item_group + item) : 0u;
  const uint32_t f = const_load(A.item_field + item);
  const bool touched = (const_load(A.state + item) & S_TOUCHED) != 0;
  const FieldDesc F = const_load(A.fields + f);
  const BoardView B = const_load(A.boards + F.board);
  TileResult res;
  if constexpr (SP)
    res = wave_tile<TW, TH, LM, ADD_NONE, PK, false, false, false, true>(A, item, f, touched, F, B, smem, A.round);
  else if (B.add16)
    res = wave_visit<TW, TH, LM, PK, HO, EX, ADD_U16>(A, item, f, touched, F, B, smem);
  else if (B.add)
    res = wave_visit<TW, TH, LM, PK, HO, EX, ADD_U32>(A, item, f, touched, F, B, smem);
  else
    res = wave_visit<TW, TH, LM, PK, HO, EX, ADD_NONE>(A, item, f, touched, F, B, smem);
  if (threadIdx.x >= 64) return;  // the first wave finishes the visit
  round_tail(A, item, grp, res, false, 0u);
}
