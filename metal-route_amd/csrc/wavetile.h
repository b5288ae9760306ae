// Wave form of the tile visit: one wavefront per (field, tile) item.
// Included by engine.hip after relax_round (it uses RelaxArgs, TileResult,
// round_tail and the device helpers defined there).
//
// The workgroup form (process_tile) runs a tile on four wavefronts; in its
// row and column phases a 16 x 16 x 4 tile has 64 lines, so one wavefront
// sweeps while three wait at the barrier, and the tile's 25.8 KB of LDS
// (labels with a halo ring plus six step planes) lets six visits share a CU.
// Here one wavefront owns the tile:
//   * labels live in an LDS buffer [layer][y][x] (row stride TW + 1, odd, so
//     a lane per row and a lane per column both hit distinct banks); each
//     phase reads the lines it sweeps into registers in its own layout --
//     a lane per row (layer, y), a lane per column (layer, x), or TW*TH/64
//     via columns per lane -- sweeps them forward and backward and writes a
//     changed line back.  With L*TH = L*TW = 64 every lane sweeps in every
//     phase.
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
// Scope: lean fields (no halo-owner or legal planes, overrides, own pads or
// via-exempt pads) on boards without a weight plane whose step prices fit the
// narrow (15-bit) form, in the round form without delta stepping.  A forced
// endpoint on an obstacle makes a field wide, so every passable cell here is
// a cell that is neither a trace node nor an obstacle and every open step
// price is e + a.  The host picks this kernel only for such launches.

template <int TW, int TH, int LM>
struct WaveTile {
  static_assert(LM * TH <= 64 && LM * TW <= 64, "a lane per row and a lane per column");
  static_assert((TW * TH) % 64 == 0 && 64 % TW == 0, "whole via columns per lane; a fixed x per lane in block order");
  static_assert(TH == TW, "the column layout reads rows cl * TH + y of the 64-row buffer");
  static constexpr uint32_t RS = TW + 1;         // label buffer row stride
  static constexpr uint32_t PT = TW * TH;        // cells of one layer of the tile
  static constexpr uint32_t NV = PT / 64;        // via columns per lane
  static constexpr uint32_t NS = LM * PT / 64;   // cells per lane in block order
  // label slots (u64): 64 rows whatever the layer bound, so the row layout's
  // lane index is a row of the buffer for every lane
  static constexpr uint32_t LAB = 64 * RS;
  // border labels as loaded: left [LM*TH], right [LM*TH], top [LM*TW], bottom [LM*TW]
  static constexpr uint32_t BORDER = 2 * LM * TH + 2 * LM * TW;
  static constexpr size_t lds = (size_t)(LAB + BORDER) * 8;
};

__host__ __device__ inline size_t wave_lds_bytes(uint32_t tw, uint32_t th, uint32_t lm) {
  return (64ull * (tw + 1) + 2ull * lm * th + 2ull * lm * tw) * 8;
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

// Narrow step price of entering a cell with board mask bm and additive
// price a through the edge whose open bit is `edge`, over gap price e; the
// extra condition `ok` closes it too (layer bounds).
__device__ __forceinline__ uint32_t wave_step(uint8_t bm, uint8_t edge, bool ok, uint32_t e, uint32_t a) {
  const bool open = ok && !(bm & (M_TRACE | M_OBST)) && (bm & edge);
  return open ? e + a : kMax;
}

// The plane elements of four consecutive cells (element alignment only).
__device__ __forceinline__ void load_row_planes(const BoardView& B, uint32_t g0, uint8_t* bm, uint32_t* a, int k0) {
  const u8x4 b4 = load4(B.bmask + g0);
  u32x4 a4 = u32x4{0, 0, 0, 0};
  if (B.add16) {
    const u16x4 h4 = load4(B.add16 + g0);
    a4 = u32x4{h4[0], h4[1], h4[2], h4[3]};
  } else if (B.add) {
    a4 = load4(B.add + g0);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    bm[k0 + k] = b4[k];
    a[k0 + k] = a4[k];
  }
}

template <int TW, int TH, int LM>
__device__ TileResult wave_tile(const RelaxArgs& A, uint32_t item, char* smem, uint32_t round) {
  using WT = WaveTile<TW, TH, LM>;
  constexpr uint32_t RS = WT::RS, PT = WT::PT, NV = WT::NV, NS = WT::NS, LAB = WT::LAB;
  uint64_t* lab = (uint64_t*)smem;
  uint32_t* lab32 = (uint32_t*)smem;  // step staging before the labels go in: y prices [LAB], via prices [LAB]
  uint64_t* b0 = lab + LAB;
  const uint32_t lane = threadIdx.x;

  const uint32_t f = const_load(A.item_field + item);
  const bool touched = (const_load(A.state + item) & S_TOUCHED) != 0;
  const FieldDesc F = const_load(A.fields + f);
  const BoardView B = const_load(A.boards + F.board);
  const uint32_t t = item - F.item0;
  const uint32_t txi = t % F.tx, tyi = t / F.tx;
  const uint32_t L = B.layers;
  const uint32_t ncell = L * PT;
  const uint32_t wx0 = txi * TW, wy0 = tyi * TH;
  const uint64_t TB = tile_block_words(L, TW, TH);
  const uint64_t blockbase = F.lab_off + (uint64_t)t * TB;
  const bool has_l = txi > 0, has_r = txi + 1 < F.tx, has_t = tyi > 0, has_b = tyi + 1 < F.ty;

  // neighbour states (lanes 1..4): whether their border labels exist
  uint32_t stv = 0;
  if (lane >= 1 && lane < 5) {
    uint32_t nb = item;
    bool ok = true;
    if (lane == 1) ok = has_l, nb = item - 1;
    if (lane == 2) ok = has_r, nb = item + 1;
    if (lane == 3) ok = has_t, nb = item - F.tx;
    if (lane == 4) ok = has_b, nb = item + F.tx;
    stv = ok ? __hip_atomic_load(&A.state[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
  }

  // ---- loads, all issued before the first use ---------------------------------
  // The tile's labels in block order: lane holds cells lane + 64 j, i.e.
  // column x = lane % TW of row (64 / TW) * (j % NV) + lane / TW of layer
  // j / NV -- the same x for every j, so each cell's LDS slot is a constant
  // offset from one per-lane base (`sbase`, also the via layout's base).
  constexpr uint32_t RPJ = 64 / TW;  // rows per 64 cells
  const uint32_t sx_ = lane % TW, sy_ = lane / TW;
  const uint32_t sbase = sy_ * RS + sx_;
  uint64_t sl[NS];
#pragma unroll
  for (uint32_t j = 0; j < NS; ++j) sl[j] = kUnreached;
  if (touched) {
    const auto* lp = gbl(A.labels) + blockbase + lane;
#pragma unroll
    for (uint32_t j = 0; j < NS; ++j)
      if (j / NV < L) sl[j] = lp[64u * j];
  }
  // row layout: lane (rl, ry) holds row ry of layer rl
  const uint32_t rl = lane / TH, ry = lane - rl * TH;
  const bool rrow = rl < L;
  const bool row_in = rrow && wy0 + ry < F.wh;
  const uint32_t rgy = F.y0 + min(wy0 + ry, F.wh - 1);
  const uint32_t rowg = ((rrow ? rl : 0u) * B.h + rgy) * B.w + F.x0;  // board cell of window column 0
  uint8_t bm[TW];
  uint32_t av[TW];
  if (wx0 + TW <= F.ww) {
    const uint32_t g0 = rowg + wx0;
#pragma unroll
    for (int q = 0; q < TW / 4; ++q) load_row_planes(B, g0 + 4 * q, bm, av, 4 * q);
  } else {
#pragma unroll
    for (int x = 0; x < TW; ++x) {
      const uint32_t g = rowg + min(wx0 + x, F.ww - 1);
      bm[x] = gbl(B.bmask)[g];
      av[x] = add_at(B, g);
    }
  }
  // column layout: lane (cl, cx) holds column cx of layer cl
  const uint32_t cl = lane / TW, cx = lane - cl * TW;
  const bool ccol = cl < L;
  const bool col_in = ccol && wx0 + cx < F.ww;
  // Halo labels (kept only if the neighbour's first visit ran in an earlier
  // round) and the board cells of the halo (the price of stepping from the
  // tile into them).  Every load is issued: where a neighbour or the cell
  // does not exist the address falls back to one inside this tile, and the
  // value is dropped below.
  const bool el = row_in && has_l, er = row_in && has_r, et = col_in && has_t, eb = col_in && has_b;
  const uint64_t own = blockbase + lane % TW;
  const uint64_t hla = el ? blockbase - TB + ncell + L * TH + rl * TH + ry : own;
  const uint64_t hra = er ? blockbase + TB + ncell + rl * TH + ry : own;
  const uint64_t hta = et ? blockbase - (uint64_t)F.tx * TB + cl * PT + (TH - 1) * TW + cx : own;
  const uint64_t hba = eb ? blockbase + (uint64_t)F.tx * TB + cl * PT + cx : own;
  uint64_t hl = __hip_atomic_load(&A.labels[hla], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t hr = __hip_atomic_load(&A.labels[hra], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t ht = __hip_atomic_load(&A.labels[hta], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t hb = __hip_atomic_load(&A.labels[hba], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint8_t hbm[4];
  uint32_t ha[4];
  {
    const uint32_t cgx = F.x0 + wx0 + cx;
    const uint32_t g[4] = {el ? rowg + wx0 - 1 : rowg, er ? rowg + wx0 + TW : rowg,
                           et ? (cl * B.h + F.y0 + wy0 - 1) * B.w + cgx : rowg,
                           eb ? (cl * B.h + F.y0 + wy0 + TH) * B.w + cgx : rowg};
    const bool e[4] = {el, er, et, eb};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint8_t m = gbl(B.bmask)[g[k]];
      const uint32_t a = add_at(B, g[k]);
      hbm[k] = e[k] ? m : (uint8_t)M_TRACE;
      ha[k] = a;
    }
  }
  // gap prices: x gaps uniform (lane i <= TW holds the price of entering
  // tile column i from column i-1), y gaps and via prices per lane
  uint32_t xgv = 0;
  {
    const uint32_t gx = F.x0 + wx0 + lane;
    if (lane <= TW && gx >= 1 && gx < B.w) xgv = gbl(B.xe)[gx - 1];
  }
  const uint32_t gyy = F.y0 + wy0 + ry;
  const uint32_t yg0 = (gyy >= 1 && gyy < B.h) ? gbl(B.ye)[gyy - 1] : 0u;
  const uint32_t yg1 = (gyy + 1 < B.h) ? gbl(B.ye)[gyy] : 0u;
  const uint32_t vlo = (rrow && rl > 0) ? gbl(B.ve)[rl - 1] : 0u;
  const uint32_t vhi = (rrow && rl + 1 < L) ? gbl(B.ve)[rl] : 0u;
  const uint32_t gyt = F.y0 + wy0, gyb = F.y0 + wy0 + TH;
  const uint32_t ytop = (gyt >= 1 && gyt < B.h) ? gbl(B.ye)[gyt - 1] : 0u;
  const uint32_t ybot = (gyb >= 1 && gyb < B.h) ? gbl(B.ye)[gyb - 1] : 0u;

  auto earlier = [&](uint32_t st) { return (st & S_TOUCHED) && (st >> 1) < round; };
  const uint32_t nbv = (earlier(__builtin_amdgcn_readlane(stv, 1)) ? 1u : 0u) |
                       (earlier(__builtin_amdgcn_readlane(stv, 2)) ? 2u : 0u) |
                       (earlier(__builtin_amdgcn_readlane(stv, 3)) ? 4u : 0u) |
                       (earlier(__builtin_amdgcn_readlane(stv, 4)) ? 8u : 0u);
  hl = (nbv & 1u) ? lab_in(hl) : kUnreached;
  hr = (nbv & 2u) ? lab_in(hr) : kUnreached;
  ht = (nbv & 4u) ? lab_in(ht) : kUnreached;
  hb = (nbv & 8u) ? lab_in(hb) : kUnreached;
  // prices of stepping out of the tile into the halo cells: the cell's edge
  // facing the tile must be open and the cell passable (narrow: at most 0x7FFE)
  const uint32_t xleft = __builtin_amdgcn_readlane(xgv, 0), xright = __builtin_amdgcn_readlane(xgv, TW);
  const uint32_t p_lr = pack_steps(wave_step(hbm[0], M_PX, true, xleft, ha[0]), wave_step(hbm[1], M_NX, true, xright, ha[1]));
  const uint32_t p_tb = pack_steps(wave_step(hbm[2], M_PY, true, ytop, ha[2]), wave_step(hbm[3], M_NY, true, ybot, ha[3]));

  // ---- step prices of every cell (row layout) ---------------------------------
  uint32_t sx[TW];  // row layout: from x-1 | from x+1
  {
    uint32_t xg[TW + 1];
#pragma unroll
    for (int i = 0; i <= TW; ++i) xg[i] = __builtin_amdgcn_readlane(xgv, i);
#pragma unroll
    for (int x = 0; x < TW; ++x) {
      const uint8_t m = (row_in && wx0 + x < F.ww) ? bm[x] : (uint8_t)M_TRACE;  // outside the window: closed
      const uint32_t a = av[x];
      sx[x] = pack_steps(wave_step(m, M_NX, true, xg[x], a), wave_step(m, M_PX, true, xg[x + 1], a));
      // staged for the column and via layouts; lanes without a row
      // (layers beyond the board's) stage closed prices in rows no lane reads
      const uint32_t at = lane * RS + x;  // lane = rl * TH + ry
      lab32[at] = pack_steps(wave_step(m, M_NY, true, yg0, a), wave_step(m, M_PY, true, yg1, a));
      lab32[LAB + at] = pack_steps(wave_step(m, M_DN, rl > 0, vlo, a), wave_step(m, M_UP, rl + 1 < L, vhi, a));
    }
  }
  __syncthreads();
  uint32_t sy[TH];        // column layout: from y-1 | from y+1
  uint32_t sv[NV][LM];    // via layout: from layer-1 | from layer+1
#pragma unroll
  for (int y = 0; y < TH; ++y) sy[y] = ccol ? lab32[(cl * TH + y) * RS + cx] : kMax;
#pragma unroll
  for (uint32_t k = 0; k < NV; ++k) {
#pragma unroll
    for (int l = 0; l < LM; ++l) sv[k][l] = (uint32_t)l < L ? lab32[LAB + sbase + (l * TH + RPJ * k) * RS] : kMax;
  }
  __syncthreads();
  // ---- labels into the buffer -------------------------------------------------
  const bool sx_in = wx0 + sx_ < F.ww;
#pragma unroll
  for (uint32_t j = 0; j < NS; ++j) {
    const uint32_t l = j / NV, y = RPJ * (j % NV);  // + sy_
    const bool in = sx_in && wy0 + y + sy_ < F.wh;
    lab[sbase + (l * TH + y) * RS] = in ? sl[j] : kUnreached;
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
    for (uint32_t k = lane; k < F.n_src; k += 64) {
      const uint32_t g = A.pool[F.src_off + k];
      const uint32_t l = g / plane, r = g - l * plane, gy = r / B.w, gx = r - gy * B.w;
      if (l >= L || gx < F.x0 + wx0 || gy < F.y0 + wy0) continue;
      const uint32_t x = gx - F.x0 - wx0, y = gy - F.y0 - wy0;
      if (x >= TW || y >= TH || wx0 + x >= F.ww || wy0 + y >= F.wh) continue;
      if (!(gbl(B.bmask)[g] & (M_TRACE | M_OBST))) {
        lab[(l * TH + y) * RS + x] = 0ull;
        put = true;
      }
    }
    seeded = __ballot(put) != 0;
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
      if (rrow) {
        const uint32_t base = (rl * TH + ry) * RS;
        uint64_t v[TW];
#pragma unroll
        for (int x = 0; x < TW; ++x) v[x] = lab[base + x];
        uint64_t p = hl;
#pragma unroll
        for (int x = 0; x < TW; ++x) {
          ch |= relax_w(p, step_lo(sx[x]), v[x]);
          p = v[x];
        }
        p = hr;
#pragma unroll
        for (int x = TW - 1; x >= 0; --x) {
          ch |= relax_w(p, step_hi(sx[x]), v[x]);
          p = v[x];
        }
        if (ch) {
#pragma unroll
          for (int x = 0; x < TW; ++x) lab[base + x] = v[x];
        }
      }
      const bool c = __ballot(ch) != 0;
      __syncthreads();
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
      if (ccol) {
        const uint32_t base = (cl * TH) * RS + cx;
        uint64_t v[TH];
#pragma unroll
        for (int y = 0; y < TH; ++y) v[y] = lab[base + y * RS];
        uint64_t p = ht;
#pragma unroll
        for (int y = 0; y < TH; ++y) {
          ch |= relax_w(p, step_lo(sy[y]), v[y]);
          p = v[y];
        }
        p = hb;
#pragma unroll
        for (int y = TH - 1; y >= 0; --y) {
          ch |= relax_w(p, step_hi(sy[y]), v[y]);
          p = v[y];
        }
        if (ch) {
#pragma unroll
          for (int y = 0; y < TH; ++y) lab[base + y * RS] = v[y];
        }
      }
      const bool c = __ballot(ch) != 0;
      __syncthreads();
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
      for (uint32_t kk = 0; kk < NV; ++kk) {
        const uint32_t base = sbase + RPJ * kk * RS;
        uint64_t v[LM];
#pragma unroll
        for (int l = 0; l < LM; ++l) v[l] = (uint32_t)l < L ? lab[l * TH * RS + base] : kUnreached;
        bool cc = false;
#pragma unroll
        for (int l = 1; l < LM; ++l) cc |= relax_w(v[l - 1], step_lo(sv[kk][l]), v[l]);
#pragma unroll
        for (int l = LM - 2; l >= 0; --l) cc |= relax_w(v[l + 1], step_hi(sv[kk][l]), v[l]);
        if (cc) {
#pragma unroll
          for (int l = 0; l < LM; ++l)
            if ((uint32_t)l < L) lab[l * TH * RS + base] = v[l];
        }
        ch = ch || cc;
      }
      const bool c = __ballot(ch) != 0;
      __syncthreads();
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
    const uint64_t v0 = lab[base], v1 = lab[base + TW - 1];
    if (has_l && v0 < b0[rl * TH + ry] && step_lo(sx[0]) != kMax && offer(v0, step_lo(p_lr)) < hl) fl |= T_LEFT;
    if (has_r && v1 < b0[LM * TH + rl * TH + ry] && step_hi(sx[TW - 1]) != kMax && offer(v1, step_hi(p_lr)) < hr)
      fl |= T_RIGHT;
  }
  if (col_in) {
    const uint32_t base = (cl * TH) * RS + cx;
    const uint64_t v0 = lab[base], v1 = lab[base + (TH - 1) * RS];
    if (has_t && v0 < b0[2 * LM * TH + cl * TW + cx] && step_lo(sy[0]) != kMax && offer(v0, step_lo(p_tb)) < ht)
      fl |= T_TOP;
    if (has_b && v1 < b0[2 * LM * TH + LM * TW + cl * TW + cx] && step_hi(sy[TH - 1]) != kMax &&
        offer(v1, step_hi(p_tb)) < hb)
      fl |= T_BOTTOM;
  }
  // ---- write back: every label of the tile, then the border-column copies ------
  {
    uint64_t* lp = A.labels + blockbase + lane;
#pragma unroll
    for (uint32_t j = 0; j < NS; ++j) {
      const uint32_t l = j / NV, y = RPJ * (j % NV);
      if (l < L) __hip_atomic_store(lp + 64u * j, lab[sbase + (l * TH + y) * RS], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (rrow) {
    const uint32_t base = (rl * TH + ry) * RS;
    __hip_atomic_store(&A.labels[blockbase + ncell + rl * TH + ry], lab[base], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&A.labels[blockbase + ncell + L * TH + rl * TH + ry], lab[base + TW - 1], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  TileResult res;
  res.flags = (__ballot(fl & T_LEFT) ? T_LEFT : 0u) | (__ballot(fl & T_RIGHT) ? T_RIGHT : 0u) |
              (__ballot(fl & T_TOP) ? T_TOP : 0u) | (__ballot(fl & T_BOTTOM) ? T_BOTTOM : 0u) |
              (any ? T_ANY : 0u) | (capped ? T_CAPPED : 0u);
  res.tx = F.tx;
  res.kl = res.kr = res.kt = res.kb = res.ka = kMax;  // keys: the wave form runs without delta stepping
  if (lane == 0) {
    if (!touched)
      __hip_atomic_fetch_or(&A.state[item], S_TOUCHED | (round << 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd(&A.counters[(C_VISITS) * kCtrStride], 1u);
    atomicAdd(&A.counters[(C_ROUNDS) * kCtrStride], rounds);
  }
  return res;
}

// Waves per SIMD the wave form is compiled for.
template <int TW, int LM>
struct WaveMin {
  static constexpr int waves = 4;
};

template <int TW, int TH, int LM>
__global__ void __launch_bounds__(64, (WaveMin<TW, LM>::waves)) relax_wave(RelaxArgs A) {
  extern __shared__ __align__(16) char smem[];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (A.clear_ctr != A.next_ctr)
      __hip_atomic_store(&A.counters[A.clear_ctr * kCtrStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const uint32_t item = const_load(A.list_cur + blockIdx.x);
  const uint32_t grp = A.group_active ? const_load(A.item_group + item) : 0u;
  const TileResult res = wave_tile<TW, TH, LM>(A, item, smem, A.round);
  round_tail(A, item, grp, res, false, 0u);
}
