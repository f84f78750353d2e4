// Tile operators of the MI355X summed-area-table kernels.
//
// A tile is TH rows x 256 columns of the input. A workgroup of NW waves loads
// it with one coalesced 4-element access per lane and row (wave w holds rows
// [w*RPT, (w+1)*RPT), lane l holds columns [4l, 4l+4)), scans it in registers,
// and joins its neighbours through decoupled look-back on tagged carry words.
// For the 2-D SAT the value at tile-local (i, j) of tile (I, J) is
//
//   S = L(i,j) + sum_{j'<=j} Cab(j') + sum_{i'<=i} Rab(i') + Corner
//
// with L the tile-local SAT, Cab(j) the column sums above the tile, Rab(i)
// the row sums left of the tile (rows of this tile row only) and Corner the
// sum of everything above and to the left. The three carries come from three
// independent look-backs, so none waits for another:
//
//   vertical words   per tile column: the tile's column sums (AGG), the column
//                    sums down to and including the tile (INC); plus one
//                    corner slot: the tile's Sum_i Rab(i) (AGG), and the corner
//                    of the tile below (INC);
//   horizontal words per tile row: the tile's row sums (AGG), the row sums up
//                    to and including the tile (INC).
//
// Workgroups map to tiles through tile_of_block (below), under which the
// workgroups each XCD starts (in increasing index) take tiles in row-major
// order. The first unfinished tile in row-major order therefore has every
// earlier tile of its XCD finished, so it is resident, and its predecessors
// (above and to the left, earlier in row-major order) are finished: the
// launch always progresses.
#pragma once

#include "sat_device.hpp"

namespace satdev {

enum : int { MODE_ROW = 1, MODE_COL = 2, MODE_SAT = 3 };

// predecessors whose carry words one look-back round reads at once (a lane
// that follows two column chains reads LB_WINDOW_V per chain). Carry reads
// cost HBM bandwidth like pixel data, and a wide window mostly reads words
// past the inclusive one, so the windows are short.
constexpr int LB_WINDOW = 2;
constexpr int LB_WINDOW_V = 1;
// Horizontal window per payload width: an 8-byte payload is two words, and 4
// predecessors per round trip measured 8-12 % faster than 2 there (f64
// 4096^2 and 8192^2; 8 was slower again), while 4-byte payloads lost with 4.
template <typename A> constexpr int lb_window_h() { return sizeof(A) == 8 ? 4 : LB_WINDOW; }

// Second carry level of the vertical chains of tall images: the tiles of a
// tile column are grouped in runs of `len` tile rows, and the last tile of a
// run publishes the run's column sums (AGG) and its inclusive column sums
// (INC) in `carry` (one entry per run, tile column and column, planes after
// each other). A tile's vertical look-back walks its own run's tiles (one
// window) and then the earlier runs' words, so a chain of n tiles takes about
// len + n / len steps instead of up to n. len == 0: one level. Only 4-byte
// payloads take it: the extra code made the 8-byte kernels spill registers
// (f64 8192^2 246 -> 254 us with no run words in use).
struct VGroups {
  uint64_t* carry = nullptr;
  uint32_t len = 0;
};

template <bool C, typename T, typename F> struct cond { using type = T; };
template <typename T, typename F> struct cond<false, T, F> { using type = F; };

// ---------------------------------------------------------------------------
// Register tile operator. MODE_SAT: 2-D inclusive prefix sum (one pass).
// MODE_ROW: row prefix sums (IncrementScan::ScanX semantics, or with TOUT the
// transposed row scan of BlockScan::blockScan). MODE_COL: column prefix sums
// (IncrementScan::ScanY semantics). REV mirrors both axes (suffix sums), the
// adjoint of the forward SAT. ACC is the arithmetic type (default: that of
// TDst); a wider one accumulates and carries in it and rounds once on store.
// ---------------------------------------------------------------------------
template <typename TSrc, typename TDst, int MODE, bool REV, bool TOUT, int SCAN, int NW, int RPT,
          typename ACC = arith_t<TDst>>
struct TileOp {
  using Src = TSrc;
  using Dst = TDst;
  using A = ACC;
  static constexpr int TW = 256;
  static constexpr int TH = NW * RPT;
  static constexpr int NT = 64 * NW;
  static constexpr bool DO_ROW = (MODE & MODE_ROW) != 0;
  static constexpr bool DO_COL = (MODE & MODE_COL) != 0;
  static constexpr bool CORNER = DO_ROW && DO_COL;
  // vertical look-backs run on the first NVW waves (all waves for a column
  // pass), the horizontal and corner look-backs on the last wave
  static constexpr int NVW = !DO_COL ? 1 : (DO_ROW ? NW / 2 : NW);
  static constexpr int CPV = TW / (64 * NVW);  // columns per vertical lane
  // two adjacent columns per vertical lane, their words published and read as
  // one 16-byte pair, so both chains take one round trip per predecessor
  // (f32 8192^2 107.3 -> 106.2 us, u32 -0.6 %, u8 -> u32 8192^2 -1.7 %; the
  // columns of a lane were 128 apart, read one chain after the other)
  static constexpr bool VPAIR = CPV == 2 && Carry<A>::WORDS == 1;
  __device__ static __forceinline__ int vcol(int q, int tid) { return VPAIR ? 2 * tid + q : q * 64 * NVW + tid; }
  static constexpr int HWAVE = NW - 1;
  // corner look-back: its own wave when there is one to spare
  static constexpr int CWAVE = (NW >= 3 && NVW <= NW - 2) ? NW - 2 : HWAVE;
  // vertical entries per tile: the tile's columns, then its corner entry padded
  // to a whole 64-byte line, so every tile's block of column words starts on a
  // line and its wave-wide stores write whole lines
  static constexpr int64_t VS = TW + (CORNER ? 8 : 0);
  static constexpr int64_t VWORDS = DO_COL ? VS * Carry<A>::WORDS : 0;  // per tile
  static constexpr int64_t HWORDS = DO_ROW ? (int64_t)TH * Carry<A>::WORDS : 0;  // per tile
  // run words (VGroups) for tiles whose vertical lanes follow one column each
  static constexpr bool GROUPS = DO_COL && CPV == 1 && Carry<A>::WORDS == 1;
  static constexpr int64_t GWORDS = GROUPS ? (int64_t)TW * Carry<A>::WORDS : 0;  // per run and tile column
  static_assert(!(TOUT && (REV || DO_COL)), "transposed store is a row-scan pass only");
  // the horizontal look-back wave follows KH rows per lane: rows l, l+64, ...
  // (lanes of the last chain past the tile's rows take no part)
  static constexpr int KH = (TH + 63) / 64;
  static_assert(!TOUT || TH <= 64, "one lane per tile row in the transposed store");
  static_assert(!CORNER || NW >= 2, "the 2-D pass splits look-backs over waves");
  static_assert(TW % (64 * NVW) == 0, "vertical lanes");

  struct SmemCarry {
    // per-wave column sums; for 8-byte values also the store stage (16-byte
    // aligned for its 16-byte LDS accesses)
    alignas(sizeof(A) == 8 ? 16 : alignof(A)) A ct[DO_COL ? NW : 1][TW];
    A cab[DO_COL ? TW : 1];     // column sums above the tile
    A rs[DO_ROW ? TH : 1];      // row sums of the tile
    A rc[DO_ROW ? TH : 1];      // row carry of each tile row
    A corner;
    A ksum;                     // sum of the tile's row carries (corner-slot aggregate)
  };
  struct SmemStage {
    A st[TH * TW];  // XOR-swizzled, see stage_put4 / stage_get
    A rs[TH];
    A rc[TH];
    A corner;
  };
  using Smem = typename cond<TOUT, SmemStage, SmemCarry>::type;

  // Physical placement of logical tile t (row-major over the tile grid).
  struct Coords {
    uint32_t I, J;        // logical tile row / column
    int64_t py0, px0;     // physical origin of the tile
    int64_t pcol;         // physical first column of this lane's 4-chunk
  };
  __device__ static __forceinline__ Coords coords(const ImgArgs& g, uint32_t t) {
    Coords k;
    const int lane = threadIdx.x & 63;
    k.I = t / g.nTx;
    k.J = t - k.I * g.nTx;
    const uint32_t PI = REV ? g.nTy - 1 - k.I : k.I;
    const uint32_t PJ = REV ? g.nTx - 1 - k.J : k.J;
    k.py0 = (int64_t)PI * TH;
    k.px0 = (int64_t)PJ * TW;
    k.pcol = k.px0 + 4 * (REV ? 63 - lane : lane);
    return k;
  }
  // physical row of this wave's logical tile row r
  __device__ static __forceinline__ int64_t prow_of(const Coords& k, int r) {
    const int li = (threadIdx.x >> 6) * RPT + r;
    return k.py0 + (REV ? TH - 1 - li : li);
  }

  // A tile that lies wholly inside the image, with aligned rows, is loaded
  // and stored with straight-line vector accesses, so every lane keeps all
  // RPT row accesses in flight at once; edge tiles take the bounds-checked
  // path. The test is uniform over the workgroup.
  __device__ static __forceinline__ bool interior(const ImgArgs& g, const Coords& k) {
    return k.py0 + TH <= (int64_t)g.H && k.px0 + TW <= (int64_t)g.W;
  }

  __device__ static __forceinline__ void run(const ImgArgs& g, uint32_t t, uint32_t ep4, Smem& sm,
                                             const VGroups& vg = VGroups{}) {
    const Coords k = coords(g, t);
    const TSrc* __restrict__ in = (const TSrc*)g.in;
    // ---- load (each wave: RPT coalesced row segments of 256 elements) ----
    A v[RPT][4];
    if (interior(g, k)) {
      typename VecG<TSrc>::type q[RPT];
      const TSrc* p0 = in + prow_of(k, 0) * g.ldi + k.pcol;
      const int64_t step = REV ? -g.ldi : g.ldi;
      // 4-byte inputs are read nontemporally: the input is streamed once, and
      // keeping it out of the caches measured 1-3 % faster (f32 8192^2,
      // 65536 x 256; neutral or slower for 1- and 8-byte inputs)
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const typename VecG<TSrc>::type* a = (const typename VecG<TSrc>::type*)(p0 + r * step);
        if constexpr (sizeof(TSrc) == 4) q[r] = __builtin_nontemporal_load(a);
        else q[r] = *a;
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const A x[4] = {cvt_in<TSrc, TDst, A>(q[r].x), cvt_in<TSrc, TDst, A>(q[r].y),
                        cvt_in<TSrc, TDst, A>(q[r].z), cvt_in<TSrc, TDst, A>(q[r].w)};
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] = REV ? x[3 - c] : x[c];
      }
    } else if (g.W >= 4) {
      // edge tile: clamped rows and chunks, one vector load per row, every
      // row's load issued before any is used
      typename VecG<TSrc>::type q[RPT];
      const int64_t base = edge_base(k.pcol, g.W);
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int64_t prow = prow_of(k, r);
        const int64_t crow = prow < (int64_t)g.H ? prow : (int64_t)g.H - 1;
        q[r] = *(const typename VecG<TSrc>::type*)(in + crow * g.ldi + base);
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        A x[4];
        edge_pick<TSrc, TDst>(q[r], k.pcol, base, g.W, prow_of(k, r) < (int64_t)g.H, x);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] = REV ? x[3 - c] : x[c];
      }
    } else {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int64_t prow = prow_of(k, r);
        A x[4];
        load4<TSrc, TDst>(in + prow * g.ldi, k.pcol, g.W, prow < (int64_t)g.H, x);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] = REV ? x[3 - c] : x[c];
      }
    }
    finish(g, t, ep4, sm, k, v, vg);
  }

  // The tile's column sums, per wave in sm.ct, summed over the waves: the
  // first NVW waves publish them (AGG; INC on the first tile row) and keep
  // their columns' sums in colsum. Starts with the barrier that makes sm.ct
  // complete.
  __device__ static __forceinline__ void publish_colsums(const ImgArgs& g, uint32_t t, uint32_t I, uint32_t ep4,
                                                         Smem& sm, A (&colsum)[CPV]) {
    const int tid = threadIdx.x, w = tid >> 6;
    __syncthreads();
    if (w < NVW) {
#pragma unroll
      for (int q = 0; q < CPV; ++q) {
        const int j = vcol(q, tid);
        A cs = sm.ct[0][j];
#pragma unroll
        for (int u = 1; u < NW; ++u) cs += sm.ct[u][j];
        colsum[q] = cs;
        if constexpr (!VPAIR) Carry<A>::put(g.vcarry, (int64_t)t * VS + j, cs, ep4 | (I == 0 ? ST_INC : ST_AGG));
      }
      if constexpr (VPAIR) {
        const uint64_t tag = (uint64_t)(ep4 | (I == 0 ? ST_INC : ST_AGG)) << 32;
        u64x2 pw;
        pw.x = tag | __builtin_bit_cast(uint32_t, colsum[0]);
        pw.y = tag | __builtin_bit_cast(uint32_t, colsum[1]);
        st_pair(g.vcarry + (int64_t)t * VS + 2 * tid, pw);
      }
    }
  }

  // The three look-backs (after the tile's column and row sums are
  // published, sm.rs filled and a barrier passed): vertical on the first NVW
  // waves into sm.cab, horizontal on wave HWAVE into sm.rc (row carries,
  // prefixed over the tile's rows for the 2-D SAT, whose total is the corner
  // slot's aggregate), the corner on wave CWAVE into sm.corner; each
  // publishes its inclusive words. Ends after a barrier.
  // Vertical look-back of a one-column lane over two carry levels (VGroups):
  // acc = column sum of every tile above. The last tile of a run publishes
  // the run's aggregate before it waits on earlier runs, and the run's
  // inclusive sums after.
  __device__ static __forceinline__ A group_look_back(const ImgArgs& g, const VGroups& vg, uint32_t t, uint32_t I,
                                                      uint32_t J, uint32_t ep4, A colsum) {
    const int tid = threadIdx.x;
    const uint32_t gi = I / vg.len, m = I - gi * vg.len;  // run index, predecessors inside the run
    const int64_t gstep = (int64_t)g.nTx * TW;
    const int64_t ge = ((int64_t)gi * g.nTx + J) * TW + tid;
    A acc = zero<A>();
    bool inc = false;
    if (m > 0)
      acc = look_back<A, 8>(g.vcarry, (int64_t)(t - g.nTx) * VS + tid, (int64_t)g.nTx * VS, m, ep4, g.fault, &inc);
    const bool last = m == vg.len - 1;
    if (!inc) {
      // every tile of the run above this one was an aggregate (run 0 always
      // ends at the inclusive first tile row)
      if (last) Carry<A>::put(vg.carry, ge, acc + colsum, ep4 | ST_AGG);
      if (gi > 0) acc += look_back<A, 4>(vg.carry, ge - gstep, gstep, gi, ep4, g.fault);
    }
    if (last) Carry<A>::put(vg.carry, ge, acc + colsum, ep4 | ST_INC);
    return acc;
  }

  __device__ static __forceinline__ void lookbacks(const ImgArgs& g, uint32_t t, uint32_t I, uint32_t J, uint32_t ep4,
                                                   Smem& sm, const A (&colsum)[CPV], const VGroups& vg = VGroups{}) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if constexpr (DO_COL) {
      if (w < NVW) {
        A acc[CPV];
#pragma unroll
        for (int q = 0; q < CPV; ++q) acc[q] = zero<A>();
        if (I > 0) {
          int64_t e[CPV];
#pragma unroll
          for (int q = 0; q < CPV; ++q) e[q] = (int64_t)(t - g.nTx) * VS + vcol(q, tid);
          bool grouped = false;
          if constexpr (GROUPS) {
            if (vg.len) {
              acc[0] = group_look_back(g, vg, t, I, J, ep4, colsum[0]);
              grouped = true;
            }
          }
          // two one-word column chains of a lane: adjacent words, one 16-byte
          // load per predecessor for both. Otherwise the chains are read one
          // after the other: two separate loads per round for both at once
          // measured 1-3 % slower on 64-row f32 and u32 tiles
          if constexpr (VPAIR) {
            look_back_pair<A>(g.vcarry, e[0], (int64_t)g.nTx * VS, I, ep4, g.fault, acc);
            const uint64_t tag = (uint64_t)(ep4 | ST_INC) << 32;
            u64x2 pw;
            pw.x = tag | __builtin_bit_cast(uint32_t, acc[0] + colsum[0]);
            pw.y = tag | __builtin_bit_cast(uint32_t, acc[1] + colsum[1]);
            st_pair(g.vcarry + (int64_t)t * VS + 2 * tid, pw);
          } else {
            if (!grouped)
              look_back_k<A, LB_WINDOW_V, CPV, false>(g.vcarry, e, (int64_t)g.nTx * VS, I, ep4, g.fault, acc);
#pragma unroll
            for (int q = 0; q < CPV; ++q)
              Carry<A>::put(g.vcarry, (int64_t)t * VS + vcol(q, tid), acc[q] + colsum[q], ep4 | ST_INC);
          }
        }
#pragma unroll
        for (int q = 0; q < CPV; ++q) sm.cab[vcol(q, tid)] = acc[q];
      }
    }
    if constexpr (DO_ROW) {
      if (w == HWAVE) {
        A rab[KH];
#pragma unroll
        for (int k = 0; k < KH; ++k) rab[k] = zero<A>();
        if (J > 0) {
          if constexpr (KH == 1) {
            if (lane < TH) {
              const int64_t e = (int64_t)t * TH + lane;
              // the wider window of 8-byte payloads only for 8-byte outputs: the
              // f32 gradient accumulated in f64 (32-row tiles) measured 1 %
              // faster with 2
              rab[0] = look_back<A, (sizeof(TDst) == 8 ? lb_window_h<A>() : LB_WINDOW)>(g.hcarry, e - TH, TH, J, ep4,
                                                                                      g.fault);
              Carry<A>::put(g.hcarry, e, rab[0] + sm.rs[lane], ep4 | ST_INC);
            }
          } else {
            int64_t e[KH];
            bool skip[KH];
#pragma unroll
            for (int k = 0; k < KH; ++k) {
              skip[k] = 64 * k + lane >= TH;
              e[k] = (int64_t)(t - 1) * TH + (skip[k] ? 0 : 64 * k + lane);
            }
            // two row chains per lane are read together (f32 16384^2 with
            // 128-row tiles: 383 -> 362 us); four (256-row packed tiles)
            // one after the other, which measured 1-4 % faster there. Tiles
            // of 1-byte inputs into 4-byte outputs read one predecessor per
            // round: u8 -> u32
            // 16384^2 273.2 -> 268.5 us and 8192^2 -1.7 % (2 for the rest:
            // f32 16384^2 +0.7 % and 256 x 65536 +2 % with 1)
            look_back_k<A, (sizeof(TSrc) == 1 && sizeof(A) == 4 ? 1 : lb_window_h<A>()), KH, (KH <= 2)>(
                g.hcarry, e, TH, J, ep4, g.fault, rab, skip);
#pragma unroll
            for (int k = 0; k < KH; ++k)
              if (!skip[k])
                Carry<A>::put(g.hcarry, (int64_t)t * TH + 64 * k + lane, rab[k] + sm.rs[64 * k + lane],
                              ep4 | ST_INC);
          }
        }
        if constexpr (CORNER) {
          // row carry = prefix of Rab over the tile's rows; its total is the
          // corner-slot aggregate the tiles below need
          A base = zero<A>();
#pragma unroll
          for (int k = 0; k < KH; ++k) {
            const A inc = wave_scan<SCAN>(rab[k], lane);
            if (64 * k + lane < TH) sm.rc[64 * k + lane] = base + inc;
            base += readlane(inc, 63);
          }
          if (lane == 0) {
            Carry<A>::put(g.vcarry, (int64_t)t * VS + TW, base, ep4 | (I == 0 ? ST_INC : ST_AGG));
            sm.ksum = base;
          }
        } else {
#pragma unroll
          for (int k = 0; k < KH; ++k)
            if (64 * k + lane < TH) sm.rc[64 * k + lane] = rab[k];
        }
      }
      if constexpr (CORNER) {
        // the corner look-back only reads the tiles above, so it runs
        // beside this tile's row look-back
        if (w == CWAVE && lane == 0) {
          A corner = zero<A>();
          if (I > 0) {
            const int64_t e = (int64_t)t * VS + TW;
            corner = look_back<A, LB_WINDOW>(g.vcarry, e - (int64_t)g.nTx * VS, (int64_t)g.nTx * VS, I, ep4,
                                               g.fault);
          }
          sm.corner = corner;
        }
      }
    }
    __syncthreads();
    if constexpr (CORNER) {
      if (I > 0 && tid == 0)
        Carry<A>::put(g.vcarry, (int64_t)t * VS + TW, sm.corner + sm.ksum, ep4 | ST_INC);
    }

  }

  // colAdd[c] for this lane's column 4*lane + c: the column sums above the
  // tile (sm.cab) and above this wave within the tile, prefixed over the
  // columns for the 2-D SAT.
  __device__ static __forceinline__ void col_add(const Smem& sm, A (&colAdd)[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 4; ++c) colAdd[c] = zero<A>();
    if constexpr (DO_COL) {
      A U[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = 4 * lane + c;
        A u = sm.cab[j];
        for (int q = 0; q < w; ++q) u += sm.ct[q][j];
        U[c] = u;
      }
      if constexpr (DO_ROW) {
        U[1] += U[0];
        U[2] += U[1];
        U[3] += U[2];
        const A inc = wave_scan<SCAN>(U[3], lane);
        const A exc = wave_shr1(inc);
#pragma unroll
        for (int c = 0; c < 4; ++c) colAdd[c] = U[c] + exc;
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) colAdd[c] = U[c];
      }
    }
  }

  // Everything after the load: scans, look-backs, store. `v` holds the tile
  // in logical order (mirrored when REV).
  __device__ static __forceinline__ void finish(const ImgArgs& g, uint32_t t, uint32_t ep4, Smem& sm,
                                                const Coords& k, A (&v)[RPT][4], const VGroups& vg = VGroups{}) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t I = k.I, J = k.J;
    const int64_t py0 = k.py0, px0 = k.px0, pcol = k.pcol;

    // ---- per-wave column sums; the tile's column sums are published first ----
    A colsum[CPV];
    if constexpr (DO_COL) {
      if constexpr (!DO_ROW) {
        // column pass: the in-lane column prefix is the result within the wave
#pragma unroll
        for (int r = 1; r < RPT; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) v[r][c] += v[r - 1][c];
#pragma unroll
        for (int c = 0; c < 4; ++c) sm.ct[w][4 * lane + c] = v[RPT - 1][c];
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          A sum = v[0][c];
#pragma unroll
          for (int r = 1; r < RPT; ++r) sum += v[r][c];
          sm.ct[w][4 * lane + c] = sum;
        }
      }
      publish_colsums(g, t, I, ep4, sm, colsum);
    }

    // ---- row prefix: in-lane over 4 columns, then across the 64 lanes; the
    //      tile's row sums are published as soon as they are known ----
    if constexpr (DO_ROW) {
      A myrs = zero<A>();  // lane r < RPT: sum of this wave's row r
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        v[r][1] += v[r][0];
        v[r][2] += v[r][1];
        v[r][3] += v[r][2];
        const A inc = wave_scan<SCAN>(v[r][3], lane);
        const A exc = wave_shr1(inc);
        const A rsum = readlane(inc, 63);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] += exc;
        if (lane == r) myrs = rsum;
      }
      if (lane < RPT) {
        // one store instruction per wave publishes its RPT row sums
        const int i = w * RPT + lane;
        Carry<A>::put(g.hcarry, (int64_t)t * TH + i, myrs, ep4 | (J == 0 ? ST_INC : ST_AGG));
        sm.rs[i] = myrs;
      }
      if constexpr (DO_COL) {
        // tile-local column prefix of the row prefixes: the wave-local SAT
#pragma unroll
        for (int r = 1; r < RPT; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) v[r][c] += v[r - 1][c];
      }
      __syncthreads();
    }

    lookbacks(g, t, I, J, ep4, sm, colsum, vg);

    // ---- per-column and per-row additions ----
    A colAdd[4];
    col_add(sm, colAdd);
    A rp[RPT];
    if constexpr (DO_ROW) {
      A corner = zero<A>();
      if constexpr (CORNER) corner = sm.corner;
#pragma unroll
      for (int r = 0; r < RPT; ++r) rp[r] = sm.rc[w * RPT + r] + corner;
    }

    // ---- store ----
    if constexpr (!TOUT) {
      TDst* __restrict__ out = (TDst*)g.out;
      const bool fast = interior(g, k);
      // 8-byte outputs of interior tiles go through the wave's row of sm.ct
      // (free once every wave is past col_add), so that each store
      // instruction writes 1 KB of whole lines (lane l: columns 2l, 2l+1, then
      // 128 + 2l, 129 + 2l) instead of 16 of every 32 bytes, and can be
      // nontemporal like the 4-byte outputs' stores: f32 -> f64 96 x 1024^2
      // 286 -> 263 us, f64 4096^2 -2 %, f64 8192^2 equal, the 512-image
      // batch with f64 outputs +0.7 %
      constexpr bool STAGE8 = sizeof(TDst) == 8 && DO_COL;
      if constexpr (STAGE8) {
        if (fast) __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int li = w * RPT + r;
        const int64_t prow = py0 + (REV ? TH - 1 - li : li);
        A y[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          A s = v[r][c];
          if constexpr (DO_COL) s += colAdd[c];
          if constexpr (DO_ROW) s += rp[r];
          y[REV ? 3 - c : c] = s;
        }
        if (STAGE8 && fast) {
          typedef TDst d2 __attribute__((ext_vector_type(2)));
          A* st = &sm.ct[w][0];
          const int qc = REV ? 63 - lane : lane;  // physical chunk of this lane's 4 columns
          *(d2*)&st[4 * qc] = d2{cvt_out<TDst>(y[0]), cvt_out<TDst>(y[1])};
          *(d2*)&st[4 * qc + 2] = d2{cvt_out<TDst>(y[2]), cvt_out<TDst>(y[3])};
          // a wave's LDS accesses complete in order; the barriers keep the
          // compiler from moving the reads above the writes, or the next
          // row's writes above these reads
          __builtin_amdgcn_wave_barrier();
          const d2 a = *(const d2*)&st[2 * lane];
          const d2 b = *(const d2*)&st[128 + 2 * lane];
          __builtin_amdgcn_wave_barrier();
          TDst* orow = out + prow * g.ldo + px0;
          __builtin_nontemporal_store(a, (d2*)(orow + 2 * lane));
          __builtin_nontemporal_store(b, (d2*)(orow + 128 + 2 * lane));
        } else if (fast) {
          typename VecG<TDst>::type q;
          q.x = cvt_out<TDst>(y[0]);
          q.y = cvt_out<TDst>(y[1]);
          q.z = cvt_out<TDst>(y[2]);
          q.w = cvt_out<TDst>(y[3]);
          TDst* o0 = out + (py0 + (REV ? TH - 1 - w * RPT : w * RPT)) * g.ldo + pcol;
          // streaming output: nontemporal stores keep it from displacing
          // the carry words and the input in the caches. An 8-byte output
          // row chunk of the row pass (no stage) is two 16-byte stores per
          // lane, each covering half of every 64-byte line: plain stores,
          // which L2 can merge, measured 3-5 % faster than nontemporal ones
          // there (on the 2-D SAT before it had the stage)
          typename VecG<TDst>::type* oq = (typename VecG<TDst>::type*)(o0 + r * (REV ? -g.ldo : g.ldo));
          if constexpr (sizeof(TDst) == 8) *oq = q;
          else __builtin_nontemporal_store(q, oq);
        } else {
          store4<TDst>(out + prow * g.ldo, pcol, g.W, prow < (int64_t)g.H, y);
        }
      }
    } else {
      // Block-register-local transpose: stage the scanned tile in LDS, then
      // write output rows (= input columns). Each lane gathers E consecutive
      // tile rows of one column and writes them as one 16-byte vector, so
      // TH/E lanes cover an output row's TH elements without gaps.
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        A y[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) y[c] = v[r][c] + rp[r];
        stage_put4<A, TW>(sm.st, w * RPT + r, 4 * lane, y);
      }
      __syncthreads();
      constexpr int E = kVecOut<TDst>;  // elements per lane store
      constexpr int LPR = TH / E;       // lanes per output row segment
      constexpr int RPI = NT / LPR;     // output rows written per pass
      static_assert(TH % E == 0 && NT % LPR == 0 && TW % RPI == 0, "transpose passes");
      TDst* __restrict__ out = (TDst*)g.out;
      const int seg = tid % LPR, rsub = tid / LPR;
      const int64_t oy0 = py0 + E * seg;
      const bool full = oy0 + E <= (int64_t)g.H;
#pragma unroll 4
      for (int xo = rsub; xo < TW; xo += RPI) {
        const int64_t orow = px0 + xo;
        A y[E];
#pragma unroll
        for (int c = 0; c < E; ++c) y[c] = stage_get<A, TW>(sm.st, E * seg + c, xo);
        if (orow < (int64_t)g.W) {
          if (full) {
            typename VecOut<TDst>::type q;
#pragma unroll
            for (int c = 0; c < E; ++c) q[c] = cvt_out<TDst>(y[c]);
            __builtin_nontemporal_store(q, (typename VecOut<TDst>::type*)(out + orow * g.ldo + oy0));
          } else {
#pragma unroll
            for (int c = 0; c < E; ++c)
              if (oy0 + c < (int64_t)g.H) out[orow * g.ldo + oy0 + c] = cvt_out<TDst>(y[c]);
          }
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// One-pass 2-D SAT of 1-byte inputs (u8, i8) into 4-byte outputs with the
// tile kept packed: a lane holds the 4 input bytes of each of its rows in one
// register until the output pass, so a wave holds RPT rows in the registers
// an unpacked tile spends on RPT / 4. With NW = 4 and RPT = 64 a tile is
// 256 x 256 pixels, four times the pixels of the unpacked 64 x 256 tile for
// the same carry words per tile (a quarter of the vertical words per pixel).
//
// Before the look-backs, the column sums (two 16-bit lanes per register:
// bytes 0 and 2, bytes 1 and 3) and the row sums (v_dot4 per lane, then a
// wave scan) come from the packed bytes and are published; the look-backs are
// TileOp's. After them every row is unpacked, scanned and stored in turn.
// Tile-local sums are exact 32-bit integers (|sum| <= 256 * 256 * 255 < 2^24,
// so they are also exact once converted to f32); carries add in the output's
// arithmetic type, so u8 -> u32 and i8 -> i32 wrap modulo 2^32 as the other
// operators do.
// ---------------------------------------------------------------------------
template <typename TSrc, typename TDst, int SCAN, int NW, int RPT>
struct PackedTileOp {
  using Base = TileOp<TSrc, TDst, MODE_SAT, false, false, SCAN, NW, RPT>;
  using Src = TSrc;
  using Dst = TDst;
  using A = typename Base::A;
  using Smem = typename Base::Smem;
  static constexpr int TW = Base::TW, TH = Base::TH, NT = Base::NT;
  static constexpr int64_t VWORDS = Base::VWORDS, HWORDS = Base::HWORDS;
  static constexpr bool GROUPS = false;
  static constexpr int64_t GWORDS = 0;
  static constexpr bool SIGNED = (TSrc)(-1) < (TSrc)0;
  static_assert(sizeof(TSrc) == 1 && sizeof(A) == 4, "1-byte inputs, 4-byte outputs");
  static_assert(RPT <= 128, "16-bit column partial sums");

  __device__ static __forceinline__ uint32_t byte_of(uint32_t p, int c) {
    if constexpr (SIGNED) return (uint32_t)((int32_t)(p << (24 - 8 * c)) >> 24);
    else return (p >> (8 * c)) & 0xffu;
  }
  __device__ static __forceinline__ A to_a(uint32_t x) {
    if constexpr (sizeof(A) == 4 && (A)0.5f != (A)0) return SIGNED ? (A)(int32_t)x : (A)x;  // f32
    else return (A)x;
  }

  // The 4 bytes of columns [col, col+4) of one row; zero outside the image.
  __device__ static __forceinline__ uint32_t load_edge(const TSrc* __restrict__ row, int64_t col, int64_t W,
                                                       bool row_ok) {
    if (!row_ok || col >= W) return 0u;
    if (W >= 4) {
      // the last whole chunk [W-4, W) for a chunk crossing the right edge,
      // shifted so that byte c is column col + c (zeros shift in past W)
      const int64_t base = col + 3 < W ? col : W - 4;
      const uint32_t raw = __builtin_bit_cast(uint32_t, *(const typename VecG<TSrc>::type*)(row + base));
      return raw >> (8 * (int)(col - base));
    }
    uint32_t p = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (col + c < W) p |= (uint32_t)(uint8_t)row[col + c] << (8 * c);
    return p;
  }

  __device__ static __forceinline__ void run(const ImgArgs& g, uint32_t t, uint32_t ep4, Smem& sm) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const typename Base::Coords k = Base::coords(g, t);
    const uint32_t I = k.I, J = k.J;
    const TSrc* __restrict__ in = (const TSrc*)g.in;
    const bool inner = Base::interior(g, k);

    // ---- load: RPT packed rows per lane ----
    uint32_t p[RPT];
    if (inner) {
      const TSrc* p0 = in + Base::prow_of(k, 0) * g.ldi + k.pcol;
#pragma unroll
      for (int r = 0; r < RPT; ++r)
        p[r] = __builtin_bit_cast(uint32_t, *(const typename VecG<TSrc>::type*)(p0 + r * g.ldi));
    } else {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int64_t prow = Base::prow_of(k, r);
        const bool ok = prow < (int64_t)g.H;
        p[r] = load_edge(in + (ok ? prow : (int64_t)g.H - 1) * g.ldi, k.pcol, g.W, ok);
      }
    }

    // ---- column sums of the wave's rows: bytes 0 / 2 and 1 / 3 summed in
    //      the two 16-bit halves of a register (i8 biased by 128 to u8) ----
    {
      const uint32_t bias = SIGNED ? 0x80808080u : 0u;
      uint32_t even = 0, odd = 0;
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const uint32_t q = p[r] ^ bias;
        even += q & 0x00ff00ffu;
        odd += (q >> 8) & 0x00ff00ffu;
      }
      const uint32_t off = SIGNED ? 128u * RPT : 0u;
      const uint32_t cs[4] = {(even & 0xffffu) - off, (odd & 0xffffu) - off, (even >> 16) - off, (odd >> 16) - off};
#pragma unroll
      for (int c = 0; c < 4; ++c) sm.ct[w][4 * lane + c] = to_a(cs[c]);
    }
    A colsum[Base::CPV];
    Base::publish_colsums(g, t, I, ep4, sm, colsum);

    // ---- row sums, published as soon as they are known ----
    {
      uint32_t mine = 0;  // lane r: sum of this wave's row r (r < 64)
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        uint32_t s4;
        if constexpr (SIGNED) s4 = (uint32_t)__builtin_amdgcn_sdot4((int)p[r], 0x01010101, 0, false);
        else s4 = __builtin_amdgcn_udot4(p[r], 0x01010101u, 0u, false);
        const uint32_t tot = readlane(wave_scan_ks(s4), 63);
        if (lane == r % 64) mine = tot;
        if (r % 64 == 63 || r == RPT - 1) {
          if (lane <= r % 64) {
            const int i = w * RPT + (r / 64) * 64 + lane;
            const A rs = to_a(mine);
            Carry<A>::put(g.hcarry, (int64_t)t * TH + i, rs, ep4 | (J == 0 ? ST_INC : ST_AGG));
            sm.rs[i] = rs;
          }
        }
      }
    }
    __syncthreads();
    Base::lookbacks(g, t, I, J, ep4, sm, colsum);
    A colAdd[4];
    Base::col_add(sm, colAdd);
    const A corner = sm.corner;

    // ---- output: unpack, scan and store row by row ----
    TDst* __restrict__ out = (TDst*)g.out;
    uint32_t colrun[4] = {0u, 0u, 0u, 0u};  // wave-local column prefix (the tile-local SAT)
    TDst* o0 = out + (k.py0 + w * RPT) * g.ldo + k.pcol;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      uint32_t x[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) x[c] = byte_of(p[r], c);
      x[1] += x[0];
      x[2] += x[1];
      x[3] += x[2];
      const uint32_t exc = wave_shr1(wave_scan_ks(x[3]));
      const A rp = sm.rc[w * RPT + r] + corner;
      A y[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        colrun[c] += x[c] + exc;
        y[c] = to_a(colrun[c]) + colAdd[c] + rp;
      }
      if (inner) {
        typename VecG<TDst>::type q;
        q.x = cvt_out<TDst>(y[0]);
        q.y = cvt_out<TDst>(y[1]);
        q.z = cvt_out<TDst>(y[2]);
        q.w = cvt_out<TDst>(y[3]);
        __builtin_nontemporal_store(q, (typename VecG<TDst>::type*)(o0 + r * g.ldo));
      } else {
        const int64_t prow = Base::prow_of(k, r);
        store4<TDst>(out + prow * g.ldo, k.pcol, g.W, prow < (int64_t)g.H, y);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// BRLT-then-serial-scan operator (SerielScan::serielScan semantics: transposed
// row scan). The tile (64 rows x NW*CW columns) is loaded row-major, moved
// through LDS so that lane = tile row and registers = a run of CW consecutive
// columns owned by the wave, scanned serially in registers, joined across the
// waves' column runs and across tiles (horizontal look-back), and moved back
// through LDS so that every lane stores consecutive output elements (tile
// rows of one column) as one 16-byte vector.
// ---------------------------------------------------------------------------
template <typename TSrc, typename TDst, int NW, int CW>
struct SerialTileOp {
  using Src = TSrc;
  using Dst = TDst;
  using A = arith_t<TDst>;
  static constexpr int TH = 64;        // one tile row per lane
  static constexpr int TW = NW * CW;   // columns per tile
  static constexpr int NT = 64 * NW;
  static constexpr int LPR = TW / 4;   // lanes per row segment at load time
  static constexpr int RPL = NT / LPR; // rows loaded per pass
  static constexpr int64_t VWORDS = 0;
  static constexpr int64_t HWORDS = (int64_t)TH * Carry<A>::WORDS;
  static constexpr bool GROUPS = false;
  static constexpr int64_t GWORDS = 0;
  static_assert(TW % 32 == 0 && NT % LPR == 0 && TH % RPL == 0, "tile shape");

  struct Smem {
    A st[TH * TW];  // XOR-swizzled, see stage_put4 / stage_get
    A ctot[NW][TH];
    A rp[TH];
  };

  __device__ static __forceinline__ void run(const ImgArgs& g, uint32_t t, uint32_t ep4, Smem& sm) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t I = t / g.nTx, J = t - I * g.nTx;
    const int64_t y0 = (int64_t)I * TH, x0 = (int64_t)J * TW;
    const TSrc* __restrict__ in = (const TSrc*)g.in;

    // ---- coalesced row-major load into LDS (all of a lane's loads are
    //      issued before the first is staged) ----
    {
      constexpr int NLOAD = TH / RPL;
      const int seg = tid % LPR, r0 = tid / LPR;
      const int64_t col = x0 + 4 * seg;
      A x[NLOAD][4];
      if (y0 + TH <= (int64_t)g.H && x0 + TW <= (int64_t)g.W) {
        // interior tile: straight-line vector loads
#pragma unroll
        for (int i = 0; i < NLOAD; ++i) {
          const typename VecG<TSrc>::type q =
              *(const typename VecG<TSrc>::type*)(in + (y0 + r0 + i * RPL) * g.ldi + col);
          x[i][0] = cvt_in<TSrc, TDst>(q.x);
          x[i][1] = cvt_in<TSrc, TDst>(q.y);
          x[i][2] = cvt_in<TSrc, TDst>(q.z);
          x[i][3] = cvt_in<TSrc, TDst>(q.w);
        }
      } else if (g.W >= 4) {
        // edge tile: clamped rows and chunks, one vector load per row, every
        // row's load issued before any is used
        typename VecG<TSrc>::type q[NLOAD];
        const int64_t base = edge_base(col, g.W);
#pragma unroll
        for (int i = 0; i < NLOAD; ++i) {
          const int64_t row = y0 + r0 + i * RPL;
          q[i] = *(const typename VecG<TSrc>::type*)(in + (row < (int64_t)g.H ? row : (int64_t)g.H - 1) * g.ldi + base);
        }
#pragma unroll
        for (int i = 0; i < NLOAD; ++i)
          edge_pick<TSrc, TDst>(q[i], col, base, g.W, y0 + r0 + i * RPL < (int64_t)g.H, x[i]);
      } else {
#pragma unroll
        for (int i = 0; i < NLOAD; ++i) {
          const int64_t row = y0 + r0 + i * RPL;
          load4<TSrc, TDst>(in + row * g.ldi, col, g.W, row < (int64_t)g.H, x[i]);
        }
      }
#pragma unroll
      for (int i = 0; i < NLOAD; ++i) stage_put4<A, TW>(sm.st, r0 + i * RPL, 4 * seg, x[i]);
    }
    __syncthreads();

    // ---- transposed read: lane = row, registers = CW columns; serial scan ----
    A v[CW];
#pragma unroll
    for (int k = 0; k < CW; ++k) v[k] = stage_get<A, TW>(sm.st, lane, w * CW + k);
#pragma unroll
    for (int k = 1; k < CW; ++k) v[k] += v[k - 1];
    sm.ctot[w][lane] = v[CW - 1];
    __syncthreads();

    A base = zero<A>();
    for (int q = 0; q < w; ++q) base += sm.ctot[q][lane];

    // ---- the last wave holds the row total: publish, look back, share ----
    if (w == NW - 1) {
      const A a = base + v[CW - 1];
      const int64_t e = (int64_t)t * TH + lane;
      A rp = zero<A>();
      if (J == 0) {
        Carry<A>::put(g.hcarry, e, a, ep4 | ST_INC);
      } else {
        Carry<A>::put(g.hcarry, e, a, ep4 | ST_AGG);
        rp = look_back<A, LB_WINDOW>(g.hcarry, e - TH, TH, J, ep4, g.fault);
        Carry<A>::put(g.hcarry, e, rp + a, ep4 | ST_INC);
      }
      sm.rp[lane] = rp;
    }
    __syncthreads();
    base += sm.rp[lane];

    // ---- transposed store: the results go back through LDS so that every
    //      lane writes E consecutive y of one output row x as one 16-byte
    //      vector ----
#pragma unroll
    for (int m = 0; m < CW / 4; ++m) {
      A y4[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) y4[c] = v[4 * m + c] + base;
      stage_put4<A, TW>(sm.st, lane, w * CW + 4 * m, y4);
    }
    __syncthreads();
    constexpr int E = kVecOut<TDst>;  // elements per lane store
    constexpr int SLPR = TH / E;      // lanes per output row segment
    constexpr int SRPI = NT / SLPR;   // output rows written per pass
    static_assert(TH % E == 0 && NT % SLPR == 0 && TW % SRPI == 0, "transposed store passes");
    TDst* __restrict__ out = (TDst*)g.out;
    const int seg = tid % SLPR, rsub = tid / SLPR;
    const int64_t oy0 = y0 + E * seg;
    const bool full = oy0 + E <= (int64_t)g.H;
#pragma unroll 4
    for (int xo = rsub; xo < TW; xo += SRPI) {
      const int64_t x = x0 + xo;
      A qv[E];
#pragma unroll
      for (int c = 0; c < E; ++c) qv[c] = stage_get<A, TW>(sm.st, E * seg + c, xo);
      if (x < (int64_t)g.W) {
        if (full) {
          typename VecOut<TDst>::type q;
#pragma unroll
          for (int c = 0; c < E; ++c) q[c] = cvt_out<TDst>(qv[c]);
          __builtin_nontemporal_store(q, (typename VecOut<TDst>::type*)(out + x * g.ldo + oy0));
        } else {
#pragma unroll
          for (int c = 0; c < E; ++c)
            if (oy0 + c < (int64_t)g.H) out[x * g.ldo + oy0 + c] = cvt_out<TDst>(qv[c]);
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// One tile per workgroup. The hardware hands workgroup b to XCD b % nx and each
// XCD starts its workgroups in increasing b. With order.run > 0 (tile rows
// split into runs of `run` tiles, run q of a row on XCD q % nx), XCD x takes
// its runs of every tile row, top to bottom: a tile's upper neighbour, and
// its left one inside a run, then sit on its own XCD, started just before it,
// and the row chains cross XCDs only between runs (from XCD x-1 to x, so the
// dies settle into a staggered pipeline instead of waiting on each other in a
// cycle). The map
// keeps every XCD's sequence in row-major tile order, which guarantees
// progress (see the top of this file); the bounded look-back turns a
// violation of the hardware's dispatch order into a reported fault instead of
// a hang.
// order.run == 0 keeps tile = blockIdx.x. Tile t covers plane t / tpp, tile
// t % tpp of that plane (planes stack vertically in the tile grid).
// ---------------------------------------------------------------------------
struct TileOrder {
  uint32_t nx;      // XCDs the workgroups are dealt over
  uint32_t run;     // tile columns of one run (0: identity order); divides nTx / nx
  uint32_t vgroup;  // tile rows per run of the second vertical carry level (0: none)
};

// Runs of `run` tile columns are dealt to the XCDs round-robin (run q of a
// tile row to XCD q % nx); each XCD walks its runs row by row, left to right.
__device__ __forceinline__ uint32_t tile_of_block(uint32_t b, uint32_t nTx, TileOrder o) {
  if (o.run == 0) return b;
  const uint32_t x = b % o.nx, k = b / o.nx;
  const uint32_t per = nTx / o.nx;  // tiles of one tile row on one XCD
  const uint32_t I = k / per, j = k - I * per;
  const uint32_t q = (j / o.run) * o.nx + x;
  return I * nTx + q * o.run + (j % o.run);
}

// WPE waves per SIMD (at most 512 / WPE registers per lane): the tile kernels
// are latency-bound on their look-backs, and the tile data the register file
// holds (workgroups per CU x tile bytes) bounds their throughput.
template <class Op, int WPE>
__global__ void __launch_bounds__(Op::NT) __attribute__((amdgpu_waves_per_eu(WPE)))
    tiles_kernel(ImgArgs g, uint32_t tpp, int64_t in_plane, int64_t out_plane, uint32_t ep_arg, const WsHeader* hdr,
                 TileOrder order) {
  __shared__ typename Op::Smem sm;
  const uint32_t ep4 = launch_ep4(ep_arg, hdr);
  const uint32_t t = tile_of_block(blockIdx.x, g.nTx, order);
  const uint32_t p = t / tpp, lt = t - p * tpp;
  ImgArgs gp = g;
  if (p != 0) {
    gp.in = (const typename Op::Src*)g.in + (int64_t)p * in_plane;
    gp.out = (typename Op::Dst*)g.out + (int64_t)p * out_plane;
    gp.vcarry = g.vcarry + (int64_t)p * tpp * Op::VWORDS;
    gp.hcarry = g.hcarry + (int64_t)p * tpp * Op::HWORDS;
  }
  if constexpr (Op::GROUPS) {
    VGroups vg;
    if (order.vgroup) {
      // run words follow the horizontal words of all planes, plane by plane
      const uint32_t runs = (g.nTy + order.vgroup - 1) / order.vgroup;
      vg.carry = g.hcarry + (int64_t)gridDim.x * Op::HWORDS + (int64_t)p * runs * g.nTx * Op::GWORDS;
      vg.len = order.vgroup;
    }
    Op::run(gp, lt, ep4, sm, vg);
  } else {
    Op::run(gp, lt, ep4, sm);
  }
}

}  // namespace satdev
