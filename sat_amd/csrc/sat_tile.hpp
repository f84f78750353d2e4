// Tile operators of the MI355X summed-area-table kernels.
//
// A tile is TH rows x 256 columns of the input. A workgroup of NW waves loads
// it with one coalesced 4-element access per lane and row (wave w holds rows
// [w*RPT, (w+1)*RPT), lane l holds columns [4l, 4l+4)), scans it in registers,
// and joins its neighbours through decoupled look-back on tagged carry words:
//
//   vertical carries   one word per tile column: the tile's column sums (AGG)
//                      and the column sums of everything down to and including
//                      the tile (INC);
//   horizontal carries one word per tile row: the sum of the row prefix of the
//                      tile's rows restricted to the tile's columns (AGG) and
//                      the row's SAT value at the tile's right edge (INC).
//
// Tiles are numbered by tickets taken from a counter when a workgroup starts,
// row-major over the tile grid, so every predecessor a tile waits for has
// already been handed to a running workgroup.
#pragma once

#include "sat_device.hpp"

namespace satdev {

enum : int { MODE_ROW = 1, MODE_COL = 2, MODE_SAT = 3 };

template <bool C, typename T, typename F> struct cond { using type = T; };
template <typename T, typename F> struct cond<false, T, F> { using type = F; };

// ---------------------------------------------------------------------------
// Register tile operator. MODE_SAT: 2-D inclusive prefix sum (one pass).
// MODE_ROW: row prefix sums (IncrementScan::ScanX semantics, or with TOUT the
// transposed row scan of BlockScan::blockScan). MODE_COL: column prefix sums
// (IncrementScan::ScanY semantics). REV mirrors both axes (suffix sums), the
// adjoint of the forward SAT.
// ---------------------------------------------------------------------------
template <typename TSrc, typename TDst, int MODE, bool REV, bool TOUT, int SCAN, int NW, int RPT>
struct TileOp {
  using Src = TSrc;
  using Dst = TDst;
  using A = arith_t<TDst>;
  static constexpr int TW = 256;
  static constexpr int TH = NW * RPT;
  static constexpr int NT = 64 * NW;
  static constexpr bool DO_ROW = (MODE & MODE_ROW) != 0;
  static constexpr bool DO_COL = (MODE & MODE_COL) != 0;
  static constexpr int CPT_V = (TW + NT - 1) / NT;
  static constexpr int64_t VWORDS = DO_COL ? (int64_t)TW * Carry<A>::WORDS : 0;  // per tile
  static constexpr int64_t HWORDS = DO_ROW ? (int64_t)TH * Carry<A>::WORDS : 0;  // per tile
  static_assert(!(TOUT && (REV || DO_COL)), "transposed store is a row-scan pass only");
  static_assert(RPT <= 64, "one lane per tile row in the horizontal phase");

  struct SmemCarry {
    A ct[DO_COL ? NW : 1][TW];  // per-wave column totals
    A cab[DO_COL ? TW : 1];     // column sums above the tile
  };
  struct SmemStage {
    A st[TH * TW];  // XOR-swizzled, see stage_put4 / stage_get
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

  __device__ static __forceinline__ void run(const ImgArgs& g, uint32_t t, uint32_t ep4, Smem& sm) {
    const Coords k = coords(g, t);
    const TSrc* __restrict__ in = (const TSrc*)g.in;
    const bool vin = g.vin != 0;
    // ---- load (each wave: RPT coalesced row segments of 256 elements) ----
    A v[RPT][4];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int64_t prow = prow_of(k, r);
      A x[4];
      load4<TSrc, TDst>(in + prow * g.ldi, k.pcol, g.W, prow < (int64_t)g.H, vin, x);
#pragma unroll
      for (int c = 0; c < 4; ++c) v[r][c] = REV ? x[3 - c] : x[c];
    }
    finish(g, t, ep4, sm, k, v);
  }

  // Everything after the load: scans, look-backs, store. `v` holds the tile
  // in logical order (mirrored when REV).
  __device__ static __forceinline__ void finish(const ImgArgs& g, uint32_t t, uint32_t ep4, Smem& sm,
                                                const Coords& k, A (&v)[RPT][4]) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t I = k.I, J = k.J;
    const int64_t py0 = k.py0, px0 = k.px0, pcol = k.pcol;
    const bool vout = g.vout != 0;

    // ---- column prefix inside the wave; publish the tile's column sums ----
    A colsum[CPT_V];
    if constexpr (DO_COL) {
#pragma unroll
      for (int r = 1; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] += v[r - 1][c];
#pragma unroll
      for (int c = 0; c < 4; ++c) sm.ct[w][4 * lane + c] = v[RPT - 1][c];
      __syncthreads();
#pragma unroll
      for (int s = 0; s < CPT_V; ++s) {
        const int j = tid + s * NT;
        if (j < TW) {
          A cs = sm.ct[0][j];
#pragma unroll
          for (int q = 1; q < NW; ++q) cs += sm.ct[q][j];
          colsum[s] = cs;
          Carry<A>::put(g.vcarry, (int64_t)t * TW + j, cs, ep4 | (I == 0 ? ST_INC : ST_AGG));
        }
      }
    }

    // ---- row prefix: in-lane over 4 columns, then across the 64 lanes ----
    A rowtot[RPT];
    if constexpr (DO_ROW) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        v[r][1] += v[r][0];
        v[r][2] += v[r][1];
        v[r][3] += v[r][2];
        const A inc = wave_scan<SCAN>(v[r][3], lane);
        const A exc = wave_shr1(inc);
        rowtot[r] = readlane(inc, 63);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] += exc;
      }
    }

    // ---- vertical look-back; per-column carry into this wave's rows ----
    A colAdd[4] = {zero<A>(), zero<A>(), zero<A>(), zero<A>()};
    A Tw = zero<A>();
    if constexpr (DO_COL) {
#pragma unroll
      for (int s = 0; s < CPT_V; ++s) {
        const int j = tid + s * NT;
        if (j < TW) {
          A acc = zero<A>();
          if (I > 0) {
            acc = look_back<A>(g.vcarry, (int64_t)(t - g.nTx) * TW + j, (int64_t)g.nTx * TW, I, ep4);
            Carry<A>::put(g.vcarry, (int64_t)t * TW + j, acc + colsum[s], ep4 | ST_INC);
          }
          sm.cab[j] = acc;
        }
      }
      __syncthreads();
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
        Tw = readlane(inc, 63);
#pragma unroll
        for (int c = 0; c < 4; ++c) colAdd[c] = U[c] + exc;
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) colAdd[c] = U[c];
      }
    }

    // ---- horizontal look-back: lane r of wave w owns tile row w*RPT + r ----
    A rp[RPT];
    if constexpr (DO_ROW) {
      A my_a = zero<A>();
#pragma unroll
      for (int r = 0; r < RPT; ++r)
        if (lane == r) my_a = rowtot[r] + Tw;
      A my_rp = zero<A>();
      if (lane < RPT) {
        const int64_t e = (int64_t)t * TH + w * RPT + lane;
        if (J == 0) {
          Carry<A>::put(g.hcarry, e, my_a, ep4 | ST_INC);
        } else {
          Carry<A>::put(g.hcarry, e, my_a, ep4 | ST_AGG);
          my_rp = look_back<A>(g.hcarry, e - TH, TH, J, ep4);
          Carry<A>::put(g.hcarry, e, my_rp + my_a, ep4 | ST_INC);
        }
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r) rp[r] = readlane(my_rp, r);
    }

    // ---- store ----
    if constexpr (!TOUT) {
      TDst* __restrict__ out = (TDst*)g.out;
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
        store4<TDst>(out + prow * g.ldo, pcol, g.W, prow < (int64_t)g.H, vout, y);
      }
    } else {
      // Block-register-local transpose: stage the scanned tile in LDS, then
      // write output rows (= input columns); the TH lanes of a group write TH
      // consecutive output elements.
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        A y[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) y[c] = v[r][c] + rp[r];
        stage_put4<A, TW>(sm.st, w * RPT + r, 4 * lane, y);
      }
      __syncthreads();
      constexpr int GROUPS = NT / TH;  // output rows written per pass
      static_assert(NT % TH == 0 && TW % GROUPS == 0, "transpose passes");
      TDst* __restrict__ out = (TDst*)g.out;
      const int yl = tid % TH, grp = tid / TH;
      const int64_t oy = py0 + yl;
      const bool y_ok = oy < (int64_t)g.H;
#pragma unroll 8
      for (int xo = grp; xo < TW; xo += GROUPS) {
        const int64_t orow = px0 + xo;
        const A val = stage_get<A, TW>(sm.st, yl, xo);
        if (y_ok && orow < (int64_t)g.W) out[orow * g.ldo + oy] = cvt_out<TDst>(val);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// BRLT-then-serial-scan operator (SerielScan::serielScan semantics: transposed
// row scan). The tile (64 rows x NW*CW columns) is loaded row-major, moved
// through LDS so that lane = tile row and registers = a run of CW consecutive
// columns owned by the wave, scanned serially in registers, joined across the
// waves' column runs and across tiles (horizontal look-back), and stored so
// that for every register the 64 lanes write 64 consecutive output elements.
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

    // ---- coalesced row-major load into LDS ----
    {
      const int seg = tid % LPR, r0 = tid / LPR;
      const int64_t col = x0 + 4 * seg;
#pragma unroll 4
      for (int r = r0; r < TH; r += RPL) {
        const int64_t row = y0 + r;
        A x[4];
        load4<TSrc, TDst>(in + row * g.ldi, col, g.W, row < (int64_t)g.H, g.vin != 0, x);
        stage_put4<A, TW>(sm.st, r, 4 * seg, x);
      }
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
        rp = look_back<A>(g.hcarry, e - TH, TH, J, ep4);
        Carry<A>::put(g.hcarry, e, rp + a, ep4 | ST_INC);
      }
      sm.rp[lane] = rp;
    }
    __syncthreads();
    base += sm.rp[lane];

    // ---- transposed store: out[x][y], 64 lanes = 64 consecutive y ----
    TDst* __restrict__ out = (TDst*)g.out;
    const int64_t y = y0 + lane;
    const bool y_ok = y < (int64_t)g.H;
#pragma unroll
    for (int k = 0; k < CW; ++k) {
      const int64_t x = x0 + w * CW + k;
      if (y_ok && x < (int64_t)g.W) out[x * g.ldo + y] = cvt_out<TDst>(v[k] + base);
    }
  }
};

// ---------------------------------------------------------------------------
// One tile per workgroup; the workgroup count equals the tile count of all
// planes. Ticket t covers plane t / tpp, tile t % tpp of that plane.
// ---------------------------------------------------------------------------
template <class Op>
__global__ void __launch_bounds__(Op::NT) tiles_kernel(ImgArgs g, uint32_t tpp, int64_t in_plane, int64_t out_plane,
                                                       WsHeader* hdr, uint64_t* carries, uint64_t words_used) {
  __shared__ typename Op::Smem sm;
  __shared__ uint32_t s_ticket, s_ep, s_flag;
  if (threadIdx.x == 0) {
    s_ticket = take_ticket(hdr);
    s_ep = read_epoch(hdr);
  }
  __syncthreads();
  const uint32_t t = s_ticket;
  const uint32_t p = t / tpp, lt = t - p * tpp;
  ImgArgs gp = g;
  if (p != 0) {
    gp.in = (const typename Op::Src*)g.in + (int64_t)p * in_plane;
    gp.out = (typename Op::Dst*)g.out + (int64_t)p * out_plane;
    gp.vcarry = g.vcarry + (int64_t)p * tpp * Op::VWORDS;
    gp.hcarry = g.hcarry + (int64_t)p * tpp * Op::HWORDS;
  }
  Op::run(gp, lt, s_ep << 2, sm);
  finish_launch(hdr, carries, words_used, &s_flag);
}

}  // namespace satdev
