// Host-side launch templates shared by the kernel translation units.
#pragma once

#include "sat_host.hpp"
#include "sat_tile.hpp"

namespace sathost {

using satdev::arith_t;

// Default tile shapes: 4 waves; 16 rows per wave for 4-byte outputs (64x256
// tiles, 64 accumulators per lane), 8 rows per wave for 8-byte outputs.
template <typename TDst> constexpr int kRPT = sizeof(arith_t<TDst>) == 4 ? 16 : 8;
// serielScan pass: 8 waves of kSerialCW columns each (tiles of 64 rows x 256
// columns, 64 x 128 for 8-byte outputs); its LDS stage bounds it to two
// workgroups per CU, so 8 waves keep four per SIMD.
template <typename TDst> constexpr int kSerialCW = sizeof(arith_t<TDst>) == 4 ? 32 : 16;
constexpr int kNW_S = 8;
constexpr int kNW = 4;
// The transposed row pass (blockScan) stages its whole tile in LDS for the
// transpose, so LDS and not registers bounds its workgroups per CU: it uses
// 8 waves of half the rows (tiles of the same height, four waves per SIMD).
constexpr int kNW_T = 8;
template <typename TDst> constexpr int kRPT_T = kRPT<TDst> / 2;

// Tile configurations of the one-pass SAT (chosen per launch by the tuning
// table):
//   0: 4 waves x kRPT rows  (64 x 256 tiles for 4-byte outputs, 32 x 256 for 8-byte)
//   1: 8 waves x kRPT rows  (128 x 256 / 64 x 256): half the carry words per pixel
//      in the vertical direction, for tall images
//   2: 8 waves x kRPT/2 rows (64 x 256 / 32 x 256): half the accumulators per lane
//   3: 8 waves x 8 rows at 8 waves per SIMD (64 registers): configuration 2
//      with twice the resident waves
//   4: 4 waves x 24 rows at 3 waves per SIMD (96 x 256 / 48 x 256): three
//      workgroups per CU hold 288 KB of tile data instead of 256 KB
// NW waves per workgroup, RPT4 rows per wave for 4-byte outputs (half for
// 8-byte ones), WPE waves per SIMD the register budget is sized for.
template <bool C, typename T, typename F> using cond_t = typename satdev::cond<C, T, F>::type;

template <int CFG> struct TileCfg;
template <> struct TileCfg<0> { static constexpr int NW = 4, RPT4 = 16, WPE = 4; };
template <> struct TileCfg<1> { static constexpr int NW = 8, RPT4 = 16, WPE = 4; };
template <> struct TileCfg<2> { static constexpr int NW = 8, RPT4 = 8, WPE = 4; };
template <> struct TileCfg<3> { static constexpr int NW = 8, RPT4 = 8, WPE = 8; };
template <> struct TileCfg<4> { static constexpr int NW = 4, RPT4 = 24, WPE = 3; };
// 5: 1-byte inputs into 4-byte outputs take PackedTileOp (4 waves x 64 packed
//    rows, 256 x 256 tiles); every other pair takes configuration 1's tile
template <> struct TileCfg<5> { static constexpr int NW = 8, RPT4 = 16, WPE = 4; };
// 6: 4 waves x 32 rows (16 for 8-byte outputs) at 2 waves per SIMD
template <> struct TileCfg<6> { static constexpr int NW = 4, RPT4 = 32, WPE = 2; };
// 7: 1-byte inputs into 4-byte outputs take PackedTileOp with 4 waves x 32
//    packed rows (128 x 256 tiles); every other pair takes configuration 0's
//    tile
template <> struct TileCfg<7> { static constexpr int NW = 4, RPT4 = 16, WPE = 4; };
// packed rows per wave of configuration CFG for 1-byte inputs (0: unpacked)
template <int CFG> constexpr int packed_rpt = CFG == 5 ? 64 : CFG == 7 ? 32 : 0;
template <typename S, typename D> constexpr bool packable = sizeof(S) == 1 && sizeof(arith_t<D>) == 4;
template <typename S, typename D, int SC, int RPT> using PackedOp = satdev::PackedTileOp<S, D, SC, 4, RPT>;
template <int CFG, typename TDst>
constexpr int cfg_rpt = sizeof(arith_t<TDst>) == 4 ? TileCfg<CFG>::RPT4 : TileCfg<CFG>::RPT4 / 2;

template <class Op>
inline Geometry geom_of() {
  Geometry g;
  g.th = Op::TH;
  g.tw = Op::TW;
  g.vwords = Op::VWORDS;
  g.hwords = Op::HWORDS;
  g.gwords = Op::GWORDS;
  g.threads = Op::NT;
  return g;
}

template <class Op>
inline uint64_t op_carry_words(const Launch& L) {
  const uint64_t nTx = (L.width + Op::TW - 1) / Op::TW, nTy = (L.height + Op::TH - 1) / Op::TH;
  const uint64_t ntiles = nTx * nTy * L.planes;
  return ntiles * (uint64_t)(Op::VWORDS + Op::HWORDS);
}

// Tile columns of one XCD run, given the `per` tile columns of a tile row
// that one XCD takes: at most 4 (the largest divisor of `per` up to 4). A row
// chain crosses XCDs between runs; runs of 4 measured 4 % faster than 8 on
// u8 -> u32 16384^2 (64 tile columns) and than 32 on f32 256 x 65536, and
// equal on the 8- and 16-wide XCD shares of f32 8192^2 and 16384^2.
inline uint32_t run_length(uint32_t per) {
  for (uint32_t r = 4; r > 1; --r)
    if (per % r == 0) return r;
  return 1;
}

template <class Op, int WPE = 4>
inline sat_status launch_op(const Launch& L, void* ws, size_t ws_bytes, hipStream_t stream) {
  const uint64_t nTx = (L.width + Op::TW - 1) / Op::TW, nTy = (L.height + Op::TH - 1) / Op::TH;
  const uint64_t tpp = nTx * nTy, ntiles = tpp * L.planes;
  if (ntiles == 0) return SAT_OK;
  if (ntiles > 0x7fffffffull) return SAT_ERR_INVALID;
  const uint64_t vw = ntiles * Op::VWORDS, hw = ntiles * Op::HWORDS;
  uint32_t vgroup = Op::GROUPS ? vgroup_len(nTx, nTy) : 0u;
  const uint64_t gw = vgroup ? (uint64_t)L.planes * ((nTy + vgroup - 1) / vgroup) * nTx * Op::GWORDS : 0u;
  // the run words are laid out whenever the shape qualifies (workspace sizes
  // depend on the shape only) but used only when every tile of the launch is
  // resident at once: with later generations of tiles they measured slower
  // (f32 262144 x 256 105 -> 114 us, 1048576 x 256 359 -> 445 us)
  if (vgroup && ntiles > (uint64_t)device_cu_count() * (4 * WPE / (Op::NT / 64))) vgroup = 0;
  const uint64_t need = satdev::WS_HEADER_BYTES + (vw + hw + gw) * 8;
  if (ws == nullptr) return SAT_ERR_NULL;
  if (ws_bytes < need) return SAT_ERR_WORKSPACE;
  uint64_t* carries = (uint64_t*)((char*)ws + satdev::WS_HEADER_BYTES);
  satdev::ImgArgs g;
  g.in = L.in;
  g.out = L.out;
  g.ldi = L.ldi;
  g.ldo = L.ldo;
  g.H = L.height;
  g.W = L.width;
  g.nTx = (uint32_t)nTx;
  g.nTy = (uint32_t)nTy;
  g.vcarry = carries;
  g.hcarry = carries + vw;
  g.fault.ws = &((satdev::WsHeader*)ws)->fault;
  g.fault.host = host_fault_word(L.may_alloc);
  satdev::TileOrder order;
  order.nx = (uint32_t)device_xcd_count();
  order.run = (order.nx > 1 && nTx % order.nx == 0) ? run_length((uint32_t)(nTx / order.nx)) : 0u;
  order.vgroup = vgroup;
  hipLaunchKernelGGL((satdev::tiles_kernel<Op, WPE>), dim3((uint32_t)ntiles), dim3(Op::NT), 0, stream, g, (uint32_t)tpp,
                     L.in_plane, L.out_plane, L.epoch, (const satdev::WsHeader*)ws, order);
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

// The reference dispatchers' dtype pairs (ScanRow-BRLT.cu:274-376) plus u32->u32.
// X(tin_enum, tout_enum, TSrc, TDst, float_out)
#define SAT_FOR_EACH_PAIR(X)                  \
  X(SAT_U8, SAT_U32, uint8_t, uint32_t, 0)    \
  X(SAT_U8, SAT_F32, uint8_t, float, 1)       \
  X(SAT_U8, SAT_F64, uint8_t, double, 1)      \
  X(SAT_I8, SAT_I32, int8_t, int32_t, 0)      \
  X(SAT_I8, SAT_F32, int8_t, float, 1)        \
  X(SAT_I8, SAT_F64, int8_t, double, 1)       \
  X(SAT_I32, SAT_I32, int32_t, int32_t, 0)    \
  X(SAT_I32, SAT_F64, int32_t, double, 1)     \
  X(SAT_F32, SAT_F32, float, float, 1)        \
  X(SAT_F32, SAT_F64, float, double, 1)       \
  X(SAT_F64, SAT_F64, double, double, 1)      \
  X(SAT_U32, SAT_U32, uint32_t, uint32_t, 0)

// Dispatch F<TSrc, TDst, SCAN>::go(L, ws, bytes, stream) on (L.tin, L.tout,
// L.scan). Integer outputs are order-independent, so for them only the
// Kogge-Stone order is instantiated.
template <bool FLOAT_OUT> struct ScanPick;
template <> struct ScanPick<true> {
  template <template <typename, typename, int> class F, typename S, typename D>
  static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) {
    return L.scan == SAT_SCAN_LF ? F<S, D, 1>::go(L, ws, b, st) : F<S, D, 0>::go(L, ws, b, st);
  }
};
template <> struct ScanPick<false> {
  template <template <typename, typename, int> class F, typename S, typename D>
  static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) {
    return F<S, D, 0>::go(L, ws, b, st);
  }
};

template <template <typename, typename, int> class F>
inline sat_status dispatch_pairs(const Launch& L, void* ws, size_t b, hipStream_t st) {
#define SAT_X(a, b_, S, D, fl) \
  if (L.tin == a && L.tout == b_) return ScanPick<fl != 0>::template go<F, S, D>(L, ws, b, st);
  SAT_FOR_EACH_PAIR(SAT_X)
#undef SAT_X
  return SAT_ERR_DTYPE;
}

// One-pass SAT (forward and reverse) in tile configuration CFG; every
// configuration lives in its own translation unit (k_sat*.hip).
template <int CFG> struct FamSatCfg {
  template <typename S, typename D, int SC> struct Fwd {
    using Op = cond_t<(packed_rpt<CFG> > 0) && packable<S, D>, PackedOp<S, D, SC, packed_rpt<CFG> ? packed_rpt<CFG> : 64>,
                      satdev::TileOp<S, D, satdev::MODE_SAT, false, false, SC, TileCfg<CFG>::NW, cfg_rpt<CFG, D>>>;
    static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) {
      return launch_op<Op, TileCfg<CFG>::WPE>(L, ws, b, st);
    }
  };
  template <typename S, typename D, int SC> struct Rev {
    using Op = satdev::TileOp<S, D, satdev::MODE_SAT, true, false, SC, TileCfg<CFG>::NW, cfg_rpt<CFG, D>>;
    static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) {
      return launch_op<Op, TileCfg<CFG>::WPE>(L, ws, b, st);
    }
  };
};

template <int CFG>
inline sat_status launch_sat_cfg(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return dispatch_pairs<FamSatCfg<CFG>::template Fwd>(L, ws, b, st);
}

// The adjoint is defined for same-type pairs.
template <int CFG>
inline sat_status launch_sat_rev_cfg(const Launch& L, void* ws, size_t b, hipStream_t st) {
  if (L.tin != L.tout) return SAT_ERR_DTYPE;
  using F = FamSatCfg<CFG>;
  switch (L.tin) {
    case SAT_F32: return ScanPick<true>::go<F::template Rev, float, float>(L, ws, b, st);
    case SAT_F64: return ScanPick<true>::go<F::template Rev, double, double>(L, ws, b, st);
    case SAT_U32: return ScanPick<false>::go<F::template Rev, uint32_t, uint32_t>(L, ws, b, st);
    case SAT_I32: return ScanPick<false>::go<F::template Rev, int32_t, int32_t>(L, ws, b, st);
    default: return SAT_ERR_DTYPE;
  }
}

template <int CFG>
inline Geometry sat_geometry_cfg(size_t out_size, size_t in_size) {
  if (out_size == 8) return geom_of<typename FamSatCfg<CFG>::template Fwd<double, double, 0>::Op>();
  if (in_size == 1) return geom_of<typename FamSatCfg<CFG>::template Fwd<uint8_t, uint32_t, 0>::Op>();
  return geom_of<typename FamSatCfg<CFG>::template Fwd<float, float, 0>::Op>();
}

sat_status launch_fam_sat_c1(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_rev_c1(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_c2(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_rev_c2(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_c3(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_rev_c3(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_c4(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_rev_c4(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_c5(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_rev_c5(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_c6(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_rev_c6(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_c7(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_rev_c7(const Launch& L, void* ws, size_t b, hipStream_t st);

}  // namespace sathost
