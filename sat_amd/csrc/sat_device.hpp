// Device-side building blocks of the MI355X (gfx950) summed-area-table kernels.
//
// Everything here is wave64 code: cross-lane scans use DPP row shifts and the
// GFX9 row broadcasts, tile carries travel through 64-bit "tagged" words whose
// upper half holds (epoch << 2 | status) and whose lower half holds the payload.
// A payload wider than 32 bits is split over two tagged words.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace satdev {

// ---------------------------------------------------------------------------
// Arithmetic type. Integer outputs accumulate in unsigned 32-bit arithmetic so
// that overflow wraps modulo 2^32 (the reference's `uint` accumulators,
// integralImage.h:105-124, and the wrap the u32 / u8->u32 workloads rely on);
// an int32 output is the same bit pattern reinterpreted.
// ---------------------------------------------------------------------------
template <typename T> struct arith { using type = T; };
template <> struct arith<int32_t> { using type = uint32_t; };
template <typename T> using arith_t = typename arith<T>::type;

// TSrc -> TDst -> arithmetic type A: the reference converts each pixel to
// TDst before adding it (`sum += src[i]` with a TDst accumulator,
// integralImage.h:79). A is arith_t<TDst> except where a kernel accumulates
// a narrower output in a wider type (the f32 box gradient in f64).
template <typename TSrc, typename TDst, typename A = arith_t<TDst>>
__device__ __forceinline__ A cvt_in(TSrc x) {
  return (A)(TDst)x;
}
template <typename TDst, typename A>
__device__ __forceinline__ TDst cvt_out(A x) {
  return (TDst)x;
}

template <typename A> __device__ __forceinline__ A zero() { return (A)0; }

// ---------------------------------------------------------------------------
// Cross-lane primitives (wave64).
// ---------------------------------------------------------------------------
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp32(uint32_t x) {
  // old = 0 and bound_ctrl: lanes with no source lane, or outside ROWMASK, read 0
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWMASK, 0xf, true);
}

template <int CTRL, int ROWMASK, typename A>
__device__ __forceinline__ A dpp(A x) {
  if constexpr (sizeof(A) == 4) {
    return __builtin_bit_cast(A, dpp32<CTRL, ROWMASK>(__builtin_bit_cast(uint32_t, x)));
  } else {
    static_assert(sizeof(A) == 8, "4- or 8-byte arithmetic types only");
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = dpp32<CTRL, ROWMASK>((uint32_t)b);
    const uint32_t hi = dpp32<CTRL, ROWMASK>((uint32_t)(b >> 32));
    return __builtin_bit_cast(A, ((uint64_t)hi << 32) | lo);
  }
}

// DPP controls (GFX9 encoding)
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int DPP_ROW_SHR2 = 0x112;
constexpr int DPP_ROW_SHR4 = 0x114;
constexpr int DPP_ROW_SHR8 = 0x118;
constexpr int DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

// Inclusive scan over the 64 lanes of a wave: Hillis-Steele inside each
// 16-lane DPP row, then the row totals are folded in with the two GFX9 row
// broadcasts (lane 15 -> rows 1 and 3, lane 31 -> rows 2 and 3). This is the
// Kogge-Stone order (the reference's KoggeStone_SCAN, ScanRow-BRLT.cu:58-67)
// re-expressed for 64-wide waves.
template <typename A>
__device__ __forceinline__ A wave_scan_ks(A x) {
  x += dpp<DPP_ROW_SHR1, 0xf>(x);
  x += dpp<DPP_ROW_SHR2, 0xf>(x);
  x += dpp<DPP_ROW_SHR4, 0xf>(x);
  x += dpp<DPP_ROW_SHR8, 0xf>(x);
  x += dpp<DPP_ROW_BCAST15, 0xa>(x);
  x += dpp<DPP_ROW_BCAST31, 0xc>(x);
  return x;
}

template <typename A>
__device__ __forceinline__ uint32_t bits_lo(A x) {
  if constexpr (sizeof(A) == 4) return __builtin_bit_cast(uint32_t, x);
  else return (uint32_t)__builtin_bit_cast(uint64_t, x);
}

template <int PATTERN>
__device__ __forceinline__ uint32_t swz32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, PATTERN);
}
template <int PATTERN, typename A>
__device__ __forceinline__ A swz(A x) {
  if constexpr (sizeof(A) == 4) {
    return __builtin_bit_cast(A, swz32<PATTERN>(__builtin_bit_cast(uint32_t, x)));
  } else {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = swz32<PATTERN>((uint32_t)b);
    const uint32_t hi = swz32<PATTERN>((uint32_t)(b >> 32));
    return __builtin_bit_cast(A, ((uint64_t)hi << 32) | lo);
  }
}

template <typename A>
__device__ __forceinline__ A readlane(A x, int l) {
  if constexpr (sizeof(A) == 4) {
    return __builtin_bit_cast(A, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
  } else {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(A, ((uint64_t)hi << 32) | lo);
  }
}

// ds_swizzle bit-mask mode: source lane (within 32) = ((lane & and) | or) ^ xor
constexpr int swz_mask(int and_mask, int or_mask, int xor_mask) {
  return (and_mask & 0x1f) | ((or_mask & 0x1f) << 5) | ((xor_mask & 0x1f) << 10);
}

// Inclusive scan over 64 lanes in the Ladner-Fischer (Sklansky) order: at step
// k every lane whose bit k is set adds the last lane of the preceding 2^k
// block (the reference's LF_SCAN, ScanRow-BRLT.cu:68-76). Steps k = 0..4 are
// ds_swizzle broadcasts inside 32-lane halves; step 5 broadcasts lane 31.
template <typename A>
__device__ __forceinline__ A wave_scan_lf(A x, int lane) {
  A t;
  t = swz<swz_mask(0x1e, 0x00, 0)>(x); if (lane & 1)  x += t;   // src = lane & ~1
  t = swz<swz_mask(0x1c, 0x01, 0)>(x); if (lane & 2)  x += t;   // src = (lane & ~3) | 1
  t = swz<swz_mask(0x18, 0x03, 0)>(x); if (lane & 4)  x += t;   // src = (lane & ~7) | 3
  t = swz<swz_mask(0x10, 0x07, 0)>(x); if (lane & 8)  x += t;   // src = (lane & ~15) | 7
  t = swz<swz_mask(0x00, 0x0f, 0)>(x); if (lane & 16) x += t;   // src = 15 (within the half)
  t = readlane(x, 31);                  if (lane & 32) x += t;
  return x;
}

template <int SCAN, typename A>
__device__ __forceinline__ A wave_scan(A x, int lane) {
  if constexpr (SCAN == 1) return wave_scan_lf(x, lane);
  else return wave_scan_ks(x);
}

// exclusive value = inclusive value of the previous lane (lane 0 reads 0)
template <typename A>
__device__ __forceinline__ A wave_shr1(A x) {
  return dpp<DPP_WAVE_SHR1, 0xf>(x);
}

// ---------------------------------------------------------------------------
// Tagged carry words. tag = (epoch << 2) | status, status 1 = aggregate of the
// tile alone, 2 = inclusive prefix through the tile. A word whose epoch is not
// the current launch's reads as "not yet published".
// ---------------------------------------------------------------------------
constexpr uint32_t ST_AGG = 1;
constexpr uint32_t ST_INC = 2;

__device__ __forceinline__ uint64_t ld_word(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_word(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The two words of an 8-byte payload move as one 16-byte access (a volatile
// global dwordx4 at system scope, which like the agent-scope words reaches
// memory on every store and load). Whether or not the hardware splits it, a
// reader that sees halves of two different writes finds two different tags,
// since each entry is written once per (epoch, status), and reads again. One
// request per entry instead of two measured f64 8192^2 260 -> 245 us and the
// f32 -> f64 table of 96 x 1024^2 296 -> 285 us.
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u64x2 ld_pair(const uint64_t* p) {
  return *(const volatile __attribute__((address_space(1))) u64x2*)p;
}
__device__ __forceinline__ void st_pair(uint64_t* p, u64x2 v) {
  *(volatile __attribute__((address_space(1))) u64x2*)p = v;
}

template <typename A> struct Carry {
  static constexpr int WORDS = sizeof(A) / 4;

  __device__ static __forceinline__ void put(uint64_t* base, int64_t e, A v, uint32_t tag) {
    const uint64_t t = (uint64_t)tag << 32;
    if constexpr (WORDS == 1) {
      st_word(base + e, t | __builtin_bit_cast(uint32_t, v));
    } else {
      const uint64_t b = __builtin_bit_cast(uint64_t, v);
      u64x2 w;
      w.x = t | (uint32_t)(b >> 32);
      w.y = t | (uint32_t)b;
      st_pair(base + 2 * e, w);
    }
  }

  // The raw word(s) of entry e. Loads and their use are separate steps so
  // that a look-back round issues every load of its window before it waits
  // for any (a use right after each load costs one memory round trip per
  // word).
  struct Raw {
    uint64_t w[WORDS];
  };
  __device__ static __forceinline__ Raw load(const uint64_t* base, int64_t e) {
    Raw r;
    if constexpr (WORDS == 1) {
      r.w[0] = ld_word(base + e);
    } else {
      const u64x2 w = ld_pair(base + 2 * e);
      r.w[0] = w.x;
      r.w[1] = w.y;
    }
    return r;
  }
  // 0: not published in this launch yet, otherwise ST_AGG / ST_INC
  __device__ static __forceinline__ uint32_t decode(const Raw& r, uint32_t ep4, A& v) {
    if constexpr (WORDS == 1) {
      const uint32_t tag = (uint32_t)(r.w[0] >> 32);
      if ((tag & ~3u) != ep4) return 0;
      v = __builtin_bit_cast(A, (uint32_t)r.w[0]);
      return tag & 3u;
    } else {
      const uint32_t th = (uint32_t)(r.w[0] >> 32), tl = (uint32_t)(r.w[1] >> 32);
      if (th != tl || (th & ~3u) != ep4) return 0;  // torn AGG/INC pair: retry
      v = __builtin_bit_cast(A, ((uint64_t)(uint32_t)r.w[0] << 32) | (uint32_t)r.w[1]);
      return th & 3u;
    }
  }
  __device__ static __forceinline__ uint32_t get(const uint64_t* base, int64_t e, uint32_t ep4, A& v) {
    return decode(load(base, e), ep4, v);
  }
};

// Polls of an unpublished carry word after which a look-back gives up (about
// a second). A launch whose carries stop arriving (a corrupted workspace, a
// workspace shared by overlapping calls) then ends instead of hanging the
// GPU: the give-up is recorded in the workspace header's fault word, which
// the host reads with sat_workspace_fault(), and in a word of pinned host
// memory per device, which the host reads without synchronising
// (sat_device_fault()).
constexpr uint32_t SPIN_LIMIT = 1u << 20;

// Where a give-up is recorded: the workspace's fault word and the device's
// host-visible sticky flag (null when the library could not map one).
struct Fault {
  uint32_t* ws;
  uint32_t* host;
};

__device__ __forceinline__ void raise_fault(const Fault& f) {
  __hip_atomic_fetch_or(f.ws, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (f.host) __hip_atomic_store(f.host, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Wait until a predecessor has published. Only the first calling lane of the
// wave polls (its own nearest word), so a waiting wave costs one memory
// request per poll instead of a window of words per lane; false after
// SPIN_LIMIT polls (recorded in the fault words).
template <typename A>
__device__ __forceinline__ bool wait_lead(const uint64_t* base, int64_t e, uint32_t ep4, const Fault& fault,
                                          uint32_t& spins) {
  const int lead = __builtin_ctzll(__ballot(1));
  for (;;) {
    uint32_t s = 0;
    if ((int)(threadIdx.x & 63) == lead) {
      A v;
      s = Carry<A>::get(base, e, ep4, v);
    }
    if (__builtin_amdgcn_readlane(s, lead) != 0) return true;
    if (++spins >= SPIN_LIMIT) {
      if (threadIdx.x % 64 == (uint32_t)lead) raise_fault(fault);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Decoupled look-back over `n` predecessors, nearest first. `e` is the carry
// entry of the nearest predecessor, `step` the entry distance to the next one.
// Each round reads a window of D predecessors (for one-word payloads all its
// loads at once, one memory round trip), then folds the longest published
// prefix of the window:
// aggregates are summed until an inclusive prefix ends the walk. A window
// whose nearest word is not yet published is retried once wait_lead has
// seen it published. With `got_inc`, reports whether an inclusive word ended
// the walk (false: all n predecessors were aggregates, or a give-up).
template <typename A, int D>
__device__ __forceinline__ A look_back(const uint64_t* base, int64_t e, int64_t step, uint32_t n, uint32_t ep4,
                                       const Fault& fault, bool* got_inc = nullptr) {
  A acc = zero<A>();
  uint32_t spins = 0;
  if (got_inc) *got_inc = false;
  while (n > 0) {
    A v[D];
    uint32_t st[D];
    if constexpr (Carry<A>::WORDS == 1) {
      // every load of the window first (a window slot past the last
      // predecessor re-reads that one and is ignored), then their use
      typename Carry<A>::Raw raw[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const uint32_t dd = (uint32_t)d < n ? (uint32_t)d : n - 1;
        raw[d] = Carry<A>::load(base, e - (int64_t)dd * step);
      }
#pragma unroll
      for (int d = 0; d < D; ++d) {
        v[d] = zero<A>();
        st[d] = Carry<A>::decode(raw[d], ep4, v[d]);
        if ((uint32_t)d >= n) st[d] = 0;
      }
    } else {
      // two-word payloads: one window slot after the other (all at once
      // measured 9 % slower on a 512-image f64 batch, and within 2 % on
      // single f64 images; the same with one 16-byte load per slot)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        st[d] = 0;
        v[d] = zero<A>();
        if ((uint32_t)d < n) st[d] = Carry<A>::get(base, e - (int64_t)d * step, ep4, v[d]);
      }
    }
    uint32_t used = 0;
    bool done = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (!done && used == (uint32_t)d && st[d] != 0) {
        acc += v[d];
        used = d + 1;
        done = st[d] == ST_INC;
      }
    }
    if (done) {
      if (got_inc) *got_inc = true;
      break;
    }
    if (used == 0) {
      if (!wait_lead<A>(base, e, ep4, fault, spins)) break;
      continue;
    }
    e -= (int64_t)used * step;
    n -= used;
  }
  return acc;
}

// Two look-back chains of one-word payloads in adjacent words (entries e and
// e + 1 of every predecessor), published together by one 16-byte store: each
// round reads both words of one predecessor with one 16-byte load. Both
// chains stay on the same predecessor; a chain that met an inclusive word
// ignores the rest. A pair with an unpublished word of an unfinished chain is
// read again once wait_lead has seen that word published.
template <typename A>
__device__ __forceinline__ void look_back_pair(const uint64_t* base, int64_t e, int64_t step, uint32_t n,
                                               uint32_t ep4, const Fault& fault, A (&acc)[2]) {
  static_assert(Carry<A>::WORDS == 1, "one-word payloads");
  acc[0] = zero<A>();
  acc[1] = zero<A>();
  bool fin0 = n == 0, fin1 = n == 0;
  uint32_t spins = 0;
  while (!(fin0 && fin1)) {
    const u64x2 w = ld_pair(base + e);
    typename Carry<A>::Raw r0, r1;
    r0.w[0] = w.x;
    r1.w[0] = w.y;
    A v0, v1;
    const uint32_t s0 = Carry<A>::decode(r0, ep4, v0), s1 = Carry<A>::decode(r1, ep4, v1);
    if ((!fin0 && s0 == 0) || (!fin1 && s1 == 0)) {
      if (!wait_lead<A>(base, (!fin0 && s0 == 0) ? e : e + 1, ep4, fault, spins)) break;
      continue;
    }
    if (!fin0) {
      acc[0] += v0;
      fin0 = s0 == ST_INC;
    }
    if (!fin1) {
      acc[1] += v1;
      fin1 = s1 == ST_INC;
    }
    if (--n == 0) break;
    e -= step;
  }
}

// K independent look-back chains per lane over the same predecessors (for
// example two columns of one tile). With BATCH every round issues the windows
// of all K chains before it uses any, so the chains share one memory round
// trip per round (finished chains and window slots past the last predecessor
// re-read a word they are known to reach and ignore it); without it, each
// chain's window is read and used in turn, one round trip per chain and
// window slot. Chains with skip[k] set take no part and leave acc[k] = 0.
template <typename A, int D, int K, bool BATCH = true>
__device__ __forceinline__ void look_back_k(const uint64_t* base, const int64_t (&e0)[K], int64_t step, uint32_t n0,
                                            uint32_t ep4, const Fault& fault, A (&acc)[K],
                                            const bool* skip = nullptr) {
  int64_t e[K];
  uint32_t n[K];
  bool fin[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    acc[k] = zero<A>();
    e[k] = e0[k];
    n[k] = n0;
    fin[k] = n0 == 0 || (skip && skip[k]);
  }
  uint32_t spins = 0;
  for (;;) {
    bool all = true;
#pragma unroll
    for (int k = 0; k < K; ++k) all = all && fin[k];
    if (all) break;
    A v[K][D];
    uint32_t st[K][D];
    typename Carry<A>::Raw raw[K][D];
    // every load of the round first, then their use: a finished chain
    // re-reads its first predecessor and a window slot past the last
    // predecessor re-reads that one; both are ignored
    if constexpr (BATCH) {
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const uint32_t dd = (uint32_t)d < n[k] ? (uint32_t)d : n[k] - 1;
          raw[k][d] = Carry<A>::load(base, fin[k] ? e0[k] : e[k] - (int64_t)dd * step);
        }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int d = 0; d < D; ++d) {
          v[k][d] = zero<A>();
          st[k][d] = Carry<A>::decode(raw[k][d], ep4, v[k][d]);
          if (fin[k] || (uint32_t)d >= n[k]) st[k][d] = 0;
        }
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int d = 0; d < D; ++d) {
          st[k][d] = 0;
          v[k][d] = zero<A>();
          if (!fin[k] && (uint32_t)d < n[k]) st[k][d] = Carry<A>::get(base, e[k] - (int64_t)d * step, ep4, v[k][d]);
        }
    }
    bool progress = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (fin[k]) continue;
      uint32_t used = 0;
      bool done = false;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (!done && used == (uint32_t)d && st[k][d] != 0) {
          acc[k] += v[k][d];
          used = d + 1;
          done = st[k][d] == ST_INC;
        }
      }
      if (used) progress = true;
      e[k] -= (int64_t)used * step;
      n[k] -= used;
      if (done || n[k] == 0) fin[k] = true;
    }
    if (!progress) {
      int64_t ew = e[0];
#pragma unroll
      for (int k = K - 1; k >= 0; --k)
        if (!fin[k]) ew = e[k];
      if (!wait_lead<A>(base, ew, ep4, fault, spins)) break;
    }
  }
}

// ---------------------------------------------------------------------------
// 4-wide global loads and stores of a row chunk [col, col+4).
//
// VecG<T> is 4 consecutive elements aligned to ONE element: code objects for
// amdhsa run with unaligned global access enabled, and the compiler emits a
// single dwordx4 (dword for 1-byte elements, two dwordx4 for 8-byte ones)
// for these, so rows of any pitch take vector accesses. Vec4<T> (naturally
// aligned) is kept for LDS.
// ---------------------------------------------------------------------------
template <typename T> struct Vec4;
template <> struct Vec4<uint8_t> { typedef uint8_t type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<int8_t> { typedef int8_t type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<uint32_t> { typedef uint32_t type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<int32_t> { typedef int32_t type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<double> { typedef double type __attribute__((ext_vector_type(4))); };

template <typename T> struct VecG {
  typedef T type __attribute__((ext_vector_type(4), aligned(sizeof(T))));
};

// Transposed stores write one 16-byte vector per lane (4 elements of 4 bytes
// or 2 of 8 bytes), so the lanes of a row segment cover it without gaps.
template <typename T> constexpr int kVecOut = 16 / (int)sizeof(T);
template <typename T> struct VecOut {
  typedef T type __attribute__((ext_vector_type(kVecOut<T>), aligned(sizeof(T))));
};

// Whole chunks inside the image take one vector load; a chunk that crosses
// the right edge or lies below the image is read element by element (zeros
// outside).
template <typename TSrc, typename TDst, typename A>
__device__ __forceinline__ void load4(const TSrc* __restrict__ row, int64_t col, int64_t W, bool row_ok,
                                      A (&x)[4]) {
  if (row_ok && col + 3 < W) {
    const typename VecG<TSrc>::type q = *(const typename VecG<TSrc>::type*)(row + col);
    x[0] = cvt_in<TSrc, TDst, A>(q.x);
    x[1] = cvt_in<TSrc, TDst, A>(q.y);
    x[2] = cvt_in<TSrc, TDst, A>(q.z);
    x[3] = cvt_in<TSrc, TDst, A>(q.w);
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      x[c] = (row_ok && col + c < W) ? cvt_in<TSrc, TDst, A>(row[col + c]) : zero<A>();
  }
}

// Edge chunks of images at least 4 wide, in two steps so that a caller issues
// every row's load before it uses any (branches between the loads made the
// compiler wait for each load in turn: one memory round trip per row of an
// edge tile). edge_base is the first element of the vector a lane loads: its
// chunk [col, col+4), or the row's last whole chunk [W-4, W) when the chunk
// crosses or lies past the right edge; rows are clamped into the image by the
// caller. edge_pick shifts the loaded elements into place with selects and
// zeroes those outside the image (col + c >= W, or row_ok false).
__device__ __forceinline__ int64_t edge_base(int64_t col, int64_t W) { return col + 3 < W ? col : W - 4; }

template <typename T>
__device__ __forceinline__ T sel4(const T (&e)[4], int i) {
  const T a = (i & 1) ? e[1] : e[0];
  const T b = (i & 1) ? e[3] : e[2];
  return (i & 2) ? b : a;
}

template <typename TSrc, typename TDst, typename A>
__device__ __forceinline__ void edge_pick(const typename VecG<TSrc>::type& q, int64_t col, int64_t base, int64_t W,
                                          bool row_ok, A (&x)[4]) {
  const TSrc e[4] = {q.x, q.y, q.z, q.w};
  const int s = (int)(col - base);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const A val = cvt_in<TSrc, TDst, A>(sel4(e, (c + s) & 3));
    x[c] = (row_ok && col + c < W) ? val : zero<A>();
  }
}

template <typename TDst, typename A>
__device__ __forceinline__ void store4(TDst* __restrict__ row, int64_t col, int64_t W, bool row_ok,
                                       const A (&y)[4]) {
  if (!row_ok) return;
  if (col + 3 < W) {
    typename VecG<TDst>::type q;
    q.x = cvt_out<TDst>(y[0]);
    q.y = cvt_out<TDst>(y[1]);
    q.z = cvt_out<TDst>(y[2]);
    q.w = cvt_out<TDst>(y[3]);
    // the interior path's store policy (nontemporal for 4-byte outputs):
    // batches of 256 f32 1000^2 images 0.460 -> 0.423 ms, u8 -> u32 -7 %, the
    // 512-image mix -1.7 % (plain stores let edge tiles' output displace the
    // carry words and the input in the caches). The loads stay plain: edge
    // tiles re-read their last row for rows below the image, which a
    // nontemporal load would fetch from memory every time (1056^2 +4 %).
    if constexpr (sizeof(TDst) == 8) *(typename VecG<TDst>::type*)(row + col) = q;
    else __builtin_nontemporal_store(q, (typename VecG<TDst>::type*)(row + col));
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (col + c < W) row[col + c] = cvt_out<TDst>(y[c]);
  }
}

// ---------------------------------------------------------------------------
// XOR-swizzled LDS staging of a TW-wide tile for transposes. Element (i, j)
// lives at i*TW + (j ^ (i & M)), M = 31 for 4-byte and 15 for 8-byte values,
// so a wave reading one column (consecutive i) touches distinct banks, while a
// lane's 4 consecutive elements stay inside one aligned 4-element chunk and are
// written with a single vector store after an in-register permutation.
// ---------------------------------------------------------------------------
template <typename A> constexpr int swz_m() { return sizeof(A) == 4 ? 31 : 15; }

template <typename A, int TW>
__device__ __forceinline__ void stage_put4(A* st, int i, int j0, const A (&x)[4]) {
  static_assert(TW % 32 == 0, "staging width");
  const int s = i & swz_m<A>();
  const int p = s & 3;
  typename Vec4<A>::type q;
  // q[d] = x[d ^ p]
  q.x = p == 0 ? x[0] : p == 1 ? x[1] : p == 2 ? x[2] : x[3];
  q.y = p == 0 ? x[1] : p == 1 ? x[0] : p == 2 ? x[3] : x[2];
  q.z = p == 0 ? x[2] : p == 1 ? x[3] : p == 2 ? x[0] : x[1];
  q.w = p == 0 ? x[3] : p == 1 ? x[2] : p == 2 ? x[1] : x[0];
  *(typename Vec4<A>::type*)(st + i * TW + (j0 ^ (s & ~3))) = q;
}

template <typename A, int TW>
__device__ __forceinline__ A stage_get(const A* st, int i, int j) {
  return st[i * TW + (j ^ (i & swz_m<A>()))];
}

// ---------------------------------------------------------------------------
// One image as the tile operators see it.
// ---------------------------------------------------------------------------
struct ImgArgs {
  const void* in;
  void* out;
  int64_t ldi, ldo;   // row pitches, in elements of the input / output type
  uint32_t H, W;      // input rows / columns
  uint32_t nTx, nTy;  // tiles per tile-row / per tile-column
  uint64_t* vcarry;   // vertical carry words of this image's tile 0
  uint64_t* hcarry;   // horizontal carry words of this image's tile 0
  Fault fault;        // where a look-back give-up is recorded
};

// ---------------------------------------------------------------------------
// Workspace header. Carry words follow it. Every launch tags its words with
// an epoch no earlier launch on the same carry area used:
//   - calls on a library-managed area take epochs from a host counter in
//     [1, EPOCH_DEVICE_BASE) and clear the area when the counter wraps;
//   - launches that must not depend on host state (graph-captured calls and
//     the batched queue) first run bump_epoch_kernel, which advances the
//     header's device epoch in [EPOCH_DEVICE_BASE, EPOCH_MAX] and clears the
//     carry area itself when that range wraps; their workgroups read it.
//     The header is only ever lowered together with a clear of the carry
//     area, so while it is below EPOCH_DEVICE_BASE no carry word holds a
//     device epoch;
//   - calls on a caller-provided area clear it first and count from 1.
// ---------------------------------------------------------------------------
struct WsHeader {
  uint32_t epoch;   // device epoch of the latest bumped launch (0: none yet)
  uint32_t pad0[31];
  uint32_t fault;   // nonzero once a look-back gave up (sticky until cleared by the host)
  uint32_t pad1[31];
};
constexpr uint32_t EPOCH_MAX = (1u << 30) - 1;
// Ticket counters of the persistent batch kernel: one per XCD, each on its
// own 128-byte line (stride in 32-bit words).
constexpr uint32_t BATCH_TICKET_STRIDE = 32;
constexpr uint32_t EPOCH_DEVICE_BASE = 1u << 29;
constexpr int64_t WS_HEADER_BYTES = 256;
static_assert(sizeof(WsHeader) <= WS_HEADER_BYTES, "header size");

// Epoch of the running launch: the kernel argument, or the header's device
// epoch when the argument is 0.
__device__ __forceinline__ uint32_t launch_ep4(uint32_t ep_arg, const WsHeader* h) {
  const uint32_t e = ep_arg ? ep_arg : __hip_atomic_load(&h->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return e << 2;
}

}  // namespace satdev
