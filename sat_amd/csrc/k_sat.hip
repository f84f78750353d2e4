// One-pass SAT kernels (forward and reverse) and the family geometry table.
#include "sat_launch.hpp"

namespace satdev {

__global__ void __launch_bounds__(256) clear_kernel(uint64_t* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Advances the device epoch of a workspace for the launch that follows on the
// same stream; on wrap-around it clears `words` carry words first.
__global__ void __launch_bounds__(256) bump_epoch_kernel(WsHeader* hdr, uint64_t* carries, uint64_t words) {
  __shared__ uint32_t s_next;
  if (threadIdx.x == 0) {
    const uint32_t e = hdr->epoch;
    s_next = (e < EPOCH_DEVICE_BASE || e >= EPOCH_MAX) ? EPOCH_DEVICE_BASE : e + 1;
  }
  __syncthreads();
  const uint32_t next = s_next;
  if (next == EPOCH_DEVICE_BASE)
    for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) carries[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&hdr->epoch, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace satdev

namespace sathost {

using namespace satdev;

sat_status launch_bump_epoch(void* ws, uint64_t carry_words, hipStream_t stream) {
  hipLaunchKernelGGL(bump_epoch_kernel, dim3(1), dim3(256), 0, stream, (WsHeader*)ws,
                     (uint64_t*)((char*)ws + WS_HEADER_BYTES), carry_words);
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

sat_status launch_clear(void* p, size_t bytes, hipStream_t stream) {
  const uint64_t n = bytes / 8;
  if (n == 0) return SAT_OK;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(clear_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, (uint64_t*)p, n);
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

template <typename S, typename D, int SC>
struct FamSat {
  using Op = TileOp<S, D, MODE_SAT, false, false, SC, kNW, kRPT<D>>;
  static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) { return launch_op<Op>(L, ws, b, st); }
};

template <typename S, typename D, int SC>
struct FamSatRev {
  using Op = TileOp<S, D, MODE_SAT, true, false, SC, kNW, kRPT<D>>;
  static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) { return launch_op<Op>(L, ws, b, st); }
};

sat_status launch_fam_sat(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return dispatch_pairs<FamSat>(L, ws, b, st);
}

// The adjoint is defined for same-type pairs.
sat_status launch_fam_sat_rev(const Launch& L, void* ws, size_t b, hipStream_t st) {
  if (L.tin != L.tout) return SAT_ERR_DTYPE;
  switch (L.tin) {
    case SAT_F32: return ScanPick<true>::go<FamSatRev, float, float>(L, ws, b, st);
    case SAT_F64: return ScanPick<true>::go<FamSatRev, double, double>(L, ws, b, st);
    case SAT_U32: return ScanPick<false>::go<FamSatRev, uint32_t, uint32_t>(L, ws, b, st);
    case SAT_I32: return ScanPick<false>::go<FamSatRev, int32_t, int32_t>(L, ws, b, st);
    default: return SAT_ERR_DTYPE;
  }
}

Geometry geometry(int family, size_t out_size) {
  const bool b8 = out_size == 8;
  switch (family) {
    case FAM_SAT:
    case FAM_SAT_REV:
      return b8 ? geom_of<TileOp<double, double, MODE_SAT, false, false, 0, kNW, kRPT<double>>>()
                : geom_of<TileOp<float, float, MODE_SAT, false, false, 0, kNW, kRPT<float>>>();
    case FAM_ROWS:
      return b8 ? geom_of<TileOp<double, double, MODE_ROW, false, false, 0, kNW, kRPT<double>>>()
                : geom_of<TileOp<float, float, MODE_ROW, false, false, 0, kNW, kRPT<float>>>();
    case FAM_ROWS_T:
      return b8 ? geom_of<TileOp<double, double, MODE_ROW, false, true, 0, kNW, kRPT<double>>>()
                : geom_of<TileOp<float, float, MODE_ROW, false, true, 0, kNW, kRPT<float>>>();
    case FAM_COLS:
      return b8 ? geom_of<TileOp<double, double, MODE_COL, false, false, 0, kNW, kRPT<double>>>()
                : geom_of<TileOp<float, float, MODE_COL, false, false, 0, kNW, kRPT<float>>>();
    case FAM_SERIAL_T:
      return b8 ? geom_of<SerialTileOp<double, double, kNW, kSerialCW<double>>>()
                : geom_of<SerialTileOp<float, float, kNW, kSerialCW<float>>>();
    default: {
      Geometry g{0, 0, 0, 0, 0};
      return g;
    }
  }
}

}  // namespace sathost
