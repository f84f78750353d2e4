// One-pass SAT kernels (forward and reverse) in tile configuration 0, the
// configuration dispatch, the fill / epoch kernels and the family geometry table.
#include <algorithm>

#include "sat_launch.hpp"

namespace satdev {

__global__ void __launch_bounds__(256) clear_kernel(uint64_t* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Streaming copy, the bandwidth reference the SAT kernels are measured
// against: each thread moves COPY_UNROLL 16-byte chunks, every load of a wave
// covering 1 KiB of consecutive bytes, all loads issued before the first
// store; loads and stores are nontemporal (the fastest of the copy shapes
// measured on buffers the MALL does not hold). Chunks past `n16` are
// skipped; `tail` bytes after the last whole chunk are copied by the first
// workgroup.
constexpr int COPY_UNROLL = 4;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) stream_copy_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                          uint64_t n16, const uint8_t* __restrict__ src_tail,
                                                          uint8_t* __restrict__ dst_tail, uint32_t tail) {
  const uint64_t base = (uint64_t)blockIdx.x * (256 * COPY_UNROLL) + threadIdx.x;
  u32x4 v[COPY_UNROLL];
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; ++u) {
    const uint64_t i = base + (uint64_t)u * 256;
    if (i < n16) v[u] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; ++u) {
    const uint64_t i = base + (uint64_t)u * 256;
    if (i < n16) __builtin_nontemporal_store(v[u], dst + i);
  }
  if (blockIdx.x == 0 && threadIdx.x < tail) dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

// Advances the device epoch of a workspace for the launch that follows on the
// same stream. A header below the device range (a fresh or cleared area, or
// one only host-epoch launches used) starts it at EPOCH_DEVICE_BASE with no
// clear: such an area holds no word tagged with a device epoch. When the
// range wraps, `words` carry words are cleared first, so no word of an
// earlier pass through the range can pass for a new launch's.
__global__ void __launch_bounds__(256) bump_epoch_kernel(WsHeader* hdr, uint64_t* carries, uint64_t words,
                                                         uint32_t* tickets, uint32_t ntickets) {
  if (threadIdx.x < ntickets) tickets[threadIdx.x * BATCH_TICKET_STRIDE] = 0;
  __shared__ uint32_t s_next, s_wrap;
  if (threadIdx.x == 0) {
    const uint32_t e = hdr->epoch;
    s_wrap = e >= EPOCH_MAX;
    s_next = (e < EPOCH_DEVICE_BASE || e >= EPOCH_MAX) ? EPOCH_DEVICE_BASE : e + 1;
  }
  __syncthreads();
  const uint32_t next = s_next;
  if (s_wrap)
    for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) carries[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&hdr->epoch, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Zero border of an OpenCV-style (H+1) x (W+1) table: row 0 (W+1 elements)
// and column 0 of rows 1..H, `planes` tables `plane` elements apart.
template <typename T>
__global__ void __launch_bounds__(256) border_kernel(T* out, int64_t pitch, int64_t plane, uint32_t width,
                                                     uint32_t height) {
  T* o = out + (int64_t)blockIdx.y * plane;
  const uint64_t n = (uint64_t)width + 1 + height;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i <= width) o[i] = T(0);
    else o[(int64_t)(i - width) * pitch] = T(0);
  }
}

}  // namespace satdev

namespace sathost {

sat_status launch_border(size_t out_size, void* out, int64_t pitch, int64_t plane, uint32_t width, uint32_t height,
                         uint32_t planes, hipStream_t stream) {
  const uint64_t n = (uint64_t)width + 1 + height;
  uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 1024);
  if (planes == 0) return SAT_OK;
  if (out_size == 8)
    hipLaunchKernelGGL(satdev::border_kernel<uint64_t>, dim3(blocks, planes), dim3(256), 0, stream, (uint64_t*)out, pitch,
                       plane, width, height);
  else
    hipLaunchKernelGGL(satdev::border_kernel<uint32_t>, dim3(blocks, planes), dim3(256), 0, stream, (uint32_t*)out, pitch,
                       plane, width, height);
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

using namespace satdev;

sat_status launch_bump_epoch(void* ws, uint64_t carry_words, hipStream_t stream, uint32_t* tickets,
                             uint32_t ntickets) {
  if (ntickets > 256) return SAT_ERR_INVALID;
  hipLaunchKernelGGL(bump_epoch_kernel, dim3(1), dim3(256), 0, stream, (WsHeader*)ws,
                     (uint64_t*)((char*)ws + WS_HEADER_BYTES), carry_words, tickets, ntickets);
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

sat_status launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t stream) {
  const uint64_t n16 = bytes / 16;
  const uint32_t tail = (uint32_t)(bytes % 16);
  const uint64_t per_block = 256 * COPY_UNROLL;
  const uint64_t blocks = std::max<uint64_t>(1, (n16 + per_block - 1) / per_block);
  if (blocks > 0x7fffffffu) return SAT_ERR_INVALID;
  hipLaunchKernelGGL(stream_copy_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, (const u32x4*)src,
                     (u32x4*)dst, n16, (const uint8_t*)src + n16 * 16, (uint8_t*)dst + n16 * 16, tail);
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

sat_status launch_fam_sat(const Launch& L, void* ws, size_t b, hipStream_t st) {
  switch (L.cfg) {
    case 1: return launch_fam_sat_c1(L, ws, b, st);
    case 2: return launch_fam_sat_c2(L, ws, b, st);
    case 3: return launch_fam_sat_c3(L, ws, b, st);
    case 4: return launch_fam_sat_c4(L, ws, b, st);
    case 5: return launch_fam_sat_c5(L, ws, b, st);
    case 6: return launch_fam_sat_c6(L, ws, b, st);
    case 7: return launch_fam_sat_c7(L, ws, b, st);
    default: return launch_sat_cfg<0>(L, ws, b, st);
  }
}

sat_status launch_fam_sat_rev(const Launch& L, void* ws, size_t b, hipStream_t st) {
  switch (L.cfg) {
    case 1: return launch_fam_sat_rev_c1(L, ws, b, st);
    case 2: return launch_fam_sat_rev_c2(L, ws, b, st);
    case 3: return launch_fam_sat_rev_c3(L, ws, b, st);
    case 4: return launch_fam_sat_rev_c4(L, ws, b, st);
    case 5: return launch_fam_sat_rev_c5(L, ws, b, st);
    case 6: return launch_fam_sat_rev_c6(L, ws, b, st);
    case 7: return launch_fam_sat_rev_c7(L, ws, b, st);
    default: return launch_sat_rev_cfg<0>(L, ws, b, st);
  }
}

Geometry geometry(int family, size_t out_size, int cfg, size_t in_size) {
  const bool b8 = out_size == 8;
  switch (family) {
    case FAM_SAT:
    case FAM_SAT_REV:
      switch (cfg) {
        case 1: return sat_geometry_cfg<1>(out_size, in_size);
        case 2: return sat_geometry_cfg<2>(out_size, in_size);
        case 3: return sat_geometry_cfg<3>(out_size, in_size);
        case 4: return sat_geometry_cfg<4>(out_size, in_size);
        case 5: return sat_geometry_cfg<5>(out_size, in_size);
        case 6: return sat_geometry_cfg<6>(out_size, in_size);
        case 7: return sat_geometry_cfg<7>(out_size, in_size);
        default: return sat_geometry_cfg<0>(out_size, in_size);
      }
    case FAM_ROWS:
      return b8 ? geom_of<TileOp<double, double, MODE_ROW, false, false, 0, kNW, kRPT<double>>>()
                : geom_of<TileOp<float, float, MODE_ROW, false, false, 0, kNW, kRPT<float>>>();
    case FAM_ROWS_T:
      return b8 ? geom_of<TileOp<double, double, MODE_ROW, false, true, 0, kNW_T, kRPT_T<double>>>()
                : geom_of<TileOp<float, float, MODE_ROW, false, true, 0, kNW_T, kRPT_T<float>>>();
    case FAM_COLS:
      return b8 ? geom_of<TileOp<double, double, MODE_COL, false, false, 0, kNW, kRPT<double>>>()
                : geom_of<TileOp<float, float, MODE_COL, false, false, 0, kNW, kRPT<float>>>();
    case FAM_SERIAL_T:
      return b8 ? geom_of<SerialTileOp<double, double, kNW_S, kSerialCW<double>>>()
                : geom_of<SerialTileOp<float, float, kNW_S, kSerialCW<float>>>();
    default: {
      Geometry g{0, 0, 0, 0, 0, 0};
      return g;
    }
  }
}

}  // namespace sathost
