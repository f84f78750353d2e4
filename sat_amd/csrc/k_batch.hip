// Batched SAT: one persistent launch walks a queue of (image, tile) items over
// images of uneven sizes and mixed dtype pairs. Tickets are global tile
// numbers, image-major and row-major inside an image, so the look-back of a
// tile only ever waits for tiles of its own image handed out earlier.
#include "sat_launch.hpp"

namespace satdev {

enum Kind : uint32_t {
  K_U8_U32 = 0, K_U8_F32, K_U8_F64, K_I8_I32, K_I8_F32, K_I8_F64,
  K_I32_I32, K_I32_F64, K_F32_F32, K_F32_F64, K_F64_F64, K_U32_U32, K_COUNT
};

template <int OUTB> struct BatchCfg;
template <> struct BatchCfg<4> { static constexpr int NW = 4, RPT = 16; using Ref = float; };
template <> struct BatchCfg<8> { static constexpr int NW = 4, RPT = 8; using Ref = double; };

template <int OUTB, typename S, typename D>
__device__ __forceinline__ void run_kind(const ImgArgs& g, uint32_t lt, uint32_t ep4, void* smem) {
  using Cfg = BatchCfg<OUTB>;
  using Op = TileOp<S, D, MODE_SAT, false, false, 0, Cfg::NW, Cfg::RPT>;
  Op::run(g, lt, ep4, *reinterpret_cast<typename Op::Smem*>(smem));
}

template <int OUTB>
__global__ void __launch_bounds__(64 * BatchCfg<OUTB>::NW)
    batch_kernel(WsHeader* hdr, const sathost::BatchItem* __restrict__ items, const uint32_t* __restrict__ tile_img,
                 uint32_t total, uint64_t* carries, uint64_t words_used) {
  using Cfg = BatchCfg<OUTB>;
  using Ref = typename Cfg::Ref;
  using RefOp = TileOp<Ref, Ref, MODE_SAT, false, false, 0, Cfg::NW, Cfg::RPT>;
  __shared__ typename RefOp::Smem sm;
  __shared__ uint32_t s_t, s_ep, s_flag;
  if (threadIdx.x == 0) s_ep = read_epoch(hdr);
  for (;;) {
    if (threadIdx.x == 0) s_t = take_ticket(hdr);
    __syncthreads();
    const uint32_t t = s_t;
    if (t >= total) break;
    const uint32_t ep4 = s_ep << 2;
    const uint32_t img = __builtin_amdgcn_readfirstlane(tile_img[t]);
    const sathost::BatchItem& it = items[img];
    const ImgArgs g = it.g;
    const uint32_t lt = t - it.tile_base;
    const uint32_t kind = __builtin_amdgcn_readfirstlane(it.kind);
    if constexpr (OUTB == 4) {
      switch (kind) {
        case K_U8_U32: run_kind<4, uint8_t, uint32_t>(g, lt, ep4, &sm); break;
        case K_U8_F32: run_kind<4, uint8_t, float>(g, lt, ep4, &sm); break;
        case K_I8_I32: run_kind<4, int8_t, int32_t>(g, lt, ep4, &sm); break;
        case K_I8_F32: run_kind<4, int8_t, float>(g, lt, ep4, &sm); break;
        case K_I32_I32: run_kind<4, int32_t, int32_t>(g, lt, ep4, &sm); break;
        case K_F32_F32: run_kind<4, float, float>(g, lt, ep4, &sm); break;
        case K_U32_U32: run_kind<4, uint32_t, uint32_t>(g, lt, ep4, &sm); break;
        default: break;
      }
    } else {
      switch (kind) {
        case K_U8_F64: run_kind<8, uint8_t, double>(g, lt, ep4, &sm); break;
        case K_I8_F64: run_kind<8, int8_t, double>(g, lt, ep4, &sm); break;
        case K_I32_F64: run_kind<8, int32_t, double>(g, lt, ep4, &sm); break;
        case K_F32_F64: run_kind<8, float, double>(g, lt, ep4, &sm); break;
        case K_F64_F64: run_kind<8, double, double>(g, lt, ep4, &sm); break;
        default: break;
      }
    }
    __syncthreads();
  }
  finish_launch(hdr, carries, words_used, &s_flag);
}

}  // namespace satdev

namespace sathost {

int pair_id(int tin, int tout) {
  using namespace satdev;
  if (tin == SAT_U8 && tout == SAT_U32) return K_U8_U32;
  if (tin == SAT_U8 && tout == SAT_F32) return K_U8_F32;
  if (tin == SAT_U8 && tout == SAT_F64) return K_U8_F64;
  if (tin == SAT_I8 && tout == SAT_I32) return K_I8_I32;
  if (tin == SAT_I8 && tout == SAT_F32) return K_I8_F32;
  if (tin == SAT_I8 && tout == SAT_F64) return K_I8_F64;
  if (tin == SAT_I32 && tout == SAT_I32) return K_I32_I32;
  if (tin == SAT_I32 && tout == SAT_F64) return K_I32_F64;
  if (tin == SAT_F32 && tout == SAT_F32) return K_F32_F32;
  if (tin == SAT_F32 && tout == SAT_F64) return K_F32_F64;
  if (tin == SAT_F64 && tout == SAT_F64) return K_F64_F64;
  if (tin == SAT_U32 && tout == SAT_U32) return K_U32_U32;
  return -1;
}

Geometry batch_geometry(size_t out_size) {
  using namespace satdev;
  if (out_size == 8) return geom_of<TileOp<double, double, MODE_SAT, false, false, 0, BatchCfg<8>::NW, BatchCfg<8>::RPT>>();
  return geom_of<TileOp<float, float, MODE_SAT, false, false, 0, BatchCfg<4>::NW, BatchCfg<4>::RPT>>();
}

int batch_threads(size_t out_size) { return 64 * (out_size == 8 ? satdev::BatchCfg<8>::NW : satdev::BatchCfg<4>::NW); }

int batch_blocks_per_cu(size_t out_size) {
  int n = 0;
  hipError_t e;
  if (out_size == 8)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, satdev::batch_kernel<8>, batch_threads(8), 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, satdev::batch_kernel<4>, batch_threads(4), 0);
  if (e != hipSuccess || n < 1) n = 1;
  return n;
}

sat_status launch_batch(int out_size, const BatchItem* items, const uint32_t* tile_img, uint32_t total_tiles, void* ws,
                        uint64_t* carries, uint64_t words_used, int grid, hipStream_t stream) {
  if (grid < 1) grid = 1;
  if (out_size == 8)
    hipLaunchKernelGGL(satdev::batch_kernel<8>, dim3(grid), dim3(batch_threads(8)), 0, stream,
                       (satdev::WsHeader*)ws, items, tile_img, total_tiles, carries, words_used);
  else
    hipLaunchKernelGGL(satdev::batch_kernel<4>, dim3(grid), dim3(batch_threads(4)), 0, stream,
                       (satdev::WsHeader*)ws, items, tile_img, total_tiles, carries, words_used);
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

}  // namespace sathost
