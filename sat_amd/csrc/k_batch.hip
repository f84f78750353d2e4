// Batched SAT: one launch walks a queue of (image, tile) items over images of
// uneven sizes and mixed dtype pairs. Queue positions are image-major and
// row-major inside an image, so the look-back of a tile only ever waits for
// tiles of its own image that precede it in the queue.
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

// One workgroup per queue item: item t = blockIdx.x is tile t - tile_base of
// image tile_img[t]. The launch epoch is the workspace's device epoch,
// advanced by bump_epoch_kernel before every launch, so the same launch can
// be replayed from a captured graph.
template <int OUTB>
__global__ void __launch_bounds__(64 * BatchCfg<OUTB>::NW) __attribute__((amdgpu_waves_per_eu(4)))
    batch_kernel(const WsHeader* hdr, const sathost::BatchItem* __restrict__ items,
                 const uint32_t* __restrict__ tile_img) {
  using Cfg = BatchCfg<OUTB>;
  using Ref = typename Cfg::Ref;
  using RefOp = TileOp<Ref, Ref, MODE_SAT, false, false, 0, Cfg::NW, Cfg::RPT>;
  __shared__ typename RefOp::Smem sm;
  const uint32_t ep4 = launch_ep4(0, hdr);
  const uint32_t t = blockIdx.x;
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

sat_status launch_batch(int out_size, const BatchItem* items, const uint32_t* tile_img, uint32_t total_tiles,
                        void* ws, hipStream_t stream) {
  if (total_tiles == 0) return SAT_OK;
  if (out_size == 8)
    hipLaunchKernelGGL(satdev::batch_kernel<8>, dim3(total_tiles), dim3(batch_threads(8)), 0, stream,
                       (const satdev::WsHeader*)ws, items, tile_img);
  else
    hipLaunchKernelGGL(satdev::batch_kernel<4>, dim3(total_tiles), dim3(batch_threads(4)), 0, stream,
                       (const satdev::WsHeader*)ws, items, tile_img);
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

}  // namespace sathost
