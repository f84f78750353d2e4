// Batched SAT: one launch walks a queue of (image, tile) items over images of
// uneven sizes and mixed dtype pairs. The host deals whole images to XCDs
// (the largest are laid over all XCDs instead when whole images would load
// the XCDs unevenly)
// and keeps each image's tiles row-major and in increasing queue position, so
// the look-back of a tile only ever waits for tiles of its own image that
// precede it in the queue, on its own XCD for a dealt image.
#include "sat_launch.hpp"

namespace satdev {

enum Kind : uint32_t {
  K_U8_U32 = 0, K_U8_F32, K_U8_F64, K_I8_I32, K_I8_F32, K_I8_F64,
  K_I32_I32, K_I32_F64, K_F32_F32, K_F32_F64, K_F64_F64, K_U32_U32, K_COUNT
};

// Tile shape of the batch kernel: tile configuration CFG (as for the one-pass
// SAT) for OUTB-byte outputs.
template <int OUTB, int CFG> struct BatchCfg {
  static constexpr int NW = sathost::TileCfg<CFG>::NW;
  static constexpr int RPT = OUTB == 4 ? sathost::TileCfg<CFG>::RPT4 : sathost::TileCfg<CFG>::RPT4 / 2;
  static constexpr int WPE = sathost::TileCfg<CFG>::WPE;
  using Ref = typename sathost::cond_t<OUTB == 4, float, double>;
};

template <int OUTB, int CFG, typename S, typename D>
__device__ __forceinline__ void run_kind(const ImgArgs& g, uint32_t lt, uint32_t ep4, void* smem) {
  using Cfg = BatchCfg<OUTB, CFG>;
  using Op = TileOp<S, D, MODE_SAT, false, false, 0, Cfg::NW, Cfg::RPT>;
  Op::run(g, lt, ep4, *reinterpret_cast<typename Op::Smem*>(smem));
}

// ImgArgs as the batch items hold it, with the pointers typed as global
// memory. Pointers read from memory as plain (generic) pointers make the
// compiler address the image through flat instructions, which also count
// against the LDS wait counter (an LDS wait then also waits for outstanding
// image accesses); read through this view, they are known to be global and
// take global instructions, as kernel arguments do. (On the 512-image
// workload the two forms measured the same time.)
#define SAT_GLOBAL __attribute__((address_space(1)))
struct ImgArgsG {
  SAT_GLOBAL const void* in;
  SAT_GLOBAL void* out;
  int64_t ldi, ldo;
  uint32_t H, W, nTx, nTy;
  SAT_GLOBAL uint64_t* vcarry;
  SAT_GLOBAL uint64_t* hcarry;
  SAT_GLOBAL uint32_t* fault_ws;
  SAT_GLOBAL uint32_t* fault_host;
};
static_assert(sizeof(ImgArgsG) == sizeof(ImgArgs) && offsetof(ImgArgsG, vcarry) == offsetof(ImgArgs, vcarry) &&
                  offsetof(ImgArgsG, fault_ws) == offsetof(ImgArgs, fault) &&
                  offsetof(ImgArgsG, fault_host) == offsetof(ImgArgs, fault) + sizeof(void*),
              "ImgArgsG mirrors ImgArgs");

__device__ __forceinline__ ImgArgs load_args(const ImgArgs* p) {
  const ImgArgsG a = *reinterpret_cast<const ImgArgsG*>(p);
  ImgArgs g;
  g.in = (const void*)a.in;
  g.out = (void*)a.out;
  g.ldi = a.ldi;
  g.ldo = a.ldo;
  g.H = a.H;
  g.W = a.W;
  g.nTx = a.nTx;
  g.nTy = a.nTy;
  g.vcarry = (uint64_t*)a.vcarry;
  g.hcarry = (uint64_t*)a.hcarry;
  g.fault.ws = (uint32_t*)a.fault_ws;
  g.fault.host = (uint32_t*)a.fault_host;
  return g;
}

// Runs queue item q: tile queue[2q+1] of image queue[2q] (an empty item
// returns at once). `sm` is the reference operator's shared memory, which
// every operator of the launch's output width fits in.
template <int OUTB, int CFG>
__device__ __forceinline__ void run_item(const sathost::BatchItem* __restrict__ items,
                                         const uint32_t* __restrict__ queue, uint32_t q, uint32_t ep4, void* sm) {
  const uint32_t img = __builtin_amdgcn_readfirstlane(queue[2 * q]);
  if (img == 0xffffffffu) return;
  const uint32_t lt = __builtin_amdgcn_readfirstlane(queue[2 * q + 1]);
  const sathost::BatchItem& it = items[img];
  const ImgArgs g = load_args(&it.g);
  const uint32_t kind = __builtin_amdgcn_readfirstlane(it.kind);
  if constexpr (OUTB == 4) {
    switch (kind) {
      case K_U8_U32: run_kind<4, CFG, uint8_t, uint32_t>(g, lt, ep4, sm); break;
      case K_U8_F32: run_kind<4, CFG, uint8_t, float>(g, lt, ep4, sm); break;
      case K_I8_I32: run_kind<4, CFG, int8_t, int32_t>(g, lt, ep4, sm); break;
      case K_I8_F32: run_kind<4, CFG, int8_t, float>(g, lt, ep4, sm); break;
      case K_I32_I32: run_kind<4, CFG, int32_t, int32_t>(g, lt, ep4, sm); break;
      case K_F32_F32: run_kind<4, CFG, float, float>(g, lt, ep4, sm); break;
      case K_U32_U32: run_kind<4, CFG, uint32_t, uint32_t>(g, lt, ep4, sm); break;
      default: break;
    }
  } else {
    switch (kind) {
      case K_U8_F64: run_kind<8, CFG, uint8_t, double>(g, lt, ep4, sm); break;
      case K_I8_F64: run_kind<8, CFG, int8_t, double>(g, lt, ep4, sm); break;
      case K_I32_F64: run_kind<8, CFG, int32_t, double>(g, lt, ep4, sm); break;
      case K_F32_F64: run_kind<8, CFG, float, double>(g, lt, ep4, sm); break;
      case K_F64_F64: run_kind<8, CFG, double, double>(g, lt, ep4, sm); break;
      default: break;
    }
  }
}

// One workgroup per queue item: item b = blockIdx.x. The launch epoch is the
// workspace's device epoch, advanced by bump_epoch_kernel before every
// launch, so the same launch can be replayed from a captured graph.
template <int OUTB, int CFG>
__global__ void __launch_bounds__((64 * BatchCfg<OUTB, CFG>::NW))
    __attribute__((amdgpu_waves_per_eu(BatchCfg<OUTB, CFG>::WPE)))
    batch_kernel(const WsHeader* hdr, const sathost::BatchItem* __restrict__ items,
                 const uint32_t* __restrict__ queue) {
  using Cfg = BatchCfg<OUTB, CFG>;
  using RefOp = TileOp<typename Cfg::Ref, typename Cfg::Ref, MODE_SAT, false, false, 0, Cfg::NW, Cfg::RPT>;
  __shared__ typename RefOp::Smem sm;
  run_item<OUTB, CFG>(items, queue, blockIdx.x, launch_ep4(0, hdr), &sm);
}

// Persistent form: nx * G workgroups, each serving the XCD x = blockIdx.x % nx
// it is dealt to. It draws tickets k = 0, 1, ... from that XCD's counter and
// runs queue item k * nx + x (the XCD's k-th item) until k reaches the list
// length L; the next ticket is drawn while the current item runs. Tickets are
// drawn in increasing order and every item waits only on earlier items of its
// own list, which workgroups that are running have drawn: the walk progresses
// whatever the number of resident workgroups. The counters are zeroed by
// bump_epoch_kernel before every launch.
// The item loop around the inlined tile operators needs more registers than
// the one-item kernel: at its configuration's waves per SIMD it spilled 60-135
// registers and took 1.75 ms on the 512-image workload; with a register budget
// of 2 waves per SIMD it spills none and takes 1.14 ms (the one-item form:
// 0.91 ms). The f32 form of configuration 1 takes 212 VGPRs in the loop
// against at most 128 for one item: values that depend only on the lane
// (LDS and row offsets of every dtype pair's inlined operator) are hoisted
// out of the loop and stay live across items. A call per item instead of the
// inlined operators would save and restore the callee-saved registers it
// uses through scratch memory on every tile.
template <int OUTB, int CFG>
__global__ void __launch_bounds__((64 * BatchCfg<OUTB, CFG>::NW))
    __attribute__((amdgpu_waves_per_eu(BatchCfg<OUTB, CFG>::WPE < 2 ? BatchCfg<OUTB, CFG>::WPE : 2)))
    batch_persistent_kernel(const WsHeader* hdr, const sathost::BatchItem* __restrict__ items,
                            const uint32_t* __restrict__ queue, uint32_t* tickets, uint32_t L, uint32_t nx) {
  using Cfg = BatchCfg<OUTB, CFG>;
  using RefOp = TileOp<typename Cfg::Ref, typename Cfg::Ref, MODE_SAT, false, false, 0, Cfg::NW, Cfg::RPT>;
  __shared__ typename RefOp::Smem sm;
  __shared__ uint32_t s_k[2];
  const uint32_t ep4 = launch_ep4(0, hdr);
  const uint32_t x = blockIdx.x % nx;
  uint32_t* tk = tickets + BATCH_TICKET_STRIDE * x;
  if (threadIdx.x == 0) s_k[0] = __hip_atomic_fetch_add(tk, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  uint32_t k = s_k[0];
  int cur = 0;
  while (k < L) {
    uint32_t next = 0;
    if (threadIdx.x == 0) next = __hip_atomic_fetch_add(tk, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    run_item<OUTB, CFG>(items, queue, k * nx + x, ep4, &sm);
    // every wave is done with the item's shared memory before the next item
    // writes it; the next ticket becomes visible to all waves
    if (threadIdx.x == 0) s_k[cur ^ 1] = next;
    __syncthreads();
    cur ^= 1;
    k = s_k[cur];
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

template <int OUTB, int CFG>
static Geometry batch_geom() {
  using Ref = typename satdev::BatchCfg<OUTB, CFG>::Ref;
  return geom_of<satdev::TileOp<Ref, Ref, satdev::MODE_SAT, false, false, 0, satdev::BatchCfg<OUTB, CFG>::NW,
                                satdev::BatchCfg<OUTB, CFG>::RPT>>();
}

Geometry batch_geometry(size_t out_size, int cfg) {
  if (out_size == 8)
    return cfg == 1 ? batch_geom<8, 1>() : cfg == 2 ? batch_geom<8, 2>() : cfg == 3 ? batch_geom<8, 3>()
         : cfg == 4 ? batch_geom<8, 4>() : batch_geom<8, 0>();
  return cfg == 1 ? batch_geom<4, 1>() : cfg == 2 ? batch_geom<4, 2>() : cfg == 3 ? batch_geom<4, 3>()
       : cfg == 4 ? batch_geom<4, 4>() : batch_geom<4, 0>();
}

template <int OUTB, int CFG>
static void launch_batch_cfg(const BatchItem* items, const uint32_t* queue, uint32_t blocks, void* ws,
                             const BatchPersist& pers, hipStream_t stream) {
  constexpr int threads = 64 * satdev::BatchCfg<OUTB, CFG>::NW;
  if (pers.tickets)
    hipLaunchKernelGGL((satdev::batch_persistent_kernel<OUTB, CFG>), dim3(pers.grid), dim3(threads), 0, stream,
                       (const satdev::WsHeader*)ws, items, queue, pers.tickets, pers.items_per_xcd, pers.nx);
  else
    hipLaunchKernelGGL((satdev::batch_kernel<OUTB, CFG>), dim3(blocks), dim3(threads), 0, stream,
                       (const satdev::WsHeader*)ws, items, queue);
}

template <int OUTB, int CFG>
static int resident_cfg() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, satdev::batch_persistent_kernel<OUTB, CFG>,
                                                   64 * satdev::BatchCfg<OUTB, CFG>::NW, 0) != hipSuccess)
    return 0;
  return n;
}

int batch_resident_per_cu(int out_size, int cfg) {
  if (out_size == 8)
    return cfg == 1 ? resident_cfg<8, 1>() : cfg == 2 ? resident_cfg<8, 2>() : cfg == 3 ? resident_cfg<8, 3>()
         : cfg == 4 ? resident_cfg<8, 4>() : resident_cfg<8, 0>();
  return cfg == 1 ? resident_cfg<4, 1>() : cfg == 2 ? resident_cfg<4, 2>() : cfg == 3 ? resident_cfg<4, 3>()
       : cfg == 4 ? resident_cfg<4, 4>() : resident_cfg<4, 0>();
}

sat_status launch_batch(int out_size, int cfg, const BatchItem* items, const uint32_t* queue, uint32_t blocks,
                        void* ws, const BatchPersist& pers, hipStream_t stream) {
  if (blocks == 0) return SAT_OK;
  if (out_size == 8) {
    if (cfg == 1) launch_batch_cfg<8, 1>(items, queue, blocks, ws, pers, stream);
    else if (cfg == 2) launch_batch_cfg<8, 2>(items, queue, blocks, ws, pers, stream);
    else if (cfg == 3) launch_batch_cfg<8, 3>(items, queue, blocks, ws, pers, stream);
    else if (cfg == 4) launch_batch_cfg<8, 4>(items, queue, blocks, ws, pers, stream);
    else launch_batch_cfg<8, 0>(items, queue, blocks, ws, pers, stream);
  } else {
    if (cfg == 1) launch_batch_cfg<4, 1>(items, queue, blocks, ws, pers, stream);
    else if (cfg == 2) launch_batch_cfg<4, 2>(items, queue, blocks, ws, pers, stream);
    else if (cfg == 3) launch_batch_cfg<4, 3>(items, queue, blocks, ws, pers, stream);
    else if (cfg == 4) launch_batch_cfg<4, 4>(items, queue, blocks, ws, pers, stream);
    else launch_batch_cfg<4, 0>(items, queue, blocks, ws, pers, stream);
  }
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

}  // namespace sathost
