// Host-side internal interface between the C ABI (sat_api.cpp) and the
// translation units that instantiate the kernels (k_*.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/sat/sat.h"
#include "sat_device.hpp"

namespace sathost {

// Pass / kernel families.
enum Family : int {
  FAM_SAT = 0,       // one-pass 2-D SAT (TileOp MODE_SAT)
  FAM_SAT_REV = 1,   // one-pass reverse SAT (TileOp MODE_SAT, mirrored)
  FAM_ROWS = 2,      // row scan, same layout (TileOp MODE_ROW)
  FAM_COLS = 3,      // column scan (TileOp MODE_COL)
  FAM_ROWS_T = 4,    // row scan + transposed store (TileOp MODE_ROW + TOUT)
  FAM_SERIAL_T = 5,  // transpose + serial row scan (SerialTileOp)
  FAM_COUNT = 6
};

// Tile geometry of a family for an output element size (bytes).
struct Geometry {
  int th, tw;            // tile rows / columns (of the input)
  int64_t vwords, hwords;  // carry words per tile
  int64_t gwords = 0;      // second-level vertical carry words per run and tile column (0: none)
  int threads;
};
// Tile rows per run of the second vertical carry level (satdev::VGroups) for
// an nTx x nTy tile grid, 0 for one level. An image of one tile column has a
// single vertical chain as long as the image; from 64 tile rows on, runs of
// 16 let a look-back skip a run per step: f32 65536 x 256 (512 tiles of 128
// rows) 31.7 -> 29.3 us, 16384 x 256 17.4 -> 16.6 us, u32 65536 x 256 31.7 ->
// 29.5 us (runs of 8 or 32 and other windows: equal or slower). With two tile
// columns it measured slower (f32 65536 x 512: 56.5 -> 60.0 us). Launches
// with more tiles than resident workgroups do not use it (launch_op).
inline uint32_t vgroup_len(uint64_t nTx, uint64_t nTy) { return (nTx == 1 && nTy >= 64) ? 16u : 0u; }
// Tile configurations of the one-pass families (FAM_SAT, FAM_SAT_REV), cfg in
// [0, TILE_CFGS), and of the batch kernel, cfg in [0, BATCH_CFGS); other
// families have one shape. The one-pass geometry of configuration 5 depends
// on the input width.
constexpr int TILE_CFGS = 8;
constexpr int BATCH_CFGS = 5;
Geometry geometry(int family, size_t out_size, int cfg = 0, size_t in_size = 4);

// One launch of a family over `planes` images of identical shape.
struct Launch {
  int family;
  int scan;
  int tin, tout;
  const void* in;
  void* out;
  int64_t ldi, ldo;
  int64_t in_plane, out_plane;  // plane strides (elements)
  uint32_t width, height, planes;
  uint32_t epoch;  // launch epoch; 0 = read the workspace header's device epoch
  int cfg = 0;     // tile configuration (one-pass families)
  bool may_alloc = true;  // false: the launch must not allocate (no first-use pinned host flag)
};

size_t carry_bytes(const Launch& L);
// XCDs of the current device (workgroup b runs on XCD b % count); 1 if unknown
int device_xcd_count();
sat_status launch_family(const Launch& L, void* ws, size_t ws_bytes, hipStream_t stream);

// Batched persistent kernel (k_batch.hip).
struct BatchItem {
  satdev::ImgArgs g;
  uint32_t kind;       // dtype pair id (pair_id)
  uint32_t tile_base;  // first global tile of this image
};
int pair_id(int tin, int tout);  // -1 if unsupported
// The persistent form of the batch launch (tickets null: one workgroup per
// queue item instead): grid workgroups draw tickets from nx counters.
struct BatchPersist {
  uint32_t* tickets = nullptr;
  uint32_t grid = 0;
  uint32_t items_per_xcd = 0;
  uint32_t nx = 1;
};
// queue: (image, tile) pairs, blocks = items_per_xcd * nx of them, item k of
// XCD x at k * nx + x (image 0xffffffff: empty)
sat_status launch_batch(int out_size, int cfg, const BatchItem* items, const uint32_t* queue, uint32_t blocks,
                        void* ws, const BatchPersist& pers, hipStream_t stream);
Geometry batch_geometry(size_t out_size, int cfg);
// resident workgroups per CU of the persistent batch kernel (0 if unknown)
int batch_resident_per_cu(int out_size, int cfg);

// Box filter (k_box.hip).
sat_status launch_box_sums(int dtype, const void* sat, int64_t pitch, int64_t plane_stride, uint32_t width,
                           uint32_t height, uint32_t planes, const int32_t* boxes, uint32_t nbox, void* out,
                           hipStream_t stream);
sat_status launch_box_scatter(int dtype, const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_sat,
                              int64_t pitch, int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes,
                              hipStream_t stream);
sat_status launch_box_grad(int dtype, const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_x,
                           int64_t pitch, int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes,
                           void* ws, size_t ws_bytes, void* bins, uint32_t epoch, hipStream_t stream);
// impulse bins of the fused box gradient (0: that shape scans the box list)
size_t box_grad_bin_bytes(int dtype, uint32_t width, uint32_t height, uint32_t planes, uint32_t nbox);
size_t box_grad_carry_bytes(int dtype, uint32_t width, uint32_t height, uint32_t planes);

// Zero `bytes` (rounded down to 8) at `p` with a fill kernel (k_sat.hip).
sat_status launch_clear(void* p, size_t bytes, hipStream_t stream);
sat_status launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t stream);
// Zero row 0 and column 0 of `planes` (H+1) x (W+1) tables (k_sat.hip).
sat_status launch_border(size_t out_size, void* out, int64_t pitch, int64_t plane, uint32_t width, uint32_t height,
                         uint32_t planes, hipStream_t stream);
// Advance the header's device epoch (clearing `carry_words` on wrap) and zero
// `ntickets` ticket counters (stride satdev::BATCH_TICKET_STRIDE), k_sat.hip.
sat_status launch_bump_epoch(void* ws, uint64_t carry_words, hipStream_t stream, uint32_t* tickets = nullptr,
                             uint32_t ntickets = 0);

// Device address of the current device's sticky give-up flag in pinned host
// memory (the host reads it without synchronising); null if it could not be
// mapped, which only loses the host-visible copy of a give-up. The pinned
// memory is allocated on first use; with may_alloc false, null until then.
uint32_t* host_fault_word(bool may_alloc = true);

// Misc.
int device_cu_count();
size_t dtype_size(int dt);

}  // namespace sathost
