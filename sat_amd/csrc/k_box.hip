// Box-filter readout on summed-area tables and its fused gradient.
#include "sat_launch.hpp"

namespace satdev {

// A box {y0, x0, y1, x1} (inclusive) is valid when it lies inside the image
// and is not empty. Invalid boxes read nothing: their sum is NaN and they
// contribute no gradient.
__device__ __forceinline__ bool box_ok(int4 b, uint32_t W, uint32_t H) {
  return b.x >= 0 && b.y >= 0 && b.x <= b.z && b.y <= b.w && (uint32_t)b.z < H && (uint32_t)b.w < W;
}

template <typename T>
__global__ void __launch_bounds__(256) box_sums_kernel(const T* __restrict__ sat, int64_t pitch, int64_t plane_stride,
                                                       uint32_t W, uint32_t H, uint64_t count, uint32_t nbox,
                                                       const int4* __restrict__ boxes, T* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t p = i / nbox;
  const int4 b = boxes[i];
  if (!box_ok(b, W, H)) {
    out[i] = __builtin_nan("");
    return;
  }
  const T* s = sat + p * plane_stride;
  const T s11 = s[(int64_t)b.z * pitch + b.w];
  const T s01 = b.x > 0 ? s[(int64_t)(b.x - 1) * pitch + b.w] : T(0);
  const T s10 = b.y > 0 ? s[(int64_t)b.z * pitch + (b.y - 1)] : T(0);
  const T s00 = (b.x > 0 && b.y > 0) ? s[(int64_t)(b.x - 1) * pitch + (b.y - 1)] : T(0);
  out[i] = ((s11 - s01) - s10) + s00;
}

template <typename T>
__global__ void __launch_bounds__(256) box_scatter_kernel(const int4* __restrict__ boxes, const T* __restrict__ gout,
                                                          uint64_t count, uint32_t nbox, T* gs, int64_t pitch,
                                                          int64_t plane_stride, uint32_t W, uint32_t H) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int4 b = boxes[i];
  if (!box_ok(b, W, H)) return;
  const T g = gout[i];
  T* s = gs + (i / nbox) * plane_stride;
  atomicAdd(&s[(int64_t)b.z * pitch + b.w], g);
  if (b.x > 0) atomicAdd(&s[(int64_t)(b.x - 1) * pitch + b.w], -g);
  if (b.y > 0) atomicAdd(&s[(int64_t)b.z * pitch + (b.y - 1)], -g);
  if (b.x > 0 && b.y > 0) atomicAdd(&s[(int64_t)(b.x - 1) * pitch + (b.y - 1)], g);
}

template <typename T> struct GradCfg;
template <> struct GradCfg<float> { static constexpr int NW = 4, RPT = 8; };
template <> struct GradCfg<double> { static constexpr int NW = 4, RPT = 4; };

// Reverse-SAT tile operator of the gradient. It accumulates and carries in
// f64 for both gradient types and rounds once on store: an f32 gradient sums
// many corner impulses of either sign, and f32 partial sums would leave
// errors near 1e-6 of its norm.
template <typename T>
using GradOp = TileOp<T, T, MODE_SAT, true, false, 0, GradCfg<T>::NW, GradCfg<T>::RPT, double>;

// Corner impulses of the boxes binned by the tile that holds them: one
// workgroup per plane counts the impulses per tile in LDS, scans the counts
// into offsets and places (position in tile, value) pairs, so that a gradient
// tile reads only its own impulses. off[p*(tpp+1) + i] is the first entry of
// physical tile i of plane p (entries of plane p start at p*4*nbox).
constexpr int BIN_THREADS = 1024;
constexpr uint32_t BIN_MAX_TILES = 8191;  // LDS counters per plane
static_assert(BIN_THREADS % 64 == 0 && BIN_THREADS / 64 <= 64, "one wave scans the wave totals");

template <typename T>
__global__ void __launch_bounds__(BIN_THREADS)
    box_bin_kernel(const int4* __restrict__ boxes, const T* __restrict__ gout, uint32_t nbox, uint32_t W, uint32_t H,
                   uint32_t nTx, uint32_t tpp, uint32_t* __restrict__ off, uint32_t* __restrict__ pos,
                   T* __restrict__ val) {
  using Op = GradOp<T>;
  constexpr int TH = Op::TH, TW = Op::TW;
  __shared__ uint32_t cnt[BIN_MAX_TILES + 1];
  __shared__ uint32_t wsum[BIN_THREADS / 64];
  const uint32_t p = blockIdx.x, tid = threadIdx.x;
  const int4* pb = boxes + (uint64_t)p * nbox;
  const T* pg = gout + (uint64_t)p * nbox;
  for (uint32_t i = tid; i <= tpp; i += BIN_THREADS) cnt[i] = 0;
  __syncthreads();
  // pass 1: impulses per tile
  for (uint32_t b = tid; b < nbox; b += BIN_THREADS) {
    const int4 bx = pb[b];
    if (!box_ok(bx, W, H)) continue;
    const int ys[2] = {bx.z, bx.x - 1}, xs[2] = {bx.w, bx.y - 1};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        if (ys[a] >= 0 && xs[c] >= 0) atomicAdd(&cnt[(uint32_t)ys[a] / TH * nTx + (uint32_t)xs[c] / TW], 1u);
  }
  __syncthreads();
  // exclusive scan of cnt[0..tpp): each thread folds a run, then the runs
  // are scanned across the block (each wave its 64 runs, one wave the wave
  // totals)
  const uint32_t per = (tpp + BIN_THREADS - 1) / BIN_THREADS, i0 = tid * per;
  uint32_t run = 0;
  for (uint32_t i = i0; i < i0 + per && i < tpp; ++i) run += cnt[i];
  const uint32_t lane = tid & 63, wv = tid >> 6;
  const uint32_t inc = wave_scan_ks(run);
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  if (wv == 0) {
    const uint32_t t = lane < BIN_THREADS / 64 ? wsum[lane] : 0u;
    const uint32_t ti = wave_scan_ks(t);
    if (lane < BIN_THREADS / 64) wsum[lane] = ti - t;
  }
  __syncthreads();
  uint32_t acc = wsum[wv] + inc - run;
  for (uint32_t i = i0; i < i0 + per && i < tpp; ++i) {
    const uint32_t v = cnt[i];
    cnt[i] = acc;
    off[(uint64_t)p * (tpp + 1) + i] = acc;
    acc += v;
  }
  if (i0 < tpp && i0 + per >= tpp) off[(uint64_t)p * (tpp + 1) + tpp] = acc;
  if (tpp == 0 && tid == 0) off[(uint64_t)p * (tpp + 1)] = 0;
  __syncthreads();
  // pass 2: place the impulses (cnt now holds each tile's cursor)
  uint32_t* ppos = pos + (uint64_t)p * 4 * nbox;
  T* pval = val + (uint64_t)p * 4 * nbox;
  for (uint32_t b = tid; b < nbox; b += BIN_THREADS) {
    const int4 bx = pb[b];
    if (!box_ok(bx, W, H)) continue;
    const T gv = pg[b];
    const int ys[2] = {bx.z, bx.x - 1}, xs[2] = {bx.w, bx.y - 1};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        if (ys[a] >= 0 && xs[c] >= 0) {
          const uint32_t ty = (uint32_t)ys[a] / TH, tx = (uint32_t)xs[c] / TW;
          const uint32_t slot = atomicAdd(&cnt[ty * nTx + tx], 1u);
          ppos[slot] = ((uint32_t)ys[a] - ty * TH) * TW + ((uint32_t)xs[c] - tx * TW);
          pval[slot] = (a ^ c) ? -gv : gv;
        }
  }
}

// grad_x = reverse SAT of the boxes' corner impulses (+g at (y1,x1) and
// (y0-1,x0-1), -g at (y0-1,x1) and (y1,x0-1)). Each tile builds its slice of
// the impulse image in LDS (from its binned impulses when `off` is given,
// otherwise from the plane's whole box list) instead of reading a dense
// impulse image, then runs the reverse-SAT tile operator.
template <typename T>
__global__ void __launch_bounds__(64 * GradCfg<T>::NW)
    box_grad_kernel(const int4* __restrict__ boxes, const T* __restrict__ gout, uint32_t nbox, ImgArgs g,
                    uint32_t tpp, int64_t out_plane, uint32_t ep_arg, const WsHeader* hdr,
                    const uint32_t* __restrict__ off, const uint32_t* __restrict__ pos, const T* __restrict__ val) {
  using Op = GradOp<T>;
  constexpr int TH = Op::TH, TW = Op::TW, NT = Op::NT, RPT = GradCfg<T>::RPT;
  union Smem {
    typename Op::Smem c;
    T imp[TH * TW];
  };
  __shared__ Smem sm;
  const uint32_t ep4 = launch_ep4(ep_arg, hdr);
  const uint32_t t = blockIdx.x;
  const uint32_t p = t / tpp, lt = t - p * tpp;
  ImgArgs gp = g;
  gp.out = (T*)g.out + (int64_t)p * out_plane;
  gp.vcarry = g.vcarry + (int64_t)p * tpp * Op::VWORDS;
  gp.hcarry = g.hcarry + (int64_t)p * tpp * Op::HWORDS;
  const typename Op::Coords k = Op::coords(gp, lt);

  for (int i = threadIdx.x; i < TH * TW; i += NT) sm.imp[i] = T(0);
  __syncthreads();
  const int4* pb = boxes + (uint64_t)p * nbox;
  const T* pg = gout + (uint64_t)p * nbox;
  if (off) {
    const uint32_t pt = (uint32_t)(k.py0 / TH) * g.nTx + (uint32_t)(k.px0 / TW);  // physical tile
    const uint32_t* po = off + (uint64_t)p * (tpp + 1);
    const uint64_t base = (uint64_t)p * 4 * nbox;
    const uint32_t e1 = po[pt + 1];
    for (uint32_t e = po[pt] + threadIdx.x; e < e1; e += NT) atomicAdd(&sm.imp[pos[base + e]], val[base + e]);
  }
  for (uint32_t b = threadIdx.x; !off && b < nbox; b += NT) {
    const int4 bx = pb[b];
    if (!box_ok(bx, g.W, g.H)) continue;
    const T gv = pg[b];
    const int64_t ys[2] = {bx.z, (int64_t)bx.x - 1};
    const int64_t xs[2] = {bx.w, (int64_t)bx.y - 1};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int64_t ly = ys[a] - k.py0, lx = xs[c] - k.px0;
        if (ys[a] >= 0 && xs[c] >= 0 && ly >= 0 && ly < TH && lx >= 0 && lx < TW)
          atomicAdd(&sm.imp[ly * TW + lx], (a ^ c) ? -gv : gv);
      }
  }
  __syncthreads();
  typename Op::A v[RPT][4];
  const int64_t lcol = k.pcol - k.px0;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int64_t ly = Op::prow_of(k, r) - k.py0;
#pragma unroll
    for (int c = 0; c < 4; ++c) v[r][c] = sm.imp[ly * TW + lcol + 3 - c];
  }
  __syncthreads();
  Op::finish(gp, lt, ep4, sm.c, k, v);
}

}  // namespace satdev

namespace sathost {

using namespace satdev;

sat_status launch_box_sums(int dtype, const void* sat, int64_t pitch, int64_t plane_stride, uint32_t width,
                           uint32_t height, uint32_t planes, const int32_t* boxes, uint32_t nbox, void* out,
                           hipStream_t stream) {
  const uint64_t count = (uint64_t)planes * nbox;
  if (count == 0) return SAT_OK;
  const uint64_t blocks = (count + 255) / 256;
  if (blocks > 0x7fffffffull) return SAT_ERR_INVALID;
  if (dtype == SAT_F32)
    hipLaunchKernelGGL(box_sums_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream, (const float*)sat, pitch,
                       plane_stride, width, height, count, nbox, (const int4*)boxes, (float*)out);
  else if (dtype == SAT_F64)
    hipLaunchKernelGGL(box_sums_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream, (const double*)sat,
                       pitch, plane_stride, width, height, count, nbox, (const int4*)boxes, (double*)out);
  else
    return SAT_ERR_DTYPE;
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

sat_status launch_box_scatter(int dtype, const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_sat,
                              int64_t pitch, int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes,
                              hipStream_t stream) {
  const uint64_t count = (uint64_t)planes * nbox;
  if (count == 0) return SAT_OK;
  const uint64_t blocks = (count + 255) / 256;
  if (blocks > 0x7fffffffull) return SAT_ERR_INVALID;
  if (dtype == SAT_F32)
    hipLaunchKernelGGL(box_scatter_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream, (const int4*)boxes,
                       (const float*)grad_out, count, nbox, (float*)grad_sat, pitch, plane_stride, width, height);
  else if (dtype == SAT_F64)
    hipLaunchKernelGGL(box_scatter_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream, (const int4*)boxes,
                       (const double*)grad_out, count, nbox, (double*)grad_sat, pitch, plane_stride, width, height);
  else
    return SAT_ERR_DTYPE;
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

template <typename T>
static uint64_t grad_words(uint32_t width, uint32_t height, uint32_t planes) {
  using Op = GradOp<T>;
  const uint64_t nTx = (width + Op::TW - 1) / Op::TW, nTy = (height + Op::TH - 1) / Op::TH;
  return nTx * nTy * planes * (uint64_t)(Op::VWORDS + Op::HWORDS);
}

size_t box_grad_carry_bytes(int dtype, uint32_t width, uint32_t height, uint32_t planes) {
  const uint64_t w = dtype == SAT_F64 ? grad_words<double>(width, height, planes) : grad_words<float>(width, height, planes);
  return (size_t)(WS_HEADER_BYTES + w * 8);
}

template <typename T>
static uint64_t grad_tiles_per_plane(uint32_t width, uint32_t height) {
  using Op = GradOp<T>;
  return (uint64_t)((width + Op::TW - 1) / Op::TW) * ((height + Op::TH - 1) / Op::TH);
}

// Bytes of the impulse bins (0: the shape takes the unbinned path).
size_t box_grad_bin_bytes(int dtype, uint32_t width, uint32_t height, uint32_t planes, uint32_t nbox) {
  const uint64_t tpp = dtype == SAT_F64 ? grad_tiles_per_plane<double>(width, height)
                                        : grad_tiles_per_plane<float>(width, height);
  if (tpp == 0 || tpp > BIN_MAX_TILES || (uint64_t)nbox * 4 > 0xffffffffull) return 0;
  const uint64_t n = (uint64_t)planes * nbox * 4;
  const uint64_t off_b = (planes * (tpp + 1) * 4 + 255) / 256 * 256, pos_b = (n * 4 + 255) / 256 * 256;
  return (size_t)(off_b + pos_b + n * (dtype == SAT_F64 ? 8 : 4));
}

template <typename T>
static sat_status grad_launch(const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_x, int64_t pitch,
                              int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes, void* ws,
                              size_t ws_bytes, void* bins, uint32_t epoch, hipStream_t stream) {
  using Op = GradOp<T>;
  const uint64_t nTx = (width + Op::TW - 1) / Op::TW, nTy = (height + Op::TH - 1) / Op::TH;
  const uint64_t tpp = nTx * nTy, ntiles = tpp * planes;
  if (ntiles == 0) return SAT_OK;
  if (ntiles > 0x7fffffffull) return SAT_ERR_INVALID;
  const uint64_t vw = ntiles * Op::VWORDS, hw = ntiles * Op::HWORDS;
  if (ws_bytes < WS_HEADER_BYTES + (vw + hw) * 8) return SAT_ERR_WORKSPACE;
  uint64_t* carries = (uint64_t*)((char*)ws + WS_HEADER_BYTES);
  ImgArgs g;
  g.in = nullptr;
  g.out = grad_x;
  g.ldi = 0;
  g.ldo = pitch;
  g.H = height;
  g.W = width;
  g.nTx = (uint32_t)nTx;
  g.nTy = (uint32_t)nTy;
  g.vcarry = carries;
  g.hcarry = carries + vw;
  g.fault.ws = &((WsHeader*)ws)->fault;
  g.fault.host = host_fault_word();
  uint32_t *off = nullptr, *pos = nullptr;
  T* val = nullptr;
  if (bins && nbox > 0) {
    const uint64_t n = (uint64_t)planes * nbox * 4;
    const uint64_t off_b = (planes * (tpp + 1) * 4 + 255) / 256 * 256, pos_b = (n * 4 + 255) / 256 * 256;
    off = (uint32_t*)bins;
    pos = (uint32_t*)((char*)bins + off_b);
    val = (T*)((char*)bins + off_b + pos_b);
    hipLaunchKernelGGL(box_bin_kernel<T>, dim3(planes), dim3(BIN_THREADS), 0, stream, (const int4*)boxes,
                       (const T*)grad_out, nbox, width, height, (uint32_t)nTx, (uint32_t)tpp, off, pos, val);
  }
  hipLaunchKernelGGL(box_grad_kernel<T>, dim3((uint32_t)ntiles), dim3(Op::NT), 0, stream, (const int4*)boxes,
                     (const T*)grad_out, nbox, g, (uint32_t)tpp, plane_stride, epoch, (const WsHeader*)ws,
                     (const uint32_t*)off, (const uint32_t*)pos, (const T*)val);
  return hipGetLastError() == hipSuccess ? SAT_OK : SAT_ERR_HIP;
}

sat_status launch_box_grad(int dtype, const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_x,
                           int64_t pitch, int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes,
                           void* ws, size_t ws_bytes, void* bins, uint32_t epoch, hipStream_t stream) {
  if (dtype == SAT_F32)
    return grad_launch<float>(boxes, nbox, grad_out, grad_x, pitch, plane_stride, width, height, planes, ws, ws_bytes,
                              bins, epoch, stream);
  if (dtype == SAT_F64)
    return grad_launch<double>(boxes, nbox, grad_out, grad_x, pitch, plane_stride, width, height, planes, ws, ws_bytes,
                               bins, epoch, stream);
  return SAT_ERR_DTYPE;
}

}  // namespace sathost
