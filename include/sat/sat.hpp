// sat.hpp - C++ drop-ins for pengdada/SAT's templated entry points, over the
// C ABI of sat.h. Names, template-parameter order and argument order follow
// the reference; the reference's __global__ kernels become host functions
// that enqueue HIP kernels on `stream` (a hipStream_t, default stream if
// null) and return a sat_status instead of failing silently.
//
//   BlockScan::blockScan        integralImage/ScanRow-BRLT.cu:24-25
//   SerielScan::serielScan      integralImage/BRLT-Scan.cu:14-15   (alias SerialScan)
//   IncrementScan::ScanY        integralImage/ScanRowColumn.cu:15-16 (alias IncreamentScan)
//   IncrementScan::ScanX        integralImage/ScanRowColumn.cu:144-145
//
// The reference's tuning template parameters (BLOCK_SIZE, SMEM_COUNT,
// BLOCK_DIM_X, BLOCK_DIM_Y, COUNT_PER_THREAD, WARP_COUNT) are accepted and
// ignored: tiles are chosen for wave64 inside the library.
#ifndef SAT_AMD_SAT_HPP
#define SAT_AMD_SAT_HPP

#include <stddef.h>
#include <stdint.h>

#include "sat.h"

namespace sat {

template <typename T> struct dtype_of;
template <> struct dtype_of<uint8_t> { static constexpr int value = SAT_U8; };
template <> struct dtype_of<int8_t> { static constexpr int value = SAT_I8; };
template <> struct dtype_of<char> { static constexpr int value = SAT_I8; };  // the reference's `char` (i8)
template <> struct dtype_of<uint32_t> { static constexpr int value = SAT_U32; };
template <> struct dtype_of<int32_t> { static constexpr int value = SAT_I32; };
template <> struct dtype_of<float> { static constexpr int value = SAT_F32; };
template <> struct dtype_of<double> { static constexpr int value = SAT_F64; };

// Full SAT with any of the algorithms (sat_alg); `workspace` null = library-managed.
template <typename TSrc, typename TDst, int SCAN_TYPE = SAT_SCAN_KOGGE_STONE>
inline sat_status integral(const TSrc* in, size_t in_pitch, TDst* out, size_t out_pitch, unsigned width,
                           unsigned height, int alg = SAT_ALG_SINGLE_PASS, void* workspace = nullptr,
                           size_t workspace_bytes = 0, void* stream = nullptr) {
  return (sat_status)sat_integral(alg, SCAN_TYPE, dtype_of<TSrc>::value, dtype_of<TDst>::value, in, (int64_t)in_pitch,
                                  out, (int64_t)out_pitch, width, height, workspace, workspace_bytes, stream);
}

template <typename TSrc, typename TDst>
inline size_t integral_workspace_bytes(unsigned width, unsigned height, int alg = SAT_ALG_SINGLE_PASS) {
  return sat_integral_workspace_bytes(alg, dtype_of<TSrc>::value, dtype_of<TDst>::value, width, height);
}

// SAT_ERR_LOOKBACK if a look-back of any finished launch of this library on
// the current device gave up since the last check (the results of that
// launch are undefined), else SAT_OK; clears the flags it reads. With
// `synchronize`, waits for `stream` first (through sat_workspace_fault on its
// managed workspaces), so that every launch enqueued on it is covered.
inline sat_status check_faults(bool synchronize = true, void* stream = nullptr) {
  int w = 0;
  if (synchronize) {
    w = sat_workspace_fault(nullptr, stream);
    if (w < 0) return (sat_status)(-w);
  }
  const int f = sat_device_fault(1);
  if (f < 0) return (sat_status)(-f);
  return (f || w) ? SAT_ERR_LOOKBACK : SAT_OK;
}

// Suffix sums (adjoint of the SAT); T is float, double, uint32_t or int32_t.
template <typename T, int SCAN_TYPE = SAT_SCAN_KOGGE_STONE>
inline sat_status integral_reverse(const T* in, size_t in_pitch, T* out, size_t out_pitch, unsigned width,
                                   unsigned height, void* stream = nullptr) {
  return (sat_status)sat_integral_reverse(SCAN_TYPE, dtype_of<T>::value, in, (int64_t)in_pitch, out,
                                          (int64_t)out_pitch, width, height, nullptr, 0, stream);
}

}  // namespace sat

namespace BlockScan {
// ScanRow-BRLT pass: out[x*heightStride + y] = sum_{k<=x} in[y*widthStride + k]
template <typename TSrc, typename TDst, int SCAN_TYPE, unsigned BLOCK_SIZE = 32, unsigned SMEM_COUNT = 8>
inline sat_status blockScan(const TSrc* dataIn, TDst* dataOut, unsigned width, unsigned widthStride, unsigned height,
                            unsigned heightStride, void* stream = nullptr) {
  return (sat_status)sat_block_scan(SCAN_TYPE, sat::dtype_of<TSrc>::value, sat::dtype_of<TDst>::value, dataIn, dataOut,
                                    width, widthStride, height, heightStride, stream);
}
}  // namespace BlockScan

namespace SerielScan {
// BRLT-ScanRow pass: same result layout as BlockScan::blockScan
template <typename TSrc, typename TDst, unsigned BLOCK_SIZE = 32, unsigned SMEM_COUNT = 8, unsigned BLOCK_DIM_X = 1024>
inline sat_status serielScan(const TSrc* dataIn, TDst* dataOut, unsigned width, unsigned widthStride, unsigned height,
                             unsigned heightStride, void* stream = nullptr) {
  return (sat_status)sat_seriel_scan(sat::dtype_of<TSrc>::value, sat::dtype_of<TDst>::value, dataIn, dataOut, width,
                                     widthStride, height, heightStride, stream);
}
}  // namespace SerielScan
namespace SerialScan = SerielScan;

namespace IncrementScan {
// ScanRowColumn column pass: out[y*widthStride + x] = sum_{k<=y} in[k*widthStride + x]
template <typename TSrc, typename TDst, unsigned SCAN_TYPE, unsigned COUNT_PER_THREAD = 32, unsigned WARP_COUNT = 32>
inline sat_status ScanY(const TSrc* dataIn, TDst* dataOut, unsigned width, unsigned widthStride, unsigned height,
                        void* stream = nullptr) {
  return (sat_status)sat_scan_y((int)SCAN_TYPE, sat::dtype_of<TSrc>::value, sat::dtype_of<TDst>::value, dataIn,
                                dataOut, width, widthStride, height, stream);
}
// ScanRowColumn row pass: out[y*widthStride + x] = sum_{k<=x} in[y*widthStride + k]
template <typename TSrc, typename TDst, int SCAN_TYPE, int BLOCK_DIM_X = 32, int BLOCK_DIM_Y = 32,
          int COUNT_PER_THREAD = 32>
inline sat_status ScanX(const TSrc* dataIn, TDst* dataOut, unsigned width, unsigned widthStride, unsigned height,
                        void* stream = nullptr) {
  return (sat_status)sat_scan_x(SCAN_TYPE, sat::dtype_of<TSrc>::value, sat::dtype_of<TDst>::value, dataIn, dataOut,
                                width, widthStride, height, stream);
}
}  // namespace IncrementScan
namespace IncreamentScan = IncrementScan;

#endif  // SAT_AMD_SAT_HPP
