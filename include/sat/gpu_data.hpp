// gpu_data.hpp - pitched 2-D / 3-D device buffers with the interface of the
// reference's GpuData / GpuDataRef (integralImage/cudaLib.cuh:56-190) and its
// grid helpers UpDivide / UpRound / Align (cudaLib.cuh:25-30), on the HIP
// runtime, so that code written against the reference's harness
// (DevData<T> buf(w, h); buf.CopyFromHost(p, w, w, h); kernel(buf.GetData(),
// ..., buf.DataPitch(), ...)) keeps compiling.
//
// Kept alike on purpose: member and method names, argument order and meaning
// (widths, heights and pitches in ELEMENTS; DataPitch() = pitch bytes /
// sizeof(T), cudaLib.cuh:178-180), buffer reuse in MallocBuffer when the new
// shape fits the old allocation (cudaLib.cuh:118-123). Differences:
//   * HIP calls report errors: `status` holds the last hipError_t of the
//     object's own calls (the reference's CUDA_CHECK_ERROR is empty,
//     cudaLib.cuh:21-22); methods never throw or abort;
//   * GpuDataRef also implements CopyToDevice / CopyFromDevice (device to
//     device 2-D copies), which the reference's GpuData forwards to but
//     GpuDataRef never defines;
//   * GpuStream creates and destroys its stream (the reference's holds an
//     uninitialised cudaStream_t);
//   * every copy takes an optional stream (default: the null stream,
//     synchronous as in the reference).
#ifndef SAT_AMD_GPU_DATA_HPP
#define SAT_AMD_GPU_DATA_HPP

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>

// Grid arithmetic of the reference's harnesses (cudaLib.cuh:25-30), usable on
// the host and in kernels: Align and UpRound round x up to a multiple of y,
// UpDivide is the number of y-sized pieces that cover x.
__host__ __device__ __forceinline__ int Align(int x, int y) { return (x + y - 1) / y * y; }
__host__ __device__ __forceinline__ int UpDivide(int x, int y) { return (x + y - 1) / y; }
__host__ __device__ __forceinline__ int UpRound(int x, int y) { return UpDivide(x, y) * y; }

namespace sat {

struct GpuStream {
  hipStream_t stream = nullptr;
  GpuStream() { hipStreamCreate(&stream); }
  ~GpuStream() {
    if (stream) hipStreamDestroy(stream);
  }
  GpuStream(const GpuStream&) = delete;
  GpuStream& operator=(const GpuStream&) = delete;
};

template <class T>
struct GpuDataRef {
  GpuDataRef() = default;
  GpuDataRef(T* ptr, size_t ww, size_t pitch, size_t hh, size_t dd)
      : data(ptr), width(ww), data_pitch(pitch), height(hh), depth(dd) {}

  // zero the width x (height*depth) element rectangle
  hipError_t Zero(hipStream_t st = nullptr) const {
    const hipError_t e = hipMemset2DAsync(data, data_pitch * sizeof(T), 0, width * sizeof(T), height * depth, st);
    return e != hipSuccess ? e : sync(st);
  }
  // copies of a _w x (_h*_d) element rectangle; `_pitch` is the other side's row pitch in elements
  hipError_t CopyToHost(T* ptr, size_t _w, size_t _pitch, size_t _h, size_t _d = 1, hipStream_t st = nullptr) const {
    return copy2d(ptr, _pitch, data, data_pitch, _w, _h * _d, hipMemcpyDeviceToHost, st);
  }
  hipError_t CopyFromHost(const T* ptr, size_t _w, size_t _pitch, size_t _h, size_t _d = 1,
                          hipStream_t st = nullptr) const {
    return copy2d(data, data_pitch, ptr, _pitch, _w, _h * _d, hipMemcpyHostToDevice, st);
  }
  hipError_t CopyToDevice(T* ptr, size_t _w, size_t _pitch, size_t _h, size_t _d = 1,
                          hipStream_t st = nullptr) const {
    return copy2d(ptr, _pitch, data, data_pitch, _w, _h * _d, hipMemcpyDeviceToDevice, st);
  }
  hipError_t CopyFromDevice(const T* ptr, size_t _w, size_t _pitch, size_t _h, size_t _d = 1,
                            hipStream_t st = nullptr) const {
    return copy2d(data, data_pitch, ptr, _pitch, _w, _h * _d, hipMemcpyDeviceToDevice, st);
  }

  T* data = nullptr;
  size_t width = 0;
  size_t data_pitch = 0;  // elements
  size_t height = 0;
  size_t depth = 0;

 private:
  static hipError_t sync(hipStream_t st) { return st ? hipSuccess : hipStreamSynchronize(nullptr); }
  static hipError_t copy2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t w, size_t rows,
                           hipMemcpyKind kind, hipStream_t st) {
    if (w == 0 || rows == 0) return hipSuccess;
    const hipError_t e = hipMemcpy2DAsync(dst, dpitch * sizeof(T), src, spitch * sizeof(T), w * sizeof(T), rows,
                                          kind, st);
    return e != hipSuccess ? e : sync(st);
  }
};

template <class T>
struct GpuData {
  typedef T DataType;
  typedef GpuDataRef<T> DataRef;

  GpuData() = default;
  explicit GpuData(size_t w, size_t h = 1, size_t d = 1) { MallocBuffer(w, h, d); }
  virtual ~GpuData() { FreeBuffer(); }
  GpuData(const GpuData&) = delete;
  GpuData& operator=(const GpuData& obj) {
    if (this != &obj) {
      MallocBuffer(obj.width, obj.height, obj.depth);
      if (status == hipSuccess)
        status = GetDataRef().CopyFromDevice(obj.GetData(), obj.width, obj.DataPitch(), obj.height, obj.depth);
    }
    return *this;
  }

  DataRef GetDataRef() const { return DataRef(GetData(), width, DataPitch(), height, depth); }

  // w x h x d elements; rows padded to the pitch hipMalloc3D picks. An
  // allocation whose rows and row count hold the new shape is reused.
  GpuData& MallocBuffer(size_t w, size_t h, size_t d) {
    if (width == w && height == h && depth == d) return *this;
    if (sizeof(T) * w <= dev_data.pitch && h * d > 0 && h * d <= extent.height * extent.depth) {
      width = w;
      height = h;
      depth = d;
      return *this;
    }
    FreeBuffer();
    extent = make_hipExtent(sizeof(T) * w, h, d);
    status = hipMalloc3D(&dev_data, extent);
    if (status != hipSuccess) {
      std::memset(&dev_data, 0, sizeof(dev_data));
      std::memset(&extent, 0, sizeof(extent));
      return *this;
    }
    width = w;
    height = h;
    depth = d;
    channels = d == 1 ? (h == 1 ? 1 : 2) : 3;
    return *this;
  }
  GpuData& MallocBuffer(size_t w, size_t h) { return MallocBuffer(w, h, 1); }
  GpuData& MallocBuffer(size_t w) { return MallocBuffer(w, 1, 1); }

  void FreeBuffer() {
    if (dev_data.ptr) hipFree(dev_data.ptr);
    std::memset(&dev_data, 0, sizeof(dev_data));
    std::memset(&extent, 0, sizeof(extent));
    width = height = depth = channels = 0;
  }

  GpuData& CopyToHost(T* ptr, size_t ww, size_t _pitch, size_t hh, size_t dd = 1) {
    status = GetDataRef().CopyToHost(ptr, ww, _pitch, hh, dd);
    return *this;
  }
  GpuData& CopyToDevice(T* ptr, size_t ww, size_t _pitch, size_t hh, size_t dd = 1) {
    status = GetDataRef().CopyToDevice(ptr, ww, _pitch, hh, dd);
    return *this;
  }
  GpuData& CopyFromHost(const T* ptr, size_t ww, size_t _pitch, size_t hh, size_t dd = 1) {
    status = GetDataRef().CopyFromHost(ptr, ww, _pitch, hh, dd);
    return *this;
  }
  GpuData& CopyFromDevice(const T* ptr, size_t ww, size_t _pitch, size_t hh, size_t dd = 1) {
    status = GetDataRef().CopyFromDevice(ptr, ww, _pitch, hh, dd);
    return *this;
  }
  GpuData& Zero() {
    status = GetDataRef().Zero();
    return *this;
  }

  size_t DataPitch() const { return dev_data.pitch / sizeof(T); }  // elements
  T* GetData() const { return (T*)dev_data.ptr; }

  hipPitchedPtr dev_data{};
  hipExtent extent{};
  size_t width = 0;
  size_t height = 0;
  size_t depth = 0;
  size_t channels = 0;
  hipError_t status = hipSuccess;  // last error of this object's HIP calls
};

}  // namespace sat

// the reference's spellings (cudaLib.cuh:183-185)
template <class T> using DevData = sat::GpuData<T>;
template <class T> using GpuData = sat::GpuData<T>;
template <class T> using GpuDataRef = sat::GpuDataRef<T>;
using DevStream = sat::GpuStream;
using GpuStream = sat::GpuStream;

#endif  // SAT_AMD_GPU_DATA_HPP
