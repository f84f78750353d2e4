// harness.hpp - the reference's test harness entry points over this library.
//
// pengdada/SAT drives its kernels from per-algorithm harnesses that a small
// main() reaches through `extern` declarations (integralImage/main.cpp:20-34):
//
//   BlockScan::Test<TSrc,TDst,SCAN>(w, h)     ScanRow-BRLT.cu:182-255
//   SerielScan::Test<TSrc,TDst,SCAN>(w, h)    BRLT-Scan.cu:192-266
//   SerielScan::TestX<TSrc,TDst,SCAN>(w, h)   BRLT-Scan.cu:128-190 (one pass)
//   IncrementScan::Test<TSrc,TDst,SCAN>(w, h) ScanRowColumn.cu:424-499
//   void TestBlockScan(int, char**)           ScanRow-BRLT.cu:260
//   void TestSerielScan(int, char**)          BRLT-Scan.cu:271
//   void TestIncreamentScan(int, char**)      ScanRowColumn.cu:504
//
// The templates below keep the names, template parameters, argument meaning
// and printed protocol of those harnesses: a `rand() % 2` input of
// width x height TSrc pixels in pitched DevData buffers, the algorithm's two
// passes through a transposed (h x w) temporary, timed by host clock and HIP
// events, and the line "<w>, <h>, total time = <host ms>, <event ms>". The
// check differs: the reference compares with its CPU integral through a
// Compare() that returns true before comparing (integralImage.h:268); here
// the result is compared with this library's one-pass SAT of the same input
// on the GPU (bit-exact for integer outputs, within 256 u |S| for floats),
// and the outcome is both printed ("compare = successed" / "failed", the
// reference's words) and returned. The dispatchers (TestBlockScan,
// TestSerielScan, TestIncreamentScan) live in libsat_amd.so and take the
// reference's argv: <alg> <scan> <dtype in> <dtype out> <n>, square sizes
// 1024*i for i = 1..n (GetArgs, integralImage.h:56-63); dtype pairs the
// reference's dispatch does not list run nothing, as there.
#ifndef SAT_AMD_HARNESS_HPP
#define SAT_AMD_HARNESS_HPP

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpu_data.hpp"
#include "sat.hpp"

namespace sat {
namespace harness {

template <typename T> const char* type_name();
template <> inline const char* type_name<uint8_t>() { return "uchar"; }
template <> inline const char* type_name<int8_t>() { return "char"; }
template <> inline const char* type_name<char>() { return "char"; }
template <> inline const char* type_name<uint32_t>() { return "uint32"; }
template <> inline const char* type_name<int32_t>() { return "int32"; }
template <> inline const char* type_name<float>() { return "float32"; }
template <> inline const char* type_name<double>() { return "float64"; }

inline const char* scan_name(int scan) {
  return scan == SAT_SCAN_KOGGE_STONE ? "KoggeStone_SCAN" : scan == SAT_SCAN_LF ? "LF_SCAN" : "error Scan type!";
}

template <typename T> constexpr bool is_float() { return T(0.5) != T(0); }

// Element-wise comparison: exact for integer types, |a - b| <= 256 u max(|a|, |b|)
// for floating-point ones (two summation orders of the same table).
template <typename T>
inline bool compare(const std::vector<T>& got, const std::vector<T>& ref, int width, int height) {
  for (size_t i = 0; i < got.size(); ++i) {
    bool ok;
    if constexpr (is_float<T>()) {
      const double a = (double)got[i], b = (double)ref[i];
      const double u = sizeof(T) == 4 ? 5.960464477539063e-08 : 1.1102230246251565e-16;
      ok = std::fabs(a - b) <= 256.0 * u * std::fmax(std::fabs(a), std::fabs(b));
    } else {
      ok = got[i] == ref[i];
    }
    if (!ok) {
      std::printf("error : %d %d (%.17g != %.17g)\n", (int)(i % width), (int)(i / width), (double)got[i],
                  (double)ref[i]);
      return false;
    }
  }
  (void)height;
  return true;
}

enum RunMode : int {
  TWO_PASS_TRANSPOSED = 0,  // two passes through an h x w temporary, checked against the one-pass SAT
  TWO_PASS_SAME = 1,        // two passes through a w x h temporary, checked against the one-pass SAT
  ONE_PASS_TRANSPOSED = 2   // one transposing row-scan pass into the h x w temporary (TestX)
};

// One timed run of `passes` on a rand() % 2 input. `passes(devA, devTmp,
// devB, stream)` enqueues the algorithm's passes. The check runs a different
// GPU algorithm of this library on the same input: the one-pass SAT, or for
// ONE_PASS_TRANSPOSED the same-layout row scan (ScanX), transposed on the
// host (a data movement only).
template <typename TSrc, typename TDst, int SCAN_TYPE, class Passes>
inline bool run(const char* name, int width, int height, Passes passes, RunMode mode = TWO_PASS_TRANSPOSED) {
  std::printf("%s<%s-->%s>, ScanName=%s\n", name, type_name<TSrc>(), type_name<TDst>(), scan_name(SCAN_TYPE));
  std::printf("begin : %s\n", name);
  const size_t size = (size_t)width * height;
  std::vector<TSrc> vecA(size);
  for (size_t i = 0; i < size; ++i) vecA[i] = (TSrc)(std::abs(std::rand()) % 2);
  DevData<TSrc> devA(width, height);
  DevData<TDst> devB(width, height), devRef(width, height);
  DevData<TDst> devTmp = mode == TWO_PASS_SAME ? DevData<TDst>(width, height) : DevData<TDst>(height, width);
  devA.CopyFromHost(&vecA[0], width, width, height);
  DevStream SM;
  hipEvent_t start, stop;
  hipEventCreate(&start);
  hipEventCreate(&stop);
  const auto t0 = std::chrono::steady_clock::now();
  hipEventRecord(start, SM.stream);
  sat_status st = passes(devA, devTmp, devB, SM.stream);
  hipEventRecord(stop, SM.stream);
  hipEventSynchronize(stop);
  const double host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  float inc = 0.f;
  hipEventElapsedTime(&inc, start, stop);
  hipEventDestroy(start);
  hipEventDestroy(stop);
  std::printf("%d, %d, total time = %f, %f\n", width, height, host_ms, inc);
  bool ok = st == SAT_OK && devA.status == hipSuccess && devB.status == hipSuccess && devTmp.status == hipSuccess &&
            devRef.status == hipSuccess;
  if (st != SAT_OK) std::printf("sat status = %s\n", sat_status_string(st));
  if (ok) {
    std::vector<TDst> got(size), ref(size);
    if (mode == ONE_PASS_TRANSPOSED) {
      ok = sat_scan_rows(SCAN_TYPE, sat::dtype_of<TSrc>::value, sat::dtype_of<TDst>::value, devA.GetData(),
                         (int64_t)devA.DataPitch(), devRef.GetData(), (int64_t)devRef.DataPitch(), (uint32_t)width,
                         (uint32_t)height, SM.stream) == SAT_OK;
    } else {
      ok = sat::integral<TSrc, TDst, SCAN_TYPE>(devA.GetData(), devA.DataPitch(), devRef.GetData(),
                                                devRef.DataPitch(), width, height, SAT_ALG_SINGLE_PASS, nullptr, 0,
                                                SM.stream) == SAT_OK;
    }
    hipStreamSynchronize(SM.stream);
    ok = ok && sat_workspace_fault(nullptr, SM.stream) == 0;
    devRef.CopyToHost(&ref[0], width, width, height);
    if (mode == ONE_PASS_TRANSPOSED) {
      std::vector<TDst> rt(size);
      for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) rt[(size_t)x * height + y] = ref[(size_t)y * width + x];
      devTmp.CopyToHost(&got[0], height, height, width);
      ok = ok && devTmp.status == hipSuccess && compare(got, rt, height, width);
    } else {
      devB.CopyToHost(&got[0], width, width, height);
      ok = ok && devB.status == hipSuccess && compare(got, ref, width, height);
    }
  }
  std::printf("compare = %s\n", ok ? "successed" : "failed");
  std::printf("end : %s\n", name);
  return ok;
}

}  // namespace harness
}  // namespace sat

namespace BlockScan {
// ScanRow-BRLT: SAT = blockScan(blockScan(A)) through a transposed temporary
// (ScanRow-BRLT.cu:225-226).
template <typename TSrc, typename TDst, int SCAN_TYPE>
inline bool Test(int width, int height) {
  return sat::harness::run<TSrc, TDst, SCAN_TYPE>(
      "ScanRow-BRLT", width, height, [](DevData<TSrc>& A, DevData<TDst>& Tmp, DevData<TDst>& B, hipStream_t s) {
        sat_status st = blockScan<TSrc, TDst, SCAN_TYPE>(A.GetData(), Tmp.GetData(), (unsigned)A.width,
                                                         (unsigned)A.DataPitch(), (unsigned)A.height,
                                                         (unsigned)Tmp.DataPitch(), s);
        if (st != SAT_OK) return st;
        return blockScan<TDst, TDst, SCAN_TYPE>(Tmp.GetData(), B.GetData(), (unsigned)Tmp.width,
                                                (unsigned)Tmp.DataPitch(), (unsigned)Tmp.height,
                                                (unsigned)B.DataPitch(), s);
      });
}
}  // namespace BlockScan

namespace SerielScan {
// BRLT-ScanRow: SAT = serielScan(serielScan(A)) (BRLT-Scan.cu:237-238).
template <typename TSrc, typename TDst, int SCAN_TYPE>
inline bool Test(int width, int height) {
  return sat::harness::run<TSrc, TDst, SCAN_TYPE>(
      "BRLT-ScanRow", width, height, [](DevData<TSrc>& A, DevData<TDst>& Tmp, DevData<TDst>& B, hipStream_t s) {
        sat_status st = serielScan<TSrc, TDst>(A.GetData(), Tmp.GetData(), (unsigned)A.width, (unsigned)A.DataPitch(),
                                               (unsigned)A.height, (unsigned)Tmp.DataPitch(), s);
        if (st != SAT_OK) return st;
        return serielScan<TDst, TDst>(Tmp.GetData(), B.GetData(), (unsigned)Tmp.width, (unsigned)Tmp.DataPitch(),
                                      (unsigned)Tmp.height, (unsigned)B.DataPitch(), s);
      });
}
// One transposing pass into the h x w temporary, checked like the reference's
// TestX (row scan, then transpose, BRLT-Scan.cu:177-183).
template <typename TSrc, typename TDst, int SCAN_TYPE>
inline bool TestX(int width, int height) {
  return sat::harness::run<TSrc, TDst, SCAN_TYPE>(
      "BRLT-ScanRow", width, height,
      [](DevData<TSrc>& A, DevData<TDst>& Tmp, DevData<TDst>&, hipStream_t s) {
        return serielScan<TSrc, TDst>(A.GetData(), Tmp.GetData(), (unsigned)A.width, (unsigned)A.DataPitch(),
                                      (unsigned)A.height, (unsigned)Tmp.DataPitch(), s);
      },
      sat::harness::ONE_PASS_TRANSPOSED);
}
}  // namespace SerielScan

namespace IncrementScan {
// ScanRowColumn: SAT = ScanY(ScanX(A)) through a w x h temporary of TDst
// (ScanRowColumn.cu:475-476). The reference passes devA's pitch for both
// launches; here each buffer's own pitch is used: the drop-in templates when
// the element pitches agree, the two-pitch row / column scans otherwise.
template <typename TSrc, typename TDst, int SCAN_TYPE>
inline bool Test(int width, int height) {
  return sat::harness::run<TSrc, TDst, SCAN_TYPE>(
      "ScanRowColumn", width, height,
      [](DevData<TSrc>& A, DevData<TDst>& T, DevData<TDst>& B, hipStream_t s) {
        sat_status st;
        if (A.DataPitch() == T.DataPitch() && T.DataPitch() == B.DataPitch()) {
          st = ScanX<TSrc, TDst, SCAN_TYPE>(A.GetData(), T.GetData(), (unsigned)A.width, (unsigned)A.DataPitch(),
                                            (unsigned)A.height, s);
          if (st != SAT_OK) return st;
          return ScanY<TDst, TDst, SCAN_TYPE>(T.GetData(), B.GetData(), (unsigned)A.width, (unsigned)T.DataPitch(),
                                              (unsigned)A.height, s);
        }
        st = (sat_status)sat_scan_rows(SCAN_TYPE, sat::dtype_of<TSrc>::value, sat::dtype_of<TDst>::value, A.GetData(),
                                       (int64_t)A.DataPitch(), T.GetData(), (int64_t)T.DataPitch(),
                                       (uint32_t)A.width, (uint32_t)A.height, s);
        if (st != SAT_OK) return st;
        return (sat_status)sat_scan_cols(SCAN_TYPE, sat::dtype_of<TDst>::value, sat::dtype_of<TDst>::value,
                                         T.GetData(), (int64_t)T.DataPitch(), B.GetData(), (int64_t)B.DataPitch(),
                                         (uint32_t)A.width, (uint32_t)A.height, s);
      },
      sat::harness::TWO_PASS_SAME);
}
}  // namespace IncrementScan
namespace IncreamentScan = IncrementScan;

// The reference's dispatchers (defined in libsat_amd.so). argv as main.cpp
// passes it: argv[1] = algorithm, then GetArgs' <scan> <dtype in> <dtype out>
// <n> when argc >= 6 (defaults LF_SCAN, float32, float32, 1).
void TestBlockScan(int argc, char** argv);
void TestSerielScan(int argc, char** argv);
void TestIncreamentScan(int argc, char** argv);
// Number of Test<> runs of the dispatchers above whose comparison failed,
// since the library was loaded (no reference counterpart: the reference's
// harnesses only print).
int TestFailures();

#endif  // SAT_AMD_HARNESS_HPP
