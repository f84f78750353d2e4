// sat: command-line driver of the SAT library with the reference's interface
//
//   sat <alg> <scan> <dt0> <dt1> <n> [--repeat R] [--check] [--csv FILE] [--ones]
//
// (main.cpp:6-40 and GetArgs, integralImage.h:56-63): alg 0 ScanRow-BRLT,
// 1 BRLT-ScanRow, 2 ScanRowColumn (the reference's three two-pass
// algorithms), 3 the one-pass kernel; scan 0 Kogge-Stone, 1 Ladner-Fischer;
// dt0 / dt1 the DataType codes of input and output (integralImage.h:22-32).
// For i = 1..n it runs an i*1024 x i*1024 image, as the reference's Test<>
// loops do (ScanRow-BRLT.cu:272-273), and prints the reference's line
// "w, h, total time = <host ms>, <event ms>" (ScanRow-BRLT.cu:239) plus GB/s of
// algorithmic bytes. Inputs are rand() % 2 as in the reference
// (ScanRow-BRLT.cu:205) or all ones (--ones). --check compares the result with
// the one-pass kernel's (or, for alg 3, with ScanRow-BRLT's): integer outputs
// must be identical, floating-point ones within 1e-5 relative. --csv appends
// alg,scan,dt0,dt1,width,height,ms,GBs rows. All device work goes through the
// library's C ABI (include/sat/sat.h).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sat/sat.h"

namespace {

const char* kAlgName[] = {"ScanRow-BRLT", "BRLT-ScanRow", "ScanRowColumn", "SinglePass"};

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "sat: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

bool sat_ok(int s, const char* what) {
  if (s == SAT_OK) return true;
  std::fprintf(stderr, "sat: %s: %s (%d)\n", what, sat_status_string(s), s);
  return false;
}

// host input: rand() % 2 (or 1) in the input type
void fill_input(std::vector<unsigned char>& h, int dt, size_t n, bool ones) {
  const size_t es = sat_dtype_size(dt);
  h.assign(n * es, 0);
  for (size_t i = 0; i < n; ++i) {
    const int v = ones ? 1 : (std::abs(std::rand()) % 2);
    unsigned char* p = h.data() + i * es;
    switch (dt) {
      case SAT_U8: *(uint8_t*)p = (uint8_t)v; break;
      case SAT_I8: *(int8_t*)p = (int8_t)v; break;
      case SAT_U16: *(uint16_t*)p = (uint16_t)v; break;
      case SAT_I16: *(int16_t*)p = (int16_t)v; break;
      case SAT_U32: *(uint32_t*)p = (uint32_t)v; break;
      case SAT_I32: *(int32_t*)p = (int32_t)v; break;
      case SAT_F32: *(float*)p = (float)v; break;
      case SAT_F64: *(double*)p = (double)v; break;
    }
  }
}

double value_at(const unsigned char* p, int dt, size_t i) {
  switch (dt) {
    case SAT_U32: return ((const uint32_t*)p)[i];
    case SAT_I32: return ((const int32_t*)p)[i];
    case SAT_F32: return ((const float*)p)[i];
    case SAT_F64: return ((const double*)p)[i];
    default: return 0;
  }
}

bool same_result(const std::vector<unsigned char>& a, const std::vector<unsigned char>& b, int dt, size_t n) {
  if (dt == SAT_U32 || dt == SAT_I32) return std::memcmp(a.data(), b.data(), a.size()) == 0;
  for (size_t i = 0; i < n; ++i) {
    const double x = value_at(a.data(), dt, i), y = value_at(b.data(), dt, i);
    if (std::fabs(x - y) > 1e-5 * std::fmax(std::fabs(x), std::fabs(y)) + 1e-30) return false;
  }
  return true;
}

int usage() {
  std::fprintf(stderr,
               "usage: sat <alg 0-3> <scan 0-1> <dt0> <dt1> <n> [--repeat R] [--check] [--csv FILE] [--ones]\n"
               "  alg: 0 ScanRow-BRLT, 1 BRLT-ScanRow, 2 ScanRowColumn, 3 one-pass\n"
               "  dt:  0 u8, 1 i8, 4 u32, 5 i32, 6 f32, 7 f64 (pairs as sat_dtype_pair_supported)\n"
               "  sizes i*1024 x i*1024 for i = 1..n\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  // defaults of the reference's TestBlockScan (ScanRow-BRLT.cu:262-265)
  int alg = 0, scan = SAT_SCAN_LF, dt0 = SAT_F32, dt1 = SAT_F32, n = 1, repeat = 1;
  bool check = false, ones = false;
  const char* csv = nullptr;
  std::vector<std::string> pos;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--check") check = true;
    else if (a == "--ones") ones = true;
    else if (a == "--repeat" && i + 1 < argc) repeat = std::atoi(argv[++i]);
    else if (a == "--csv" && i + 1 < argc) csv = argv[++i];
    else if (a == "-h" || a == "--help") return usage();
    else pos.push_back(a);
  }
  if (!pos.empty()) alg = std::atoi(pos[0].c_str());
  if (pos.size() >= 5) {  // GetArgs reads argv[2..5] only when all are given
    scan = std::atoi(pos[1].c_str());
    dt0 = std::atoi(pos[2].c_str());
    dt1 = std::atoi(pos[3].c_str());
    n = std::atoi(pos[4].c_str());
  }
  if (alg < 0 || alg > 3 || repeat < 1 || n < 0) return usage();
  if (!sat_dtype_pair_supported(dt0, dt1)) {
    std::fprintf(stderr, "sat: dtype pair %d -> %d is not supported\n", dt0, dt1);
    return 2;
  }
  FILE* out_csv = nullptr;
  if (csv) {
    out_csv = std::fopen(csv, "a");
    if (!out_csv) {
      std::perror(csv);
      return 1;
    }
  }
  std::printf("------------------------------------------------------\n");
  std::printf("%s scan=%d dt0=%d dt1=%d\n", kAlgName[alg], scan, dt0, dt1);
  hipStream_t st;
  if (!hip_ok(hipStreamCreate(&st), "hipStreamCreate")) return 1;
  hipEvent_t e0, e1;
  if (!hip_ok(hipEventCreate(&e0), "hipEventCreate") || !hip_ok(hipEventCreate(&e1), "hipEventCreate")) return 1;
  const size_t si = sat_dtype_size(dt0), so = sat_dtype_size(dt1);
  bool all_ok = true;
  for (int i = 1; i <= n; ++i) {
    const uint32_t w = 1024u * (uint32_t)i, h = w;
    const size_t px = (size_t)w * h;
    std::vector<unsigned char> hin;
    fill_input(hin, dt0, px, ones);
    void *din = nullptr, *dout = nullptr, *ws = nullptr;
    const size_t ws_bytes = sat_integral_workspace_bytes(alg, dt0, dt1, w, h);
    if (!hip_ok(hipMalloc(&din, px * si), "hipMalloc") || !hip_ok(hipMalloc(&dout, px * so), "hipMalloc") ||
        !hip_ok(hipMalloc(&ws, ws_bytes), "hipMalloc"))
      return 1;
    if (!hip_ok(hipMemcpy(din, hin.data(), px * si, hipMemcpyHostToDevice), "hipMemcpy")) return 1;
    auto run = [&](int a, void* o, void* wsp, size_t wsb) {
      return sat_integral(a, scan, dt0, dt1, din, w, o, w, w, h, wsp, wsb, st);
    };
    if (!sat_ok(run(alg, dout, ws, ws_bytes), "sat_integral")) return 1;  // warm-up
    hipStreamSynchronize(st);
    const auto t0 = std::chrono::steady_clock::now();
    hipEventRecord(e0, st);
    for (int k = 0; k < repeat; ++k)
      if (!sat_ok(run(alg, dout, ws, ws_bytes), "sat_integral")) return 1;
    hipEventRecord(e1, st);
    hipEventSynchronize(e1);
    const double host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / repeat;
    float ev_ms = 0;
    hipEventElapsedTime(&ev_ms, e0, e1);
    ev_ms /= repeat;
    const double gbs = (double)px * (si + so) / (ev_ms * 1e-3) / 1e9;
    std::printf("%u, %u, total time = %f, %f\n", w, h, host_ms, ev_ms);
    std::printf("GB/s (in+out bytes) = %.1f\n", gbs);
    if (out_csv) std::fprintf(out_csv, "%d,%d,%d,%d,%u,%u,%.6f,%.2f\n", alg, scan, dt0, dt1, w, h, ev_ms, gbs);
    if (check) {
      const int other = alg == SAT_ALG_SINGLE_PASS ? SAT_ALG_SCANROW_BRLT : SAT_ALG_SINGLE_PASS;
      void* dref = nullptr;
      if (!hip_ok(hipMalloc(&dref, px * so), "hipMalloc")) return 1;
      if (!sat_ok(sat_integral(other, scan, dt0, dt1, din, w, dref, w, w, h, nullptr, 0, st), "sat_integral")) return 1;
      hipStreamSynchronize(st);
      std::vector<unsigned char> a(px * so), b(px * so);
      hipMemcpy(a.data(), dout, px * so, hipMemcpyDeviceToHost);
      hipMemcpy(b.data(), dref, px * so, hipMemcpyDeviceToHost);
      const bool ok = same_result(a, b, dt1, px);
      all_ok = all_ok && ok;
      std::printf("compare = %s\n", ok ? "successed" : "failed");
      hipFree(dref);
    }
    hipFree(din);
    hipFree(dout);
    hipFree(ws);
    std::printf("end : %s\n", kAlgName[alg]);
  }
  if (out_csv) std::fclose(out_csv);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(st);
  return all_ok ? 0 : 1;
}
