// A program written the way the reference's main.cpp (integralImage/main.cpp:
// 6-41) drives its harnesses: argv[1] picks the algorithm (0 ScanRow-BRLT,
// 1 BRLT-ScanRow, 2 ScanRowColumn) and the harness reads the rest of argv
// (<scan> <dtype in> <dtype out> <n>). The harness entry points are reached
// through plain `extern` declarations, as there, and resolved from
// libsat_amd.so, the only library this program links.
//
// With "all" as argv[1] it also runs the header templates directly
// (Test<> of the three algorithms, SerielScan::TestX) on non-square sizes and
// the grid helpers. Exits 0 when every comparison succeeded.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/sat/harness.hpp"

int main(int argc, char** argv) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    std::printf("harness: no HIP device\n");
    return 77;
  }
  int idx = 0;
  if (argc > 1 && std::strcmp(argv[1], "all") != 0) idx = std::atoi(argv[1]);
  if (argc > 1 && std::strcmp(argv[1], "all") == 0) {
    int bad = 0;
    bad += !BlockScan::Test<unsigned char, unsigned int, 0>(1000, 333);
    bad += !SerielScan::Test<float, float, 1>(517, 1031);
    bad += !SerielScan::TestX<unsigned char, double, 0>(300, 700);
    bad += !IncrementScan::Test<char, int, 1>(777, 129);
    bad += !IncreamentScan::Test<int, double, 0>(64, 2000);
    const bool grid = UpDivide(1000, 32) == 32 && UpRound(1000, 32) == 1024 && Align(1024, 32) == 1024;
    if (!grid) std::printf("grid helpers: wrong values\n");
    std::printf("harness templates: %d failures\n", bad + !grid);
    return bad + !grid ? 1 : 0;
  }
  if (idx == 0) {
    extern void TestBlockScan(int argc, char** argv);
    TestBlockScan(argc, argv);
  }
  if (idx == 1) {
    extern void TestSerielScan(int argc, char** argv);
    TestSerielScan(argc, argv);
  }
  if (idx == 2) {
    extern void TestIncreamentScan(int argc, char** argv);
    TestIncreamentScan(argc, argv);
  }
  const int f = TestFailures();
  std::printf("harness: %d failed comparisons\n", f);
  return f ? 1 : 0;
}
