// The reference's harness dispatchers (TestBlockScan, TestSerielScan,
// TestIncreamentScan; ScanRow-BRLT.cu:260-382, BRLT-Scan.cu:271-395,
// ScanRowColumn.cu:504-626) over the harness templates of
// include/sat/harness.hpp: the reference's argv (GetArgs,
// integralImage.h:56-63), square sizes 1024*i for i = 1..n, and its dtype
// dispatch (the same pairs in all three; a pair it does not list runs
// nothing). Host code only.
#include <atomic>
#include <cstdio>
#include <cstdlib>

#include "../../include/sat/harness.hpp"

namespace {

std::atomic<int> g_failures{0};

struct RunBlock {
  template <class S, class D, int SC> static bool go(int w, int h) { return BlockScan::Test<S, D, SC>(w, h); }
};
struct RunSeriel {
  template <class S, class D, int SC> static bool go(int w, int h) { return SerielScan::Test<S, D, SC>(w, h); }
};
struct RunIncrement {
  template <class S, class D, int SC> static bool go(int w, int h) { return IncrementScan::Test<S, D, SC>(w, h); }
};

template <class R, int SC>
void dispatch_pair(int d0, int d1, int size) {
  bool ran = true, ok = true;
  if (d0 == SAT_U8 && d1 == SAT_U32) ok = R::template go<unsigned char, unsigned int, SC>(size, size);
  else if (d0 == SAT_U8 && d1 == SAT_F32) ok = R::template go<unsigned char, float, SC>(size, size);
  else if (d0 == SAT_U8 && d1 == SAT_F64) ok = R::template go<unsigned char, double, SC>(size, size);
  else if (d0 == SAT_I8 && d1 == SAT_I32) ok = R::template go<char, int, SC>(size, size);
  else if (d0 == SAT_I8 && d1 == SAT_F32) ok = R::template go<char, float, SC>(size, size);
  else if (d0 == SAT_I8 && d1 == SAT_F64) ok = R::template go<char, double, SC>(size, size);
  else if (d0 == SAT_I32 && d1 == SAT_I32) ok = R::template go<int, int, SC>(size, size);
  else if (d0 == SAT_I32 && d1 == SAT_F64) ok = R::template go<int, double, SC>(size, size);
  else if (d0 == SAT_F32 && d1 == SAT_F32) ok = R::template go<float, float, SC>(size, size);
  else if (d0 == SAT_F32 && d1 == SAT_F64) ok = R::template go<float, double, SC>(size, size);
  else if (d0 == SAT_F64 && d1 == SAT_F64) ok = R::template go<double, double, SC>(size, size);
  else ran = false;
  if (ran && !ok) g_failures.fetch_add(1);
}

template <class R>
void dispatch(int argc, char** argv) {
  std::printf("------------------------------------------------------\n");
  // GetArgs (integralImage.h:56-63): defaults LF_SCAN, float32 -> float32, 1
  int scan = SAT_SCAN_LF, d0 = SAT_F32, d1 = SAT_F32, repeat = 1;
  if (argc >= 6) {
    scan = std::atoi(argv[2]);
    d0 = std::atoi(argv[3]);
    d1 = std::atoi(argv[4]);
    repeat = std::atoi(argv[5]);
  }
  for (int i = 1; i <= repeat; ++i) {
    const int size = i * 1024;
    if (scan == SAT_SCAN_KOGGE_STONE) dispatch_pair<R, SAT_SCAN_KOGGE_STONE>(d0, d1, size);
    else dispatch_pair<R, SAT_SCAN_LF>(d0, d1, size);
  }
  std::printf("------------------------------------------------------\n");
  std::fflush(stdout);
}

}  // namespace

void TestBlockScan(int argc, char** argv) { dispatch<RunBlock>(argc, argv); }
void TestSerielScan(int argc, char** argv) { dispatch<RunSeriel>(argc, argv); }
void TestIncreamentScan(int argc, char** argv) { dispatch<RunIncrement>(argc, argv); }
int TestFailures() { return g_failures.load(); }
