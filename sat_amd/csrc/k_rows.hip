// Row-scan pass kernels: same layout (IncrementScan::ScanX) and transposed
// output after a block-register-local transpose (BlockScan::blockScan).
#include "sat_launch.hpp"

namespace sathost {

using namespace satdev;

template <typename S, typename D, int SC>
struct FamRows {
  using Op = TileOp<S, D, MODE_ROW, false, false, SC, kNW, kRPT<D>>;
  static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) { return launch_op<Op>(L, ws, b, st); }
};

template <typename S, typename D, int SC>
struct FamRowsT {
  using Op = TileOp<S, D, MODE_ROW, false, true, SC, kNW_T, kRPT_T<D>>;
  static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) { return launch_op<Op>(L, ws, b, st); }
};

sat_status launch_fam_rows(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return dispatch_pairs<FamRows>(L, ws, b, st);
}
sat_status launch_fam_rows_t(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return dispatch_pairs<FamRowsT>(L, ws, b, st);
}

}  // namespace sathost
