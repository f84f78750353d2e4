// Column-scan pass kernels (IncrementScan::ScanY) and the transpose-then-
// serial-scan pass (SerielScan::serielScan).
#include "sat_launch.hpp"

namespace sathost {

using namespace satdev;

template <typename S, typename D, int SC>
struct FamCols {
  using Op = TileOp<S, D, MODE_COL, false, false, SC, kNW, kRPT<D>>;
  static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) { return launch_op<Op>(L, ws, b, st); }
};

// The serial pass has no scan-order choice (BRLT-Scan.cu:14): both SC values
// map to the same kernel.
template <typename S, typename D, int SC>
struct FamSerialT {
  using Op = SerialTileOp<S, D, kNW_S, kSerialCW<D>>;
  static sat_status go(const Launch& L, void* ws, size_t b, hipStream_t st) { return launch_op<Op>(L, ws, b, st); }
};

sat_status launch_fam_cols(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return dispatch_pairs<FamCols>(L, ws, b, st);
}
sat_status launch_fam_serial_t(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return dispatch_pairs<FamSerialT>(L, ws, b, st);
}

}  // namespace sathost
