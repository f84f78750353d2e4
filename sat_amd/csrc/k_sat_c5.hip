// One-pass SAT kernels (forward and reverse) in tile configuration 5: the
// packed tiles of 1-byte inputs into 4-byte outputs.
#include "sat_launch.hpp"

namespace sathost {

sat_status launch_fam_sat_c5(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return launch_sat_cfg<5>(L, ws, b, st);
}

sat_status launch_fam_sat_rev_c5(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return launch_sat_rev_cfg<5>(L, ws, b, st);
}

}  // namespace sathost
