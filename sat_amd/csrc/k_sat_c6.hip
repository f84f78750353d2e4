// One-pass SAT kernels (forward and reverse) in tile configuration 6.
#include "sat_launch.hpp"

namespace sathost {

sat_status launch_fam_sat_c6(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return launch_sat_cfg<6>(L, ws, b, st);
}

sat_status launch_fam_sat_rev_c6(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return launch_sat_rev_cfg<6>(L, ws, b, st);
}

}  // namespace sathost
