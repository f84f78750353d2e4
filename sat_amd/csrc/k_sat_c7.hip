// One-pass SAT kernels (forward and reverse) in tile configuration 7.
#include "sat_launch.hpp"

namespace sathost {

sat_status launch_fam_sat_c7(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return launch_sat_cfg<7>(L, ws, b, st);
}

sat_status launch_fam_sat_rev_c7(const Launch& L, void* ws, size_t b, hipStream_t st) {
  return launch_sat_rev_cfg<7>(L, ws, b, st);
}

}  // namespace sathost
