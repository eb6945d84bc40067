// F-stage at half the default workgroup size (fewer items per workgroup).
#include "fstage_launch.hpp"

namespace gxc {
hipError_t launch_fstage_half(const FParams &p, int L, bool cplx, hipStream_t s) {
  return fs::launch_fstage_mult<1>(p, L, cplx, s);
}
}  // namespace gxc
