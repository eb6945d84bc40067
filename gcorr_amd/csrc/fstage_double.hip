// F-stage at double the default workgroup size (more items per workgroup).
#include "fstage_launch.hpp"

namespace gxc {
hipError_t launch_fstage_double(const FParams &p, int L, bool cplx, hipStream_t s) {
  return fs::launch_fstage_mult<4>(p, L, cplx, s);
}
}  // namespace gxc
