// Per-subband F-stage launch: one workgroup per group of items (station,
// FFT), FFT lengths 64..8192 through fstage_item (fstage_item.hpp); longer
// FFTs go to fstage_big.hip.
#include "fstage_item.hpp"

namespace gxc {
namespace {

using namespace fs;

template <int L, bool CPLX, int OUT>
__global__ void __launch_bounds__((FG<L, CPLX>::DEFAULT_BLOCK)) fstage_kernel(FParams p) {
  constexpr int BLOCK = FG<L, CPLX>::DEFAULT_BLOCK;
  __shared__ float2 smem[FLds<L, CPLX, BLOCK>::N];
  fstage_item<L, CPLX, OUT, BLOCK>(p, blockIdx.x, smem);
}

template <int L, bool CPLX>
hipError_t launch_one(const FParams &p, hipStream_t s) {
  constexpr int IPW = FLds<L, CPLX, FG<L, CPLX>::DEFAULT_BLOCK>::IPW;
  const long long items = (long long)p.nant * p.nt;
  const long long blocks = (items + IPW - 1) / IPW;
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks), block(FG<L, CPLX>::DEFAULT_BLOCK);
  switch (p.out_mode) {
    case OUT_F32: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_F32>), grid, block, 0, s, p); break;
    case OUT_F16: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_F16>), grid, block, 0, s, p); break;
    case OUT_BF16: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_BF16>), grid, block, 0, s, p); break;
    case OUT_FP8: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_FP8>), grid, block, 0, s, p); break;
    case OUT_POWER: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_POWER>), grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

bool fstage_supported(int L, bool cplx) {
  (void)cplx;
  return L >= 64 && L <= 32768 && (L & (L - 1)) == 0;
}

hipError_t launch_fstage(const FParams &p, int L, bool cplx, hipStream_t s) {
#define GXC_F_CASE(LL)                                              \
  case LL:                                                          \
    return cplx ? launch_one<LL, true>(p, s) : launch_one<LL, false>(p, s);
  switch (L) {
    GXC_F_CASE(64)
    GXC_F_CASE(128)
    GXC_F_CASE(256)
    GXC_F_CASE(512)
    GXC_F_CASE(1024)
    GXC_F_CASE(2048)
    GXC_F_CASE(4096)
    GXC_F_CASE(8192)
    default:
      return launch_fstage_big(p, L, cplx, s);
  }
#undef GXC_F_CASE
}

}  // namespace gxc
