// Per-subband F-stage launch for FFT lengths 64..8192 at a chosen workgroup
// size (items per workgroup = BLOCK / (L/16)).  Instantiated by fstage.hip
// (default size), fstage_half.hip and fstage_double.hip (the alternatives the
// autotuner times), each in its own translation unit.
#pragma once

#include "fstage_item.hpp"

namespace gxc {
namespace fs {

// Workgroup size at multiple MULT/2 of the default (MULT = 1, 2, 4), or 0 if
// that size is not valid for L: whole items, 64..1024 threads, LDS <= 160 KiB.
constexpr int fblock_for(int L, bool cplx, int mult) {
  const int T = L / 16;
  const int def = T >= 256 ? T : 256;
  const int b = def * mult / 2;
  if (b < T || b < 64 || b > 1024 || b % T) return 0;
  if ((long long)flds_slots(L, cplx, b) * 8 > 160 * 1024) return 0;
  return b;
}

template <int L, bool CPLX, int OUT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) fstage_kernel(FParams p) {
  __shared__ float2 smem[FLds<L, CPLX, BLOCK>::N];
  fstage_item<L, CPLX, OUT, BLOCK>(p, blockIdx.x, smem);
}

template <int L, bool CPLX, int BLOCK>
hipError_t launch_one(const FParams &p, hipStream_t s) {
  if constexpr (BLOCK == 0) {
    return hipErrorInvalidValue;
  } else {
    constexpr int IPW = FLds<L, CPLX, BLOCK>::IPW;
    const long long items = (long long)p.nant * p.nt;
    const long long blocks = (items + IPW - 1) / IPW;
    if (blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks), block(BLOCK);
    switch (p.out_mode) {
      case OUT_F32: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_F32, BLOCK>), grid, block, 0, s, p); break;
      case OUT_F16: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_F16, BLOCK>), grid, block, 0, s, p); break;
      case OUT_BF16: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_BF16, BLOCK>), grid, block, 0, s, p); break;
      case OUT_FP8: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_FP8, BLOCK>), grid, block, 0, s, p); break;
      case OUT_POWER: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_POWER, BLOCK>), grid, block, 0, s, p); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
}

template <int MULT>
hipError_t launch_fstage_mult(const FParams &p, int L, bool cplx, hipStream_t s) {
#define GXC_F_CASE(LL)                                                               \
  case LL:                                                                           \
    return cplx ? launch_one<LL, true, fblock_for(LL, true, MULT)>(p, s)             \
                : launch_one<LL, false, fblock_for(LL, false, MULT)>(p, s);
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
      return hipErrorInvalidValue;
  }
#undef GXC_F_CASE
}

}  // namespace fs
}  // namespace gxc
