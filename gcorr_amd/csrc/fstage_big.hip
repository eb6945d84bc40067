// Per-subband launch of the register/plane F-stage (fstage_big.hpp): one
// workgroup of L/32 threads per (station, FFT) for real input up to 8192
// points, per (station, FFT, pol) otherwise.  The only F-stage for
// L = 16384, 32768; at L = 2048..8192 an alternative to fstage_item that the
// autotuner weighs.
#include "fstage_big.hpp"

namespace gxc {
namespace {

using namespace fs;

// For real input up to 8192 points a workgroup takes both polarisations of
// one (station, FFT) in turn, so the delay split and the rotator tables are
// built once for two FFTs; the two FFTs fit 128 registers (4 waves per
// SIMD).  At 16384 and 32768 points (512 and 1024 threads), and for complex
// input, whose two copies hold full 32-bit sample words, the second copy
// spills, and a workgroup takes one polarisation.
template <int L, bool CPLX>
constexpr bool big_both() { return !CPLX && L <= 8192; }

template <int L, bool CPLX, int OUT>
__global__ void __launch_bounds__((BG<L, CPLX>::T)) __attribute__((amdgpu_waves_per_eu(4)))
fstage_big_kernel(FParams p) {
  __shared__ float2 smem[BLds<L, CPLX>::N];
  // grid = nant * nt (both polarisations per workgroup) or 2 * nant * nt
  // exactly; in the second form the two polarisations of one (station,
  // FFT) are consecutive units on one XCD, so the second finds the input
  // words in L2
  fstage_big_item<L, CPLX, OUT, big_both<L, CPLX>()>(p, (int)xcd_order(blockIdx.x, gridDim.x), smem);
}

template <int L, bool CPLX>
hipError_t launch_big(const FParams &p, hipStream_t s) {
  using G = BG<L, CPLX>;
  const long long items = (big_both<L, CPLX>() ? 1LL : 2LL) * p.nant * p.nt;
  if (items <= 0) return hipSuccess;
  if (items > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)items), block(G::T);
  switch (p.out_mode) {
    case OUT_F32: hipLaunchKernelGGL((fstage_big_kernel<L, CPLX, OUT_F32>), grid, block, 0, s, p); break;
    case OUT_F16: hipLaunchKernelGGL((fstage_big_kernel<L, CPLX, OUT_F16>), grid, block, 0, s, p); break;
    case OUT_BF16: hipLaunchKernelGGL((fstage_big_kernel<L, CPLX, OUT_BF16>), grid, block, 0, s, p); break;
    case OUT_FP8: hipLaunchKernelGGL((fstage_big_kernel<L, CPLX, OUT_FP8>), grid, block, 0, s, p); break;
    case OUT_POWER: hipLaunchKernelGGL((fstage_big_kernel<L, CPLX, OUT_POWER>), grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_fstage_big(const FParams &p, int L, bool cplx, hipStream_t s) {
  switch (L) {
    case 2048: return cplx ? launch_big<2048, true>(p, s) : launch_big<2048, false>(p, s);
    case 4096: return cplx ? launch_big<4096, true>(p, s) : launch_big<4096, false>(p, s);
    case 8192: return cplx ? launch_big<8192, true>(p, s) : launch_big<8192, false>(p, s);
    case 16384: return cplx ? launch_big<16384, true>(p, s) : launch_big<16384, false>(p, s);
    case 32768: return cplx ? launch_big<32768, true>(p, s) : launch_big<32768, false>(p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gxc
