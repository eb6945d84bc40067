// Multi-subband F-stage: all subbands of a subintegration in one persistent
// launch.  Subbands differ in FFT length, bandwidth, delays and input; the
// workgroup size is common (256 or 512 threads), so a unit of work is one
// workgroup's worth of items (station, FFT) of one subband, block/(L/16)
// items at FFT length L, or at 256 threads and L = 8192 one (station, FFT,
// polarisation) through the register/plane F-stage.  Units are strided
// over a grid sized to fill the device once; each unit runs the same item
// code as the per-subband launch (fstage_item.hpp).
#include "fstage_big.hpp"
#include "fstage_item.hpp"

namespace gxc {
namespace {

using namespace fs;

// LDS of the largest unit of a launch with workgroups of BLOCK threads:
// FFT lengths 256..8192 at 512 threads; at 256 threads 256..4096 with
// fstage_item and 8192 one polarisation per unit with fstage_big_item.
template <int BLOCK>
struct MultiLds {
  static constexpr int max_n() {
    int m = BLOCK == 256 ? BLds<8192, false>::N : 0;
    for (int L = 256; L <= (BLOCK == 512 ? 8192 : 4096); L *= 2) {
      const int n = flds_slots(L, false, BLOCK);
      m = n > m ? n : m;
    }
    return m;
  }
  static constexpr int N = max_n();
  static_assert(N * 8 <= 160 * 1024, "fits one CU's LDS");
};

// One unit at FFT length L, kept out of line: each length gets its own
// register allocation instead of one merged over every length.
template <int L, int OUT, int BLOCK>
__device__ __attribute__((noinline)) void multi_unit(const FSeg *g, int blk, float2 *smem) {
  if constexpr (L == 8192 && BLOCK == 256)
    fstage_big_item<8192, false, OUT>(g->p, blk, smem);  // blk = 2 * item + pol
  else
    fstage_item<L, false, OUT, BLOCK>(g->p, blk, smem);
}

template <int OUT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) fstage_multi_kernel(const FSeg *segs, int nseg, long long total) {
  __shared__ float2 smem[MultiLds<BLOCK>::N];
  // Units in plain order: their cost differs by subband (FFT length), and
  // striding them over the grid spreads every subband over every XCD.
  for (long long u = blockIdx.x; u < total; u += gridDim.x) {
    int k = 0;
    while (k + 1 < nseg && u >= segs[k + 1].unit0) ++k;
    const FSeg &g = segs[k];
    const int blk = (int)(u - g.unit0);
    __syncthreads();  // the previous unit's LDS reads are done
    switch (g.L) {
      case 256: multi_unit<256, OUT, BLOCK>(&g, blk, smem); break;
      case 512: multi_unit<512, OUT, BLOCK>(&g, blk, smem); break;
      case 1024: multi_unit<1024, OUT, BLOCK>(&g, blk, smem); break;
      case 2048: multi_unit<2048, OUT, BLOCK>(&g, blk, smem); break;
      case 4096: multi_unit<4096, OUT, BLOCK>(&g, blk, smem); break;
      default:
        if (g.L == 8192) multi_unit<8192, OUT, BLOCK>(&g, blk, smem);
        break;
    }
  }
}

template <int BLOCK>
hipError_t launch_multi(const FSeg *segs, int nseg, long long total, int out_mode, hipStream_t s) {
  const void *fn = nullptr;
  switch (out_mode) {
    case OUT_F32: fn = reinterpret_cast<const void *>(&fstage_multi_kernel<OUT_F32, BLOCK>); break;
    case OUT_F16: fn = reinterpret_cast<const void *>(&fstage_multi_kernel<OUT_F16, BLOCK>); break;
    case OUT_BF16: fn = reinterpret_cast<const void *>(&fstage_multi_kernel<OUT_BF16, BLOCK>); break;
    case OUT_FP8: fn = reinterpret_cast<const void *>(&fstage_multi_kernel<OUT_FP8, BLOCK>); break;
    case OUT_POWER: fn = reinterpret_cast<const void *>(&fstage_multi_kernel<OUT_POWER, BLOCK>); break;
    default: return hipErrorInvalidValue;
  }
  const long long grid = persistent_grid(fn, BLOCK, 0, total);
  void *args[] = {(void *)&segs, (void *)&nseg, (void *)&total};
  return hipLaunchKernel(fn, dim3((unsigned)grid), dim3(BLOCK), args, 0, s);
}

}  // namespace

bool fstage_multi_supported(int L, bool cplx) { return !cplx && L >= 256 && L <= 8192 && (L & (L - 1)) == 0; }

int fstage_multi_block(int maxL) {
  // 256 covers every length (8192 as one-polarisation units), but those
  // units then share the occupancy of the 4096-point items (2 workgroups
  // per CU); 512 with whole 8192-point items measured faster.
  return maxL >= 8192 ? 512 : 256;
}

long long fstage_multi_units(int L, int block, long long items) {
  if (L == 8192 && block == 256) return 2 * items;
  const long long ipw = block / (L / 16);
  return (items + ipw - 1) / ipw;
}

hipError_t launch_fstage_multi(const FSeg *segs, int nseg, long long total, int out_mode, int block,
                               hipStream_t s) {
  if (total <= 0) return hipSuccess;
  if (block == 512) return launch_multi<512>(segs, nseg, total, out_mode, s);
  if (block == 256) return launch_multi<256>(segs, nseg, total, out_mode, s);
  return hipErrorInvalidValue;
}

}  // namespace gxc
