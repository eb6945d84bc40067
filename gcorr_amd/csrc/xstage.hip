// X-stage (FP32) for gfx950, plus the row-count and finalise kernels.
//
// Cross-multiply-accumulate of fxkernel.cpp:327-356 (and the gcorr X-engines,
// gcorr/src/gxkernel.cu:353-585), reorganised for CDNA:
//   - stations are grouped in blocks of SB; one wave owns one block pair
//     (bi <= bj) and 64 consecutive channels (one per lane), so each lane
//     keeps (2*SB)^2 complex accumulators in registers and every spectrum
//     load feeds 2*SB complex multiply-adds;
//   - the FFT axis is split into `splits` ranges (split-K) so the grid fills
//     all 256 CUs; partial sums go to a small buffer that the finalise kernel
//     reduces in a fixed order (deterministic), then divides by the number of
//     FFTs in which both stations were valid (fxkernel.cpp:352,358-367).
// Invalid (station, FFT) spectra are written as zeros by the F-stage, so the
// products they contribute are exactly zero and need no masking here.
#include "gxc_internal.hpp"

namespace gxc {
namespace {

__device__ __forceinline__ void cmac_conj(float2 &acc, float2 x, float2 y) {
  // acc += x * conj(y)  (vectorAddProduct_cf32 with conjChannels, fxkernel.cpp:349,587)
  acc.x = fmaf(x.x, y.x, acc.x);
  acc.x = fmaf(x.y, y.y, acc.x);
  acc.y = fmaf(x.y, y.x, acc.y);
  acc.y = fmaf(-x.x, y.y, acc.y);
}

__device__ __forceinline__ void tile_of(int tile, int nb, int &bi, int &bj) {
  // Upper-triangle enumeration of block pairs: (0,0),(0,1)..(0,nb-1),(1,1)..
  int r = 0, rem = tile;
  while (rem >= nb - r) {
    rem -= nb - r;
    r++;
  }
  bi = r;
  bj = r + rem;
}

template <int SB>
__global__ void __launch_bounds__(512) xstage_kernel(XParams p) {
  constexpr int NI = 2 * SB;  // inputs per station block
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tile = p.tile0 + blockIdx.z * (blockDim.x >> 6) + wave;
  if (tile >= p.ntiles) return;  // whole wave exits; no block barriers below
  int bi, bj;
  tile_of(tile, p.nb, bi, bj);
  const int c = blockIdx.x * 64 + lane;
  if (c >= p.nchan) return;
  const int split = blockIdx.y;
  const long long tb = (long long)p.nt * split / p.splits;
  const long long te = (long long)p.nt * (split + 1) / p.splits;
  const int M = 2 * p.nant;
  // Inputs of block bi (rows) and bj (columns); stations past nant read as zero.
  int ri[NI], rj[NI];
#pragma unroll
  for (int q = 0; q < NI; q++) {
    const int sti = bi * SB + q / 2, stj = bj * SB + q / 2;
    ri[q] = sti < p.nant ? 2 * sti + (q & 1) : -1;
    rj[q] = stj < p.nant ? 2 * stj + (q & 1) : -1;
  }
  float2 acc[NI][NI];
#pragma unroll
  for (int x = 0; x < NI; x++)
#pragma unroll
    for (int y = 0; y < NI; y++) acc[x][y] = make_float2(0.f, 0.f);

  const size_t rowstride = (size_t)p.nchan;
  const size_t tstride = (size_t)M * p.nchan;
  const float2 *base = p.spectra + c;
  for (long long t = tb; t < te; t++) {
    const float2 *bt = base + (size_t)t * tstride;
    float2 xi[NI], xj[NI];
#pragma unroll
    for (int q = 0; q < NI; q++) {
      xi[q] = ri[q] >= 0 ? bt[(size_t)ri[q] * rowstride] : make_float2(0.f, 0.f);
      xj[q] = rj[q] >= 0 ? bt[(size_t)rj[q] * rowstride] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int x = 0; x < NI; x++)
#pragma unroll
      for (int y = 0; y < NI; y++) cmac_conj(acc[x][y], xi[x], xj[y]);
  }
  float2 *out = p.partial + (((size_t)split * p.ntiles + tile) * NI * NI) * p.nchan + c;
#pragma unroll
  for (int x = 0; x < NI; x++)
#pragma unroll
    for (int y = 0; y < NI; y++) {
      float2 *o = out + (size_t)(x * NI + y) * p.nchan;
      float2 v = acc[x][y];
      if (p.accumulate) {
        const float2 old = *o;
        v.x += old.x;
        v.y += old.y;
      }
      *o = v;
    }
}

__global__ void __launch_bounds__(256) counts_kernel(CountParams p) {
  // baselineCount (fxkernel.cpp:352): FFTs in which both stations were valid.
  const int row = blockIdx.x;
  const int2 jk = p.rowmap[row];
  const uint8_t *vj = p.valid + (size_t)jk.x * p.numffts;
  const uint8_t *vk = p.valid + (size_t)jk.y * p.numffts;
  int n = 0;
  for (int t = threadIdx.x; t < p.numffts; t += blockDim.x) n += (vj[t] & vk[t]) ? 1 : 0;
  __shared__ int red[256];
  red[threadIdx.x] = n;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.counts[row] = red[0];
}

__global__ void __launch_bounds__(256) finalise_kernel(FinParams p) {
  // Sum split partials in a fixed order, normalise by the row count
  // (fxkernel.cpp:358-367; rows with a zero count are left unnormalised,
  // which for all-zero partials leaves zeros).
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int prod = blockIdx.y;
  const int row = blockIdx.z;
  if (c >= p.nchan) return;
  const int2 jk = p.rowmap[row];
  const int l = prod >> 1, m = prod & 1;
  const int bi = jk.x / p.sb, bj = jk.y / p.sb;
  const int tile = bi * p.nb - bi * (bi - 1) / 2 + (bj - bi);
  const int ni = 2 * p.sb;
  const int x = 2 * (jk.x % p.sb) + l, y = 2 * (jk.y % p.sb) + m;
  const size_t per_split = (size_t)p.ntiles * ni * ni * p.nchan;
  const float2 *src = p.partial + ((size_t)tile * ni * ni + (size_t)x * ni + y) * p.nchan + c;
  float2 s = make_float2(0.f, 0.f);
  for (int k = 0; k < p.splits; k++) {
    const float2 v = src[(size_t)k * per_split];
    s.x += v.x;
    s.y += v.y;
  }
  const int cnt = p.counts[row];
  if (cnt != 0) {
    const float n = (float)cnt;
    s.x = s.x / n;
    s.y = s.y / n;
  }
  p.vis[((size_t)row * 4 + prod) * p.nchan + c] = s;
}

}  // namespace

int xstage_block_stations(int nant) {
  // Station block size: large enough for register reuse, small enough to
  // keep (2*SB)^2 accumulators per lane within the register budget.
  if (nant <= 2) return 1;
  if (nant <= 8) return 2;
  return 2;
}

hipError_t launch_xstage(const XParams &p, int sb, hipStream_t s) {
  const int waves_per_block = p.ntiles < 8 ? p.ntiles : 8;
  const int tiles_left = p.ntiles - p.tile0;
  const int zgroups = (tiles_left + waves_per_block - 1) / waves_per_block;
  dim3 grid((p.nchan + 63) / 64, p.splits, zgroups);
  dim3 block(64 * waves_per_block);
  switch (sb) {
    case 1: hipLaunchKernelGGL(xstage_kernel<1>, grid, block, 0, s, p); break;
    case 2: hipLaunchKernelGGL(xstage_kernel<2>, grid, block, 0, s, p); break;
    case 3: hipLaunchKernelGGL(xstage_kernel<3>, grid, block, 0, s, p); break;
    case 4: hipLaunchKernelGGL(xstage_kernel<4>, grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_counts(const CountParams &p, hipStream_t s) {
  if (p.nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(counts_kernel, dim3(p.nrows), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_finalise(const FinParams &p, hipStream_t s) {
  if (p.nrows <= 0) return hipSuccess;
  dim3 grid((p.nchan + 255) / 256, 4, p.nrows);
  hipLaunchKernelGGL(finalise_kernel, grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace gxc
