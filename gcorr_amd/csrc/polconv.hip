// Polarisation conversion on the visibilities.  gpubench's
// linear2circular_kernel (examples/chris/bench/gpubench.cu:141-166) converts
// each station's channelised spectra before the X-engine:
//   y0 = g0(c) x0 - g1(c) x1,  y1 = g0(c) x0 + g1(c) x1.
// Per station and channel that is a fixed 2x2 matrix J = [[g0, -g1], [g0, g1]],
// and a fixed matrix commutes with the sum over FFTs:
//   sum_t (J_j x_j)(J_k x_k)^H = J_j (sum_t x_j x_k^H) J_k^H,
// so it is applied once to the finished visibilities, V' = J_j V J_k^H per
// (row, channel), instead of to every spectrum: no extra pass over the
// spectra in HBM.
#include <hip/hip_runtime.h>

#include "gxc_xengine.hpp"

namespace gxc {
namespace {

__device__ __forceinline__ float2 cm(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 ca(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

__global__ void polconv_kernel(float2 *vis, const float2 *__restrict__ gains, const int2 *__restrict__ rowmap,
                               int nrows, int nchan) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nrows * nchan) return;
  const int r = (int)(i / nchan), c = (int)(i - (long long)r * nchan);
  const int2 jk = rowmap[r];
  float2 *v = vis + (size_t)r * 4 * nchan + c;
  const float2 V[2][2] = {{v[0], v[nchan]}, {v[2 * nchan], v[3 * nchan]}};  // p = 2l + m
  const float2 gj0 = gains[((size_t)jk.x * 2) * nchan + c], gj1 = gains[((size_t)jk.x * 2 + 1) * nchan + c];
  const float2 gk0 = gains[((size_t)jk.y * 2) * nchan + c], gk1 = gains[((size_t)jk.y * 2 + 1) * nchan + c];
  const float2 mgj1 = make_float2(-gj1.x, -gj1.y), mgk1 = make_float2(-gk1.x, -gk1.y);
  const float2 Jj[2][2] = {{gj0, mgj1}, {gj0, gj1}};
  const float2 Jk[2][2] = {{gk0, mgk1}, {gk0, gk1}};
  // W = J_j V, then V' = W J_k^H
  float2 W[2][2];
#pragma unroll
  for (int l = 0; l < 2; l++)
#pragma unroll
    for (int q = 0; q < 2; q++) W[l][q] = ca(cm(Jj[l][0], V[0][q]), cm(Jj[l][1], V[1][q]));
#pragma unroll
  for (int l = 0; l < 2; l++)
#pragma unroll
    for (int m = 0; m < 2; m++) v[(2 * l + m) * nchan] = ca(cmc(W[l][0], Jk[m][0]), cmc(W[l][1], Jk[m][1]));
}

}  // namespace

hipError_t launch_polconv(float2 *vis, const float2 *gains, const int2 *rowmap, int nrows, int nchan,
                          hipStream_t s) {
  const long long n = (long long)nrows * nchan;
  hipLaunchKernelGGL(polconv_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vis, gains, rowmap, nrows,
                     nchan);
  return hipGetLastError();
}

}  // namespace gxc
