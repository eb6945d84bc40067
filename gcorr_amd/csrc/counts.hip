// Row counts for the normalisation of fxkernel.cpp:352,358-367.
//
// The F-stage records per (station, FFT) whether the station's delayed
// window lay inside its buffer (fxkernel.cpp:304-309).  A row (station pair,
// or a station with itself for autocorrelations) counts the FFTs in which
// both stations were valid: fxkernel's baselineCount.
#include "gxc_internal.hpp"

namespace gxc {
namespace {

__global__ void __launch_bounds__(256) counts_kernel(CountParams p) {
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

}  // namespace

hipError_t launch_counts(const CountParams &p, hipStream_t s) {
  if (p.nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(counts_kernel, dim3(p.nrows), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace gxc
