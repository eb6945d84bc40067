// Lag spectra and their signal-to-noise ratio, the fringe check of
// util/lagsnr.py:33-36: for each series of nchan visibilities (one row and
// polarisation product), lags = |ifft(vis)|, S/N = max(lags) / std(lags)
// (population standard deviation) and the peak lag = argmax (first index).
//   lag_pass_kernel  one radix-2 Stockham pass of the inverse transform over
//                    every series at once (log2 nchan launches, ping-pong);
//   lag_snr_kernel   one workgroup per series: |lag|, max with the lowest
//                    index on ties, mean and variance in FP64 (two passes).
// The 1/nchan of ifft is left out: the S/N does not depend on it.
#include <hip/hip_runtime.h>

#include <math.h>

#include "gxc_lags.hpp"

namespace gxc {
namespace {

__global__ void lag_pass_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, int n,
                                long long total, int ns) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int half = n >> 1;
  const long long s = idx / half;
  const int j = (int)(idx - s * half);
  const float2 *x = in + s * n;
  float2 *y = out + s * n;
  const float2 a = x[j], b0 = x[j + half];
  const int k = j & (ns - 1);
  float sn, cs;
  sincospif((float)k / (float)ns, &sn, &cs);  // +pi k / ns: the inverse transform
  const float2 b = make_float2(b0.x * cs - b0.y * sn, b0.x * sn + b0.y * cs);
  const int d = (j - k) * 2 + k;
  y[d] = make_float2(a.x + b.x, a.y + b.y);
  y[d + ns] = make_float2(a.x - b.x, a.y - b.y);
}

constexpr int kSnrThreads = 256;

__global__ void __launch_bounds__(kSnrThreads) lag_snr_kernel(const float2 *__restrict__ lags, int n,
                                                              float *snr, int *peak) {
  __shared__ float smax[kSnrThreads];
  __shared__ int sidx[kSnrThreads];
  __shared__ double ssum[kSnrThreads];
  const float2 *x = lags + (long long)blockIdx.x * n;
  const int t = threadIdx.x;
  float best = -1.f;
  int bi = 0;
  double sum = 0.0;
  for (int i = t; i < n; i += kSnrThreads) {
    const float m = hypotf(x[i].x, x[i].y);
    if (m > best) {
      best = m;
      bi = i;
    }
    sum += m;
  }
  smax[t] = best;
  sidx[t] = bi;
  ssum[t] = sum;
  __syncthreads();
  for (int w = kSnrThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      const float m2 = smax[t + w];
      const int i2 = sidx[t + w];
      if (m2 > smax[t] || (m2 == smax[t] && i2 < sidx[t])) {
        smax[t] = m2;
        sidx[t] = i2;
      }
      ssum[t] += ssum[t + w];
    }
    __syncthreads();
  }
  const double mean = ssum[0] / n;
  const float mx = smax[0];
  const int mi = sidx[0];
  __syncthreads();
  double var = 0.0;
  for (int i = t; i < n; i += kSnrThreads) {
    const double dm = (double)hypotf(x[i].x, x[i].y) - mean;
    var += dm * dm;
  }
  ssum[t] = var;
  __syncthreads();
  for (int w = kSnrThreads / 2; w > 0; w >>= 1) {
    if (t < w) ssum[t] += ssum[t + w];
    __syncthreads();
  }
  if (t == 0) {
    const double sd = sqrt(ssum[0] / n);
    snr[blockIdx.x] = sd > 0.0 ? (float)(mx / sd) : (mx > 0.f ? INFINITY : NAN);
    peak[blockIdx.x] = mi;
  }
}

}  // namespace

hipError_t launch_lag_snr(float2 *a, float2 *b, int n, int nseries, float *snr, int *peak, hipStream_t s) {
  const long long total = (long long)nseries * (n / 2);
  const unsigned blocks = (unsigned)((total + 255) / 256);
  float2 *src = a, *dst = b;
  for (int ns = 1; ns < n; ns <<= 1) {
    hipLaunchKernelGGL(lag_pass_kernel, dim3(blocks), dim3(256), 0, s, src, dst, n, total, ns);
    float2 *tmp = src;
    src = dst;
    dst = tmp;
  }
  hipLaunchKernelGGL(lag_snr_kernel, dim3((unsigned)nseries), dim3(kSnrThreads), 0, s, src, n, snr, peak);
  return hipGetLastError();
}

}  // namespace gxc
