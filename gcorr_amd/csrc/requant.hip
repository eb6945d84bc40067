// Requantisers: float voltages -> the packed station layouts the F-stage
// reads.  gpubench packs filterbank output to 2 bits on the GPU
// (examples/chris/bench/gpubench.cu:64-130); here the codes are fxkernel's
// (the LUT of fxkernel/src/fxkernel.cpp:15-28 decodes 0,1,2,3 to -HI,-1,+1,+HI):
//   x < -t -> 0,  -t <= x < 0 -> 1,  0 <= x < t -> 2,  x >= t -> 3  (NaN -> 1),
// packed LSB first [A_t, B_t, A_t+1, B_t+1]; and complex 8-bit
//   q = clamp(rint(x * scale), -127, 127) (NaN -> 0), bytes [Are, Aim, Bre, Bim].
// Both are streaming kernels: 16 (2-bit) or 4 (8-bit) time samples per lane,
// 16-byte loads, one 8- or 16-byte store.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "gxc_requant.hpp"

namespace gxc {
namespace {

__device__ __forceinline__ unsigned code2(float x, float t) {
  unsigned c = 1u;
  if (x >= 0.f) c = 2u;
  if (x >= t) c = 3u;
  if (x < -t) c = 0u;
  return c;
}

__device__ __forceinline__ unsigned q8(float x, float scale) {
  const float y = rintf(x * scale);
  const float z = isnan(y) ? 0.f : fminf(fmaxf(y, -127.f), 127.f);
  return (unsigned)(int)z & 0xffu;
}

// 16 time samples per lane: bytes [8*g, 8*g+8) of the output.
__global__ void requant2_kernel(const float *__restrict__ a, const float *__restrict__ b, long long n, float t,
                                uint8_t *__restrict__ out) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long s0 = g * 16;
  if (s0 >= n) return;
  float xa[16], xb[16];
  if (s0 + 16 <= n) {
    const float4 *pa = reinterpret_cast<const float4 *>(a + s0), *pb = reinterpret_cast<const float4 *>(b + s0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const float4 u = pa[k], v = pb[k];
      xa[4 * k] = u.x, xa[4 * k + 1] = u.y, xa[4 * k + 2] = u.z, xa[4 * k + 3] = u.w;
      xb[4 * k] = v.x, xb[4 * k + 1] = v.y, xb[4 * k + 2] = v.z, xb[4 * k + 3] = v.w;
    }
    unsigned w[2] = {0u, 0u};
#pragma unroll
    for (int k = 0; k < 16; k++) w[k >> 3] |= (code2(xa[k], t) | code2(xb[k], t) << 2) << (4 * (k & 7));
    *reinterpret_cast<uint2 *>(out + 8 * g) = make_uint2(w[0], w[1]);
  } else {  // tail: n is even, one byte per two samples
    for (long long s = s0; s < n; s += 2) {
      const unsigned v = code2(a[s], t) | code2(b[s], t) << 2 | code2(a[s + 1], t) << 4 | code2(b[s + 1], t) << 6;
      out[s >> 1] = (uint8_t)v;
    }
  }
}

// 4 complex time samples per lane: bytes [16*g, 16*g+16) of the output.
__global__ void requant8_kernel(const float *__restrict__ a, const float *__restrict__ b, long long n, float scale,
                                uint8_t *__restrict__ out) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long s0 = g * 4;
  if (s0 >= n) return;
  if (s0 + 4 <= n) {
    const float4 *pa = reinterpret_cast<const float4 *>(a + 2 * s0), *pb = reinterpret_cast<const float4 *>(b + 2 * s0);
    unsigned w[4];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const float4 u = pa[k], v = pb[k];  // two complex samples of each pol
      w[2 * k] = q8(u.x, scale) | q8(u.y, scale) << 8 | q8(v.x, scale) << 16 | q8(v.y, scale) << 24;
      w[2 * k + 1] = q8(u.z, scale) | q8(u.w, scale) << 8 | q8(v.z, scale) << 16 | q8(v.w, scale) << 24;
    }
    *reinterpret_cast<uint4 *>(out + 16 * g) = make_uint4(w[0], w[1], w[2], w[3]);
  } else {
    for (long long s = s0; s < n; s++) {
      const unsigned v = q8(a[2 * s], scale) | q8(a[2 * s + 1], scale) << 8 | q8(b[2 * s], scale) << 16 |
                         q8(b[2 * s + 1], scale) << 24;
      *reinterpret_cast<unsigned *>(out + 4 * s) = v;
    }
  }
}

}  // namespace

hipError_t launch_requant2(const float *a, const float *b, long long n, float t, uint8_t *out, hipStream_t s) {
  const long long groups = (n + 15) / 16;
  if (groups == 0) return hipSuccess;
  hipLaunchKernelGGL(requant2_kernel, dim3((unsigned)((groups + 255) / 256)), dim3(256), 0, s, a, b, n, t, out);
  return hipGetLastError();
}

hipError_t launch_requant8(const float *a, const float *b, long long n, float scale, uint8_t *out, hipStream_t s) {
  const long long groups = (n + 3) / 4;
  if (groups == 0) return hipSuccess;
  hipLaunchKernelGGL(requant8_kernel, dim3((unsigned)((groups + 255) / 256)), dim3(256), 0, s, a, b, n, scale, out);
  return hipGetLastError();
}

}  // namespace gxc
