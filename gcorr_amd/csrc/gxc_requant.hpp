// Requantiser kernels (requant.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gxc {

// n real samples per polarisation (even); a, b 16-byte aligned; out (n/2
// bytes) 8-byte aligned.
hipError_t launch_requant2(const float *a, const float *b, long long n, float t, uint8_t *out, hipStream_t s);
// n complex samples per polarisation (re, im interleaved); a, b 16-byte
// aligned; out (4n bytes) 16-byte aligned.
hipError_t launch_requant8(const float *a, const float *b, long long n, float scale, uint8_t *out, hipStream_t s);

}  // namespace gxc
