// Lag S/N kernels (lags.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace gxc {

// a holds nseries series of n (a power of two, >= 2) visibilities and is
// overwritten; b is scratch of the same size.  snr[nseries], peak[nseries]
// are device arrays.
hipError_t launch_lag_snr(float2 *a, float2 *b, int n, int nseries, float *snr, int *peak, hipStream_t s);

}  // namespace gxc
