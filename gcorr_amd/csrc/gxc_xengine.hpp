// X-engine on MFMA for FP32, FP16, BF16 and FP8 (e4m3) spectra.
#pragma once

#include <cstddef>
#include <cstdint>

#include "gxc_internal.hpp"

namespace gxc {

// Partial Gram sums of one split: [nchan][2][M][M] floats, plane 0 the real
// part Re(V) (upper block triangle), plane 1 G_IR[a][b] = sum Im(x_a) Re(x_b)
// (all entries), so Im V_ab = G_IR[a][b] - G_IR[b][a].  FP32 blocks off the
// diagonal hold Im V_ab itself at [a][b] (a's block before b's) and 0 at
// [b][a], which the same difference returns.
struct XMParams {
  const void *spectra;  // [nt][M][nchan] complex (2 x 32, 16 or 8 bits)
  float *partial;       // [splits][nchan][2][M][M]
  float *comp;          // the same shape: compensation of partial across
                        // chunks (Neumaier), or null (one chunk)
  int nt, M, nchan, splits;
  int accumulate;       // add into partial instead of overwriting
};

struct FinMParams {
  const float *partial;  // [splits][nchan][2][M][M]
  const float *comp;     // compensation of partial, or null
  const float *scales;   // [M][nchan], or null for unscaled spectra
  const int2 *rowmap;    // [nrows] station pairs
  const int *counts;     // [nrows]
  float2 *vis;           // [nrows][4][nchan]
  int M, nchan, nrows, splits;
};

// One subband's share of a multi-subband launch.  Units of segment k are
// unit0[k] .. unit0[k+1]-1; segments are in increasing unit0.
struct XSeg {  // units = (nchan / channel tile) * splits
  XMParams p;
  long long unit0;
};
struct FinSeg {  // units = nchan
  FinMParams p;
  long long unit0;
};
struct ScaleSeg {  // units = ceil(n / 256)
  float *scales;
  const float *power;
  int n, nrows;
  float target;
  long long unit0;
};

// Inputs per MFMA block (rows of a 16x16x32 tile) and channels per workgroup.
int xm_blocks(int M);                 // ceil(M/16) in {1, 2, 4}; 0 if unsupported
int xm_channels_per_group(int M, int out_mode);  // workgroup channel tile
int xm_choose_splits(int M, int nchan, int nt, int out_mode);
size_t xm_partial_bytes(int M, int nchan, int splits);
int xm_element_bytes(int out_mode);   // bytes of one packed complex value

// ct: channel tile, one of xm_tiles (0 = default)
hipError_t launch_xmfma(const XMParams &p, int out_mode, int ct, hipStream_t s);
// Channel tiles available for M inputs (default first); returns the count (<= 3).
int xm_tiles(int M, int out_mode, int *out);
// scales[i] = power of two nearest target / sqrt(mean_r power[r][i]), r < nrows.
hipError_t launch_scales(float *scales, const float *power, int n, int nrows, float target, hipStream_t s);
constexpr int kPowerRows = 64;  // FFTs sampled for the scale estimate
hipError_t launch_finalise_mixed(const FinMParams &p, hipStream_t s);
// Polarisation conversion of finished visibilities [nrows][4][nchan]:
// V' = J_j V J_k^H with J = [[g0, -g1], [g0, g1]] from gains [nant][2][nchan].
hipError_t launch_polconv(float2 *vis, const float2 *gains, const int2 *rowmap, int nrows, int nchan,
                          hipStream_t s);
// acc[i] += v[i] for n floats.
hipError_t launch_add(float *acc, const float *v, size_t n, hipStream_t s);
// Fill n floats with v (a kernel, so it may sit inside a captured graph).
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s);
// Pack FP32 complex spectra [n] with unit scale into the out_mode format.
hipError_t launch_quantise(void *out, const float2 *in, size_t n, int out_mode, hipStream_t s);
// Persistent launches over device-resident segment tables (all subbands of
// one X launch share M).
hipError_t launch_xmfma_multi(const XSeg *segs, int nseg, long long total, int out_mode, int M, hipStream_t s);
hipError_t launch_finalise_multi(const FinSeg *segs, int nseg, long long total, int M, hipStream_t s);
hipError_t launch_scales_multi(const ScaleSeg *segs, int nseg, long long total, hipStream_t s);
// RMS that the per-(input, channel) scales normalise the spectra to.
float xm_target_rms(int out_mode);

}  // namespace gxc
