// Internal declarations shared by the HIP kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gxc {

constexpr double kPi = 3.14159265358979323846;  // IPP_PI (fxkernel/src/vectordefs.h:8)
constexpr float kHiMag = 3.3359f;                // fxkernel/src/fxkernel.cpp:17

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }

// Geometry shared by the host and the F-stage kernel for FFT length L.
// Generalised stride: s = 2^floor(log2(nchan)/2) (fxkernel.cpp:58-64 only
// accepts square nchan, for which this is sqrt(nchan)).
struct Geometry {
  int L, nchan, s, S2, nsteps, fracsteps;
  static Geometry make(int nchan, bool cplx) {
    Geometry g{};
    g.nchan = nchan;
    g.L = cplx ? nchan : 2 * nchan;
    g.s = 1 << (ilog2c(nchan) / 2);
    g.S2 = cplx ? g.s : 2 * g.s;
    g.nsteps = g.L / g.S2;
    g.fracsteps = nchan / g.s;
    return g;
  }
};

// Arguments of the F-stage kernel (unpack, fringe rotate, FFT, fractional
// sample correction) for FFTs [t0, t0+nt) of every station.
struct FParams {
  const uint8_t *const *station;  // [nant] device pointers to packed input
  const double *delays;           // [nant][4]: d0, d1, d2, file offset (s)
  const float *subchanfreq;       // [s]          fxkernel.cpp:173
  const float *stepchanfreq;      // [fracsteps]  fxkernel.cpp:174
  const float2 *twiddle;          // [L] exp(-2 pi i m / L)
  float2 *spectra;                // [nt][nant*2][nchan]
  uint8_t *valid;                 // [nant][numffts]
  int t0, nt, nant, numffts;
  int maxoffset;                  // (numffts-1)*L
  int stepmul;                    // s (stride law) or S2 (exact phase)
  int fraclo;                     // LO has a fractional part
  double lo, lofrac, sampletime;
};

// Arguments of the FP32 X-stage kernel.
struct XParams {
  const float2 *spectra;  // [nt][2*nant][nchan]
  float2 *partial;        // [splits][ntiles][(2*SB)^2][nchan]
  int nt, nant, nchan;
  int nb, ntiles;         // station blocks and block-pair tiles
  int tile0;              // first tile of this launch's tile group
  int splits;
  int accumulate;         // 1: add into partial, 0: overwrite
};

// Rows of the output: station pair (j,k) per row; autos have j == k.
struct FinParams {
  const float2 *partial;  // X-stage partials
  const int2 *rowmap;     // [nrows]
  const int *counts;      // [nrows]
  float2 *vis;            // [nrows][4][nchan]
  int nrows, nchan, nb, ntiles, splits, sb;
};

struct CountParams {
  const uint8_t *valid;  // [nant][numffts]
  const int2 *rowmap;    // [nrows]
  int *counts;           // [nrows]
  int numffts, nrows;
};

// Launchers (defined in the .hip files).  Return hipSuccess or the launch error.
bool fstage_supported(int L, bool cplx);
hipError_t launch_fstage(const FParams &p, int L, bool cplx, hipStream_t s);
int xstage_block_stations(int nant);
hipError_t launch_xstage(const XParams &p, int sb, hipStream_t s);
hipError_t launch_counts(const CountParams &p, hipStream_t s);
hipError_t launch_finalise(const FinParams &p, hipStream_t s);

}  // namespace gxc
