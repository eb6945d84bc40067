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

// What the F-stage writes per (station, FFT): FP32 spectra, scaled
// reduced-precision spectra (complex: 2 x 16 bits or 2 x 8 bits), or |X|^2
// accumulated per (input, channel) for the scale estimate.
enum OutMode { OUT_F32 = 0, OUT_F16 = 1, OUT_BF16 = 2, OUT_FP8 = 3, OUT_POWER = 4 };

// Arguments of the F-stage kernel (unpack, fringe rotate, FFT, fractional
// sample correction) for FFTs t0 + i*tstride, i in [0, nt), of every station.
struct FParams {
  const uint8_t *const *station;  // [nant] device pointers to packed input
  const double *delays;           // [nant][4]: d0, d1, d2, file offset (s)
  const float *subchanfreq;       // [s]          fxkernel.cpp:173
  const float *stepchanfreq;      // [fracsteps]  fxkernel.cpp:174
  const float2 *twiddle;          // [L] exp(-2 pi i m / L)
  float2 *spectra;                // OUT_F32: [nt][nant*2][nchan]
  void *out;                      // OUT_F16/BF16/FP8: [nt][nant*2][nchan] packed complex
  const float *scales;            // [nant*2][nchan] multipliers before rounding
  float *power;                   // OUT_POWER: [nt][nant*2][nchan] |X|^2 per sampled FFT
  uint8_t *valid;                 // [nant][numffts]
  // Missing VDIF frames (nullptr: none): frame_ok[st][f] != 0 when frame f,
  // samples [f, f+1) * frame_samples[st], is present; an FFT touching a
  // missing frame is invalid for that station.
  const uint8_t *const *frame_ok;
  const int *frame_samples;
  int out_mode;
  int t0, nt, tstride, nant, numffts;
  int maxoffset;                  // (numffts-1)*L
  int stepmul;                    // s (stride law) or S2 (exact phase)
  int fraclo;                     // LO has a fractional part
  double lo, lofrac, sampletime;
};

// Row counts: for each output row (station pair; autos have j == k) the
// number of FFTs in which both stations were valid.
struct CountParams {
  const uint8_t *valid;  // [nant][numffts]
  const int2 *rowmap;    // [nrows]
  int *counts;           // [nrows]
  int numffts, nrows;
};

// One subband's share of a multi-subband launch (units of segment k are
// unit0[k] .. unit0[k+1]-1, segments in increasing unit0).
struct FSeg {  // units = workgroups of the multi F launch for this subband
  FParams p;
  int L;
  long long unit0;
};
struct CountSeg {  // units = nrows
  CountParams p;
  long long unit0;
};

// Workgroups that keep every CU busy for a persistent launch of `kernel`
// (occupancy x CU count), at most `total`.
long long persistent_grid(const void *kernel, int block, size_t dyn_lds, long long total);

// Allow `kernel` dynamic LDS of `bytes` (above the 64 KiB default) on the
// current device, once per (device, kernel, bytes); safe from any thread.
hipError_t allow_dynamic_lds(const void *kernel, int bytes);

// Launchers (defined in the .hip files).  Return hipSuccess or the launch error.
bool fstage_supported(int L, bool cplx);
// F-stage workgroup sizes available at FFT length L (default first).
int fstage_blocks(int L, bool cplx, int *out /* [4] */);
// block: one of fstage_blocks (0 = default); ignored for L > 8192
hipError_t launch_fstage(const FParams &p, int L, bool cplx, int block, hipStream_t s);
hipError_t launch_fstage_big(const FParams &p, int L, bool cplx, hipStream_t s);  // L = 16384, 32768
// F-stage with the FFT passes on the matrix cores (fstage_mfma.hip): real
// input; reduced-precision outputs (and the power pass) at L = 2048..8192,
// FP32 spectra (operands as f16 hi + lo) at L = 2048..4096.
constexpr int kFBlockMfma = -1;
bool fstage_mfma_supported(int L, bool cplx, bool fp32);
hipError_t launch_fstage_mfma(const FParams &p, int L, hipStream_t s);
hipError_t launch_counts(const CountParams &p, hipStream_t s);
// Multi-subband F-stage: real 2-bit input, FFT lengths 256..8192, workgroups
// of `block` threads (256, the default, or 512); units of a subband =
// ceil(nant * nt / (block / (L/16))), or 2 * nant * nt at L = 8192 and 256.
bool fstage_multi_supported(int L, bool cplx);
int fstage_multi_block(int maxL);
long long fstage_multi_units(int L, int block, long long items);
hipError_t launch_fstage_multi(const FSeg *segs, int nseg, long long total, int out_mode, int block,
                               hipStream_t s);
hipError_t launch_counts_multi(const CountSeg *segs, int nseg, long long total, hipStream_t s);

}  // namespace gxc
