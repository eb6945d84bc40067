/*
 * fxref.h -- CPU restatement of fxkernel's per-subintegration FX correlation.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or the timed CPU baseline; the
 * product library (gcorr_amd/libgxcorr.so) never links or calls it.
 *
 * The reference (fxkernel/src/fxkernel.cpp) needs Intel IPP, which is not
 * present, so this is an IPP-free restatement written from the reference
 * source.  Every function cites the reference lines it follows.  IPP's own
 * rounding (FFT, ippsSinCos_32f_A11) cannot be reproduced: libm sincosf and
 * an own Stockham FFT stand in for them.
 *
 * Two arithmetic modes share one source (fxref_core.h):
 *   _f32  fxkernel-compatible: FP64 delay/phase model, FP32 arguments,
 *         sincosf, FP32 FFT, FP32 sequential accumulation.
 *   _f64  FP64 oracle of the same algorithm (every step in double).
 *
 * Extensions beyond fxkernel (each off unless requested):
 *   autos        autocorrelations appended after the cross baselines
 *                (fxkernel.cpp:327 has none).
 *   is_complex   complex 8-bit input [Are,Aim,Bre,Bim] per time sample
 *                (fxkernel.cpp:478-496 has an unused, mis-offset unpacker).
 *   quirk=0      exact per-sample fringe phase instead of fxkernel's
 *                stride law (fxkernel.cpp:105-114,553-558).
 *   nchan that is not a perfect square: stride s = 2^floor(log2(nchan)/2),
 *                nchan/s steps (fxkernel.cpp:58-63 refuses such nchan).
 *   polgains     per-spectrum polarisation conversion (gpubench.cu:141-166).
 *   frame_ok     missing-frame flags from a VDIF ingest: an FFT whose input
 *                samples touch a missing frame is invalid for its station
 *                (fxkernel has no missing data; gcorr's ingest,
 *                ingest/src/ingest.c:267-355, places frames but flags
 *                nothing).
 */
#ifndef FXREF_H
#define FXREF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nant;       /* stations (each dual polarisation) */
  int nchan;      /* output channels, power of two */
  int numffts;    /* FFTs per subintegration */
  int nbit;       /* 2 (real data) or 8 (complex data) */
  int is_complex; /* 0: real samples, 1: complex samples */
  double lo;      /* local oscillator, Hz */
  double bw;      /* bandwidth, Hz */
  int autos;      /* 1: append autocorrelations after the cross baselines */
  int quirk;      /* 1: fxkernel stride phase law (default), 0: exact law */
  /* Missing-frame flags (NULL: none): frame_ok[station][f] != 0 when frame
   * f, samples [f*frame_samples[station], (f+1)*frame_samples[station]),
   * is present. */
  const uint8_t *const *frame_ok;
  const int *frame_samples;
  /* Polarisation conversion of every spectrum (NULL: none), as
   * examples/chris/bench/gpubench.cu:141-166 applies it after the
   * filterbank: y0 = g0 x0 - g1 x1, y1 = g0 x0 + g1 x1 per channel, with
   * polgains [nant][2][nchan] complex (interleaved re, im). */
  const float *polgains;
} fxref_config;

typedef struct {
  int fftchannels;   /* FFT length L: 2*nchan real, nchan complex */
  int stride;        /* s (fxkernel stridesize) */
  int substride;     /* S2 = (2/cfact)*s (fxkernel substridesize) */
  int nsteps;        /* L / S2 coarse rotator steps */
  int fracsteps;     /* nchan / s coarse frac-sample steps */
  int maxoffset;     /* (numffts-1)*L */
  int nbaselines;    /* nant*(nant-1)/2 */
  int nrows;         /* nbaselines + (autos ? nant : 0) */
  double sampletime; /* seconds */
  size_t bytes_per_station; /* (numffts+1) FFTs of packed input */
} fxref_derived;

/* Derived constants; returns 0 on success, nonzero for an invalid config. */
int fxref_derive(const fxref_config *c, fxref_derived *d);

/* 2-bit LUT: lut[byte][i] = level of bits (2i, 2i+1), fxkernel.cpp:15-28. */
void fxref_lut2bit(float lut[256][4]);

/* Delay polynomial at FFT start/middle/end -> (mean, a, b); fxkernel.cpp:425-440. */
void fxref_station_delay(const double poly[3], int fftindex, double *mean,
                         double *a, double *b);

/* Integer/fractional split and validity for one (station, FFT); fxkernel.cpp:288-309.
 * Returns 1 if valid, 0 otherwise; writes the sample offset and fractional
 * delay (seconds) and the linear delay model a, b. */
int fxref_delay_split(const fxref_config *c, const double poly[3],
                      double fileoffset, int fftindex, int *offset,
                      double *fracdelay, double *a, double *b);

/* Unpack L samples of both pols at sample offset; out is [2][L] complex
 * (interleaved re,im).  Real 2-bit: fxkernel.cpp:443-476.  Complex 8-bit:
 * layout [Are,Aim,Bre,Bim] per time sample, byte offset 4*offset. */
void fxref_unpack_f32(const fxref_config *c, const uint8_t *data, int offset,
                      float *out);
void fxref_unpack_f64(const fxref_config *c, const uint8_t *data, int offset,
                      double *out);

/* Fringe rotator for delay model (a, b); out is [L] complex. fxkernel.cpp:498-567 */
void fxref_rotator_f32(const fxref_config *c, double a, double b, float *out);
void fxref_rotator_f64(const fxref_config *c, double a, double b, double *out);

/* Fractional-sample rotator [nchan] complex; fxkernel.cpp:595-634 (and the
 * channel-frequency tables of fxkernel.cpp:159-175). */
void fxref_fracrot_f32(const fxref_config *c, double fracdelay, float *out);
void fxref_fracrot_f64(const fxref_config *c, double fracdelay, double *out);

/* Forward unnormalised complex FFT of power-of-two length L (IPP
 * ippsFFTFwd_CToC_32fc with IPP_FFT_NODIV_BY_ANY, fxkernel.cpp:569-580). */
void fxref_fft_f32(int L, const float *in, float *out);
void fxref_fft_f64(int L, const double *in, double *out);

/* Spectrum of one (station, FFT): unpack, rotate, FFT, frac-correct; out is
 * [2][nchan] complex.  Returns validity (0 => out zeroed). */
int fxref_station_spectrum_f32(const fxref_config *c, const uint8_t *data,
                               const double poly[3], double fileoffset,
                               int fftindex, float *out);
int fxref_station_spectrum_f64(const fxref_config *c, const uint8_t *data,
                               const double poly[3], double fileoffset,
                               int fftindex, double *out);

/* One subintegration, FxKernel::process (fxkernel.cpp:253-368).
 * data[nant] packed station buffers; delays [nant][3]; offsets [nant].
 * vis: [nrows][4][nchan] complex (interleaved); counts: [nrows].
 * Returns 0 on success. */
int fxref_process_f32(const fxref_config *c, const uint8_t *const *data,
                      const double *delays, const double *offsets, float *vis,
                      int *counts);
int fxref_process_f64(const fxref_config *c, const uint8_t *const *data,
                      const double *delays, const double *offsets, double *vis,
                      int *counts);

/* FP32 process over FFTs [t0, t1) only, accumulating unnormalised sums into
 * vis/counts (no zeroing, no normalisation).  Used to build full-size
 * checks from slices. */
int fxref_accumulate_range_f32(const fxref_config *c,
                               const uint8_t *const *data,
                               const double *delays, const double *offsets,
                               int t0, int t1, float *vis, int *counts);

/* Cross (and optional auto) multiply-accumulate of fxkernel.cpp:327-356 over
 * given spectra [nfft][2*nant][nchan] complex: sequential FP32 sums
 * vis[row][2l+m] += X_j,l * conj(X_k,m), unnormalised.  Used for the
 * integer known-answer test of gcorr/test/unit.cc:42-159. */
void fxref_xmac_f32(const float *spectra, int nant, int nchan, int nfft,
                    int autos, float *vis);

/* ---- VDIF placement (gcorr ingest, ingest/src/ingest.c:267-355) ----------
 * Header fields as examples/mark/dump_vdif.py:36-55 decodes them (VDIF
 * words little-endian): w0 = invalid:1 legacy:1 seconds:30; w1 = epoch:6
 * (bits 24-29) frame:24; w2 = log2(channels):5 (bits 24-28) length/8:24;
 * w3 = complex:1 (bit 31) bits-1:5 (26-30) thread:10 (16-25) station:16.
 * Slot of a frame: (seconds - ref_second) * frames_per_second + frame -
 * ref_frame (ingest.c:313).  Frames are read at a fixed stride of
 * frame_bytes; checks in order: thread (other_thread), format (malformed:
 * length, legacy bit differing from the first frame, channels != 2, bits,
 * data type, frame >= frames_per_second), invalid bit (invalid), epoch and
 * slot range (out_of_window), slot already taken by an earlier frame
 * (duplicate).  Slots with a frame get its payload (8-bit offset-binary
 * samples XOR 0x80 when offset_binary8), the others zeros; frame_ok[slot]
 * tells which. */
typedef struct {
  int frame_bytes; /* 0: from the first frame's length field */
  int thread_id;   /* -1: every thread */
  int frames_per_second;
  int ref_epoch;
  long long ref_second;
  int ref_frame;
  int offset_binary8;
  int header_bytes; /* 16 or 32; 0: from the first frame's legacy bit */
} fxref_vdif_spec;

typedef struct {
  long long frames, placed, missing, duplicate, invalid, other_thread,
      out_of_window, malformed;
} fxref_vdif_stats;

/* Places the frames of `stream` into one station buffer of the config's
 * bytes_per_station.  frame_ok has room for the returned slot count (at
 * most bytes_per_station / 8 + 1).  Returns the slot count and the time
 * samples per frame in *frame_samples, or -1 for an unusable stream. */
long long fxref_vdif_place(const fxref_config *c, const uint8_t *stream,
                           size_t nbytes, const fxref_vdif_spec *spec,
                           uint8_t *station, uint8_t *frame_ok,
                           int *frame_samples, fxref_vdif_stats *stats);

/* bench_fxkernel analogue (fxkernel/src/bench_fxkernel.cpp:93-139,238-285):
 * nthreads threads, each with its own private copy of the input and its own
 * workspace, released together by a start barrier; each runs nloop FP32
 * process() calls.  Returns wall seconds from release to last join, or a
 * negative value on error. */
double fxref_bench_f32(const fxref_config *c, const uint8_t *const *data,
                       const double *delays, const double *offsets,
                       int nthreads, int nloop);

#ifdef __cplusplus
}
#endif

#endif
