/*
 * gxcorr.h -- C-ABI of the MI355X FX-correlation hot path.
 *
 * Drop-in for the per-subintegration processing interface of fxkernel
 * (fxkernel/src/fxkernel.h:18-183) and for the GPU call sequence of gcorr
 * (gcorr/src/testgpukernel.cu:462-534).  All arguments are plain pointers,
 * sizes and PODs; no C++ or framework types cross this boundary.
 *
 * Data layouts (identical to the reference):
 *   input       one packed buffer per station, (numffts+1)*fftchannels time
 *               samples (fxkernel/src/common.cpp:16).  Real 2-bit: 4 samples
 *               per byte, LSB first, [A_t, B_t, A_t+1, B_t+1]
 *               (fxkernel/src/fxkernel.cpp:443-476).  Complex 8-bit
 *               (extension): [Are, Aim, Bre, Bim] int8 per time sample.
 *   delays      per station a quadratic d(t)=d[0]*t^2+d[1]*t+d[2] seconds,
 *               t in FFT intervals, plus a file start offset in seconds
 *               (fxkernel/src/fxkernel.h:42-47, common.cpp:77-89).
 *   visibilities [row][product][channel] complex float, rows 0..nbl-1 the
 *               cross baselines (0,1),(0,2)..(0,N-1),(1,2).. and product
 *               p = 2*l+m = X_j,l * conj(X_k,m) (fxkernel.cpp:327-356);
 *               when autos are enabled rows nbl..nbl+nant-1 hold station
 *               autocorrelations with the same product order.
 *
 * Errors are returned as gxc_status; gxc_last_error() gives text.
 * A handle owns one HIP stream and its device buffers; handles are
 * independent and may be used from different threads; one handle must not
 * be used from two threads at once (fxkernel objects are one per thread,
 * fxkernel/src/bench_fxkernel.cpp:217-263).
 */
#ifndef GXCORR_H
#define GXCORR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GXC_ABI_VERSION 1

/* Layout-compatible with fxkernel's cf32 (Ipp32fc, fxkernel/src/vectordefs.h:16). */
typedef struct {
  float re;
  float im;
} gxc_cf32;

typedef enum {
  GXC_OK = 0,
  GXC_ERR_INVALID_ARG = 1, /* bad config or argument (fxkernel exits, fxkernel.cpp:47-63) */
  GXC_ERR_UNSUPPORTED = 2, /* valid but not supported by this build */
  GXC_ERR_NO_DEVICE = 3,   /* no usable gfx950 GPU */
  GXC_ERR_HIP = 4,         /* a HIP runtime call failed */
  GXC_ERR_OUT_OF_MEMORY = 5,
  GXC_ERR_STATE = 6,       /* input or delays not set before process */
  GXC_ERR_IO = 7           /* file could not be read or written */
} gxc_status;

typedef enum {
  GXC_PREC_FP32 = 0, /* drop-in: FP32 spectra, FP32 X-engine */
  GXC_PREC_FP16 = 1, /* FP16 spectra (per-channel RMS scaled), MFMA X-engine */
  GXC_PREC_BF16 = 2, /* BF16 spectra, MFMA X-engine */
  GXC_PREC_FP8 = 3   /* FP8 e4m3 spectra with per-channel/per-station scales */
} gxc_precision;

/* Correlator configuration: the FxKernel constructor arguments
 * (fxkernel/src/fxkernel.h:29) plus the .conf keywords NPOL
 * (fxkernel/src/common.cpp:94-95) and this build's extensions. */
typedef struct {
  int nant;        /* NANT */
  int nchan;       /* NCHAN, power of two */
  int numffts;     /* NUMFFTS per subintegration */
  int nbit;        /* NBIT: 2 (real) or 8 (complex) */
  int npol;        /* NPOL: must be 2 (fxkernel.cpp:77-81 assumes dual pol) */
  int is_complex;  /* COMPLEX */
  double lo;       /* LO, Hz */
  double bw;       /* BANDWIDTH, Hz */
  int precision;   /* gxc_precision; GXC_PREC_FP32 is the drop-in */
  int autos;       /* 1: append autocorrelation rows */
  int exact_phase; /* 0: fxkernel's stride fringe-phase law (default);
                      1: exact per-sample phase */
  int chunk_ffts;  /* FFTs per device pass; 0 = choose automatically: the
                      whole subintegration while its spectra fit in 32 GiB
                      of device memory, else as many passes as keep them
                      there (with overlap: eight passes, at most 4 GiB per
                      spectra buffer) */
  int overlap;     /* 1: the F-stage of chunk k+1 runs on a second stream
                      while the X-stage of chunk k runs (double-buffered
                      spectra); 0: one stream, one chunk after another */
  int graph;       /* 1: capture the device work of one subintegration as a
                      hipGraph on first use and replay it afterwards */
  /* Launch shapes (0 = default; gxc_autotune measures and sets them): */
  int fblock;      /* F-stage workgroup size, i.e. FFT blocks per workgroup
                      = fblock / (fftchannels/16); or GXC_FBLOCK_MFMA */
  int xtile;       /* X-stage channel tile (channels per workgroup) */
  int xsplits;     /* X-stage FFT splits (workgroups per channel tile) */
  int reserved[1];
} gxc_config;

/* gxc_config.fblock value selecting the F-stage whose FFT passes run on the
 * matrix cores (f16 operands, FP32 accumulation): real input; fftchannels
 * 2048..8192 for the reduced precisions, 2048..4096 for FP32 (operands as
 * f16 pairs plus remainders, within the FP32 bound). */
#define GXC_FBLOCK_MFMA (-1)
/* In a batch's tuning result: the subbands of 2048 points and more have
 * launches of their own on the register F-stage (gxc_batch_config.fstage_mfma
 * 2; the name is from when its workgroups took one polarisation each, as
 * they still do above 8192 points).  No counterpart in the reference (its
 * subbands are separate runs). */
#define GXC_FBLOCK_ONEPOL (-2)

typedef struct gxc_handle gxc_handle;

/* Per-stage device time of the most recent gxc_process, milliseconds,
 * measured with HIP events on the handle's stream (profiling enabled). */
typedef struct {
  double h2d_ms;       /* host-to-device input copy (0 for device input) */
  double fstage_ms;    /* unpack+rotate+FFT+frac (all chunks) */
  double xstage_ms;    /* cross-multiply-accumulate (all chunks); with overlap
                          on, F and X times are measured on their own streams
                          and overlap in wall time */
  double finalise_ms;  /* count + normalise */
  double total_ms;     /* first to last event */
  int fstage_launches; /* number of F-stage kernel launches */
  int xstage_launches;
} gxc_timings;

/* Fill cfg with defaults: nbit 2, npol 2, real, FP32, no autos, stride law,
 * automatic chunking, overlap off, graph replay on. */
void gxc_config_init(gxc_config *cfg);

/* FxKernel::FxKernel (fxkernel/src/fxkernel.cpp:30-180) */
gxc_status gxc_create(const gxc_config *cfg, gxc_handle **out);

/* FxKernel::~FxKernel (fxkernel/src/fxkernel.cpp:182-240) */
void gxc_destroy(gxc_handle *h);

/* Text of the last error on this handle (h may be NULL: last error of the
 * calling thread, e.g. from a failed gxc_create). */
const char *gxc_last_error(const gxc_handle *h);
const char *gxc_status_string(gxc_status s);

/* FxKernel::setInputData (fxkernel/src/fxkernel.cpp:242-245): host buffers,
 * one per station, each gxc_station_bytes() long.  Pointers are stored, not
 * copied; they must stay valid until gxc_process returns. */
gxc_status gxc_set_input(gxc_handle *h, const uint8_t *const *per_station);

/* As gxc_set_input but the buffers already live in device memory (HBM) of
 * the current device: no copy is made (gcorr keeps packed data on the device,
 * gcorr/src/testgpukernel.cu:468-470). */
gxc_status gxc_set_input_device(gxc_handle *h, const void *const *per_station);

/* Polarisation conversion (e.g. linear to circular), as gpubench's
 * linear2circular_kernel (examples/chris/bench/gpubench.cu:141-166) applies it
 * to each station's channelised spectra: y0 = g0 x0 - g1 x1,
 * y1 = g0 x0 + g1 x1 with complex gains g0, g1 per station and channel.
 * gains is [nant][2][nchan] (g0 then g1 of each station), copied at this
 * call; NULL turns the conversion off.  The result equals converting every
 * spectrum; it is computed on the visibilities as V' = J_j V J_k^H with
 * J = [[g0, -g1], [g0, g1]], which is the same sum.  The row counts are
 * unchanged. */
gxc_status gxc_set_pol_conversion(gxc_handle *h, const gxc_cf32 *gains);

/* Requantisers: float voltages in device memory -> the packed station
 * layout (gpubench packs filterbank output to 2 bits on the GPU,
 * examples/chris/bench/gpubench.cu:64-130).  The 2-bit codes are the ones
 * fxkernel's LUT decodes (fxkernel.cpp:15-28, 0..3 -> -HI, -1, +1, +HI):
 *   x < -t -> 0, -t <= x < 0 -> 1, 0 <= x < t -> 2, x >= t -> 3 (NaN -> 1),
 * packed LSB first [A_t, B_t, A_t+1, B_t+1] (gpubench's own pack2bit uses
 * the reverse code order); nsamples even, output nsamples/2 bytes.
 * 8-bit: complex samples (re, im interleaved floats) to
 * q = clamp(rint(x*scale), -127, 127) (NaN -> 0), bytes [Are, Aim, Bre, Bim],
 * output 4*nsamples bytes.  pol0/pol1 16-byte aligned; out 8- (2-bit) or
 * 16-byte (8-bit) aligned.  Queued on `stream` (a hipStream_t, NULL for the
 * default stream); the call does not wait. */
gxc_status gxc_requantise_2bit(const float *d_pol0, const float *d_pol1, size_t nsamples, float threshold,
                               uint8_t *d_out, void *stream);
gxc_status gxc_requantise_8bit(const float *d_pol0, const float *d_pol1, size_t nsamples, float scale,
                               uint8_t *d_out, void *stream);

/* ---- VDIF ingest --------------------------------------------------------
 * gcorr's ingest (ingest/src/ingest.c:267-355) reads VDIF frames, strips
 * the headers and places each payload by its time stamp:
 *   slot = (seconds - ref_second) * frames_per_second + frame - ref_frame
 * (ingest.c:313).  gxc_set_input_vdif does that on the GPU for one
 * station of the next subintegration: frames already in device memory,
 * fixed frame size, two channels (the polarisations) per frame, nbit and
 * real/complex as configured.  Header fields follow the VDIF layout that
 * examples/mark/dump_vdif.py:36-55 decodes.  The first frame of the stream
 * for each slot is kept; a frame is skipped when its thread differs from
 * thread_id (>= 0), its format differs from the configuration, its invalid
 * bit is set, or it falls outside the subintegration.  Slots that receive
 * no frame are zero-filled and flagged missing, and every FFT whose input
 * samples touch a missing frame is invalid for that station (it is left
 * out of the sums and the row counts, as fxkernel does with delays out of
 * range).  8-bit samples are offset binary in VDIF and are converted to
 * two's complement (offset_binary8 = 1). */
typedef struct {
  int frame_bytes;        /* header + payload bytes of every frame; 0: from the
                             first frame's length field */
  int thread_id;          /* VDIF thread to keep; -1 keeps every thread */
  int frames_per_second;  /* frames per second of the kept thread */
  int ref_epoch;          /* reference epoch (6-bit VDIF field) of sample 0 */
  long long ref_second;   /* seconds from the epoch of sample 0 */
  int ref_frame;          /* frame number within that second of sample 0 */
  int offset_binary8;     /* 8-bit samples: 1 = offset binary (VDIF), 0 = two's complement */
  int header_bytes;       /* 32, or 16 for legacy headers; 0: from the first
                             frame's legacy bit (then a malformed first frame
                             makes every frame of another geometry count as
                             malformed; with frame_bytes 0 the call checks the
                             first header and rejects a mismatch) */
  int reserved;
} gxc_vdif_spec;

typedef struct {
  long long frames;        /* whole frames in the stream */
  long long placed;        /* slots filled */
  long long missing;       /* slots with no frame (flagged) */
  long long duplicate;     /* frames for an already filled slot */
  long long invalid;       /* frames with the invalid bit set */
  long long other_thread;  /* frames of another thread */
  long long out_of_window; /* frames outside the subintegration or another epoch */
  long long malformed;     /* frames whose length, channel count, bits or
                              data type differ from the configuration */
} gxc_vdif_stats;

void gxc_vdif_spec_init(gxc_vdif_spec *spec); /* thread 0, offset binary, rest 0 */

/* Station `station`'s input for the next gxc_process from `nbytes` of VDIF
 * frames at device address d_frames (8-byte aligned).  The ingest is queued
 * on the handle's stream and copies the payloads into a buffer the handle
 * owns: d_frames may be reused once gxc_process, gxc_synchronize or a call
 * with stats non-NULL has returned.  With spec->frame_bytes set the call
 * does not wait for the device; with 0 it reads the first header first.
 * Every station must be ingested this way before gxc_process;
 * gxc_set_input / gxc_set_input_device switch the handle back to
 * unflagged input.  With stats non-NULL the call waits for the ingest and
 * reports its counts. */
gxc_status gxc_set_input_vdif(gxc_handle *h, int station, const void *d_frames, size_t nbytes,
                              const gxc_vdif_spec *spec, gxc_vdif_stats *stats);

/* Frames per station buffer for a spec's frame size (slots of one
 * subintegration), or 0 if frame_bytes does not hold whole samples. */
long long gxc_vdif_slots(const gxc_config *cfg, int frame_bytes, int legacy_header);

/* The missing-frame flags of station `station` from the last ingest:
 * out[f] = 1 when slot f holds a frame.  *nslots (may be NULL) gets the
 * slot count; out may be NULL to ask for the count only. */
gxc_status gxc_get_frame_flags(gxc_handle *h, int station, uint8_t *out, size_t n, long long *nslots);

/* FxKernel::setDelays (fxkernel/src/fxkernel.cpp:247-251): poly3[a] points
 * to 3 doubles, fileoffsets[a] seconds.  Values are copied at this call. */
gxc_status gxc_set_delays(gxc_handle *h, const double *const *poly3,
                          const double *fileoffsets);

/* FxKernel::process (fxkernel/src/fxkernel.cpp:253-368): one subintegration.
 * Synchronous: returns when the visibilities are complete on the device. */
gxc_status gxc_process(gxc_handle *h);

/* As gxc_process but only enqueues the work on the handle's stream. */
gxc_status gxc_process_async(gxc_handle *h);

/* Wait for work enqueued on the handle's stream. */
gxc_status gxc_synchronize(gxc_handle *h);

/* FxKernel::accumulate (fxkernel/src/fxkernel.cpp:414-423):
 * odata[row][product][channel] += visibilities, rows 0..gxc_num_rows-1. */
gxc_status gxc_accumulate(gxc_handle *h, gxc_cf32 *const *const *odata);

/* Copy the visibilities [rows][4][nchan] to host memory (nelem >= total). */
gxc_status gxc_get_visibilities(gxc_handle *h, gxc_cf32 *out, size_t nelem);

/* Number of FFTs accumulated per row (fxkernel.cpp:352, baselineCount). */
gxc_status gxc_get_row_counts(gxc_handle *h, int *counts, size_t n);

/* FxKernel::saveVisibilities (fxkernel/src/fxkernel.cpp:370-412): same text
 * format, one line per channel. */
gxc_status gxc_save_visibilities(gxc_handle *h, const char *path);

/* ---- Visibility output and the fringe check ------------------------------
 * Binary visibilities of the last gxc_process, little-endian:
 *   char  magic[8] = "GXCVIS01"
 *   int32 version (1), nant, nchan, nrows, nprod (4), autos, numffts, precision
 *   f64   bandwidth_hz, lo_hz
 *   int32 stations[nrows][2]   (j, k) of each row; autos have j == k
 *   int32 counts[nrows]        FFTs summed per row (fxkernel's baselineCount)
 *   cf32  vis[nrows][4][nchan] the [baseline][product][channel] layout
 * The text of gxc_save_visibilities (fxkernel.cpp:370-412) in binary, with
 * the metadata needed to read it back. */
gxc_status gxc_save_visibilities_binary(gxc_handle *h, const char *path);

/* DiFX-style output: one record per (row, product) in DiFX's binary
 * visibility format: a 74-byte header (sync word 0xFF00FF00, version 1,
 * baseline 256*(j+1)+(k+1), MJD, seconds, config, source and frequency
 * indices, the two polarisation names, pulsar bin 0, data weight =
 * count/numffts, u, v, w in metres) followed by nchan cf32 values. */
typedef struct {
  int mjd;              /* day of the subintegration */
  double seconds;       /* seconds into that day */
  int config_index, source_index, freq_index;
  char pols[2];         /* names of polarisations 0 and 1, e.g. 'X','Y' or 'R','L' */
  const double *uvw;    /* [rows][3] in metres, or NULL for zeros */
} gxc_difx_meta;
gxc_status gxc_save_difx(gxc_handle *h, const char *path, const gxc_difx_meta *meta);

/* Lag S/N of the last gxc_process, the fringe check of util/lagsnr.py:33-36:
 * for series i = row*4 + product, snr[i] = max|ifft(V)| / std|ifft(V)|
 * (population standard deviation over the nchan lags) and peak[i] = the lag
 * index of the maximum (the first one on ties).  n >= rows*4.  Runs on the
 * GPU (radix-2 Stockham inverse FFT of every series, then one reduction per
 * series). */
gxc_status gxc_lag_snr(gxc_handle *h, float *snr, int *peak, size_t n);
/* The same for nseries host series of nchan (a power of two) visibilities,
 * on the current device. */
gxc_status gxc_lag_snr_series(const gxc_cf32 *vis, int nseries, int nchan, float *snr, int *peak);

/* Device pointer of the visibilities [rows][4][nchan] cf32 (stream-ordered). */
const void *gxc_visibilities_device(const gxc_handle *h);

/* Bytes per station buffer, fxkernel/src/common.cpp:16 (numffts+1 FFTs). */
size_t gxc_station_bytes(const gxc_config *cfg);

/* Visibility rows: nant*(nant-1)/2 cross baselines (+ nant autos). */
int gxc_num_rows(const gxc_config *cfg);

/* Stage entry points used by tests and benchmarks; each mirrors a stage of
 * gcorr's kernel sequence (gcorr/src/gxkernel.h:69-90).
 *
 * gxc_fstage: F-stage for FFTs [t0, t0+nfft) of the handle's input and
 * delays.  spectra: host [nfft][nant*2][nchan] cf32 (may be NULL);
 * valid: host [nant][nfft] (may be NULL).  For GXC_PREC_FP32 only. */
gxc_status gxc_fstage(gxc_handle *h, int t0, int nfft, gxc_cf32 *spectra,
                      uint8_t *valid);

/* gxc_xengine: X-stage alone over host spectra [nfft][ninputs][nchan]
 * (ninputs = 2*nant), all FFTs valid, result divided by nfft as in
 * gcorr's CrossCorr+finaliseAccum (gcorr/src/gxkernel.cu:353-424) and
 * written to out [rows][4][nchan] (rows per `autos`).  The integer
 * known-answer test pattern of gcorr/test/unit.cc:73-159 runs through it. */
gxc_status gxc_xengine(const gxc_cf32 *spectra, int nant, int nchan, int nfft,
                       int autos, int precision, gxc_cf32 *out);

/* gxc_xengine_noalloc: gxc_xengine over device memory, allocating nothing.
 * It has no counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  d_spectra: device [nfft][2*nant][nchan] cf32; d_vis: device
 * [rows][4][nchan] cf32; queued on `stream` (a hipStream_t, NULL for the
 * default stream), the call does not wait.  Its working memory (row map
 * and counts, the spectra packed to 16/8 bits for the reduced precisions,
 * the split partials: 8 (2 nant)^2 nchan bytes per split, at least one)
 * comes from one fixed device array of GXC_NOALLOC_SCRATCH_BYTES that the
 * library holds on every GPU it runs on until the process ends; shapes
 * that need more return GXC_ERR_INVALID_ARG.  All calls on one GPU share
 * that array, so calls that overlap on one GPU race. */
#define GXC_NOALLOC_SCRATCH_BYTES (64ull << 20)
gxc_status gxc_xengine_noalloc(const gxc_cf32 *d_spectra, int nant, int nchan, int nfft,
                               int autos, int precision, gxc_cf32 *d_vis, void *stream);

/* Autotuner: with input and delays set, times the F-stage at each available
 * workgroup size (FFT blocks per workgroup), the X-stage at each channel
 * tile and split count, and whole subintegrations with and without graph
 * replay, on this handle's own data; keeps the fastest.  The choice is
 * returned (may be NULL) and applies to later gxc_process calls; it can be
 * given back through gxc_config.fblock/xtile/xsplits/graph of a new handle
 * with the same shape to skip the search.  No visibilities are kept from
 * the trial runs: call gxc_process afterwards. */
typedef struct {
  int fblock, xtile, xsplits, graph;  /* the choice */
  double fstage_ms, xstage_ms;        /* their times (all chunks of one subint) */
  double process_ms;                  /* one subintegration, as chosen */
  double default_ms;                  /* one subintegration at the defaults */
  int candidates;                     /* configurations timed */
} gxc_tuning;
gxc_status gxc_autotune(gxc_handle *h, gxc_tuning *out);

/* ---- Streaming subintegrations from host memory ----------------------------
 * fxkernel processes one subintegration per process() and the caller sums
 * them with accumulate() (fxkernel.cpp:414-423); gcorr's driver overlaps host
 * copies with compute on two streams (testgpukernel.cu:448-486).  These calls
 * pipeline that on one handle: gxc_submit queues the copy of a
 * subintegration's input (host, one buffer per station of gxc_station_bytes)
 * and delays into one of two device slots on a copy stream, then its
 * correlation on the handle's stream, and adds its visibilities to a device
 * sum; the copy of subintegration k+1 runs while subintegration k is
 * correlated.  Host buffers from gxc_host_alloc (pinned) copy asynchronously;
 * any other host memory works but copies synchronously.  A submission's
 * station buffers may be refilled once the next gxc_submit (or
 * gxc_integrate_end) has returned, so two sets of host buffers suffice;
 * delays are copied at the call.  The sum is
 * Σ_k vis_k in submission order, each vis_k normalised as gxc_process does. */
void *gxc_host_alloc(size_t bytes);  /* pinned host memory, NULL on failure */
void gxc_host_free(void *p);
gxc_status gxc_integrate_begin(gxc_handle *h);  /* zero the sum */
gxc_status gxc_submit(gxc_handle *h, const uint8_t *const *per_station,
                      const double *const *poly3, const double *fileoffsets);
/* Wait for every submitted subintegration; the sum [rows][4][nchan] goes to
 * out (may be NULL), the number summed to *nsubints (may be NULL). */
gxc_status gxc_integrate_end(gxc_handle *h, gxc_cf32 *out, size_t nelem, int *nsubints);

/* Per-stage timing (see gxc_timings).  Enabled by gxc_set_profiling. */
gxc_status gxc_set_profiling(gxc_handle *h, int enable);
gxc_status gxc_get_timings(const gxc_handle *h, gxc_timings *t);

/* ---- Several subbands per launch ------------------------------------------
 * fxkernel correlates one band per FxKernel object (fxkernel/src/fxkernel.h:29);
 * an observation with several subbands runs one object per subband.  A batch
 * holds every subband of a subintegration: the same stations, each subband
 * with its own channel count, bandwidth, LO, FFT count, delays and input.
 * Each stage runs for all subbands in one persistent launch (F-stage,
 * X-stage, row counts, finalise), and with graph on the whole subintegration
 * is captured as one hipGraph.  Per subband the visibilities are those of a
 * gxc_handle with the same settings (same layout, same normalisation).
 * Real 2-bit input, FFT lengths 2*nchan of 256..8192. */
typedef struct {
  int nchan;    /* NCHAN of this subband, power of two, 128..4096 */
  int numffts;  /* NUMFFTS of this subband's subintegration */
  double lo;    /* LO, Hz */
  double bw;    /* BANDWIDTH, Hz */
} gxc_subband;

typedef struct {
  int nant;                     /* stations, shared by all subbands */
  int nbit;                     /* 2 */
  int npol;                     /* 2 */
  int is_complex;               /* 0 */
  int precision;                /* gxc_precision */
  int autos;                    /* 1: append autocorrelation rows */
  int exact_phase;              /* as gxc_config.exact_phase */
  int nsub;                     /* number of subbands */
  const gxc_subband *subbands;  /* [nsub], copied by gxc_batch_create */
  int chunks;                   /* device passes per subintegration (each
                                   covers 1/chunks of every subband's FFTs);
                                   0 = automatic: as few as keep the spectra
                                   within 32 GiB (eight with overlap) */
  int overlap;                  /* as gxc_config.overlap */
  int graph;                    /* as gxc_config.graph */
  int fstage_mfma;              /* 1: subbands of 2048 points and more get a
                                   launch each, on the matrix-core F-stage
                                   (GXC_FBLOCK_MFMA) where it covers them,
                                   else on the register F-stage; 2: a launch
                                   each on the register F-stage; the others
                                   stay in the multi-subband launch.
                                   gxc_batch_autotune chooses. */
  int reserved[3];
} gxc_batch_config;

typedef struct gxc_batch gxc_batch;

/* Defaults: nbit 2, npol 2, real, FP32, no autos, automatic chunks, overlap
 * off, graph on, no subbands. */
void gxc_batch_config_init(gxc_batch_config *cfg);
gxc_status gxc_batch_create(const gxc_batch_config *cfg, gxc_batch **out);
void gxc_batch_destroy(gxc_batch *b);
const char *gxc_batch_last_error(const gxc_batch *b);
/* Per subband `sub`, as gxc_set_input / gxc_set_input_device / gxc_set_delays
 * (station buffers of gxc_station_bytes() of that subband's configuration). */
gxc_status gxc_batch_set_input(gxc_batch *b, int sub, const uint8_t *const *per_station);
gxc_status gxc_batch_set_input_device(gxc_batch *b, int sub, const void *const *per_station);
gxc_status gxc_batch_set_delays(gxc_batch *b, int sub, const double *const *poly3,
                                const double *fileoffsets);
/* Station `station` of subband `sub` from VDIF frames in device memory, as
 * gxc_set_input_vdif (gcorr's ingest, ingest/src/ingest.c:267-355): the
 * subband's frames are those of spec->thread_id; missing frames make the
 * FFTs over them invalid.  gxc_batch_set_input / _device leave VDIF mode. */
gxc_status gxc_batch_set_input_vdif(gxc_batch *b, int sub, int station, const void *d_frames, size_t nbytes,
                                    const gxc_vdif_spec *spec, gxc_vdif_stats *stats);
/* gxc_get_frame_flags for station `station` of subband `sub`. */
gxc_status gxc_batch_get_frame_flags(gxc_batch *b, int sub, int station, uint8_t *out, size_t n,
                                     long long *nslots);
/* Polarisation conversion of subband `sub`, as gxc_set_pol_conversion
 * (examples/chris/bench/gpubench.cu:141-166): gains [nant][2][nchan of that
 * subband], copied at this call; NULL turns it off. */
gxc_status gxc_batch_set_pol_conversion(gxc_batch *b, int sub, const gxc_cf32 *gains);
/* One subintegration of every subband. */
gxc_status gxc_batch_process(gxc_batch *b);
gxc_status gxc_batch_process_async(gxc_batch *b);
gxc_status gxc_batch_synchronize(gxc_batch *b);
/* Subband `sub`'s visibilities [rows][4][nchan] and row counts. */
gxc_status gxc_batch_get_visibilities(gxc_batch *b, int sub, gxc_cf32 *out, size_t nelem);
gxc_status gxc_batch_get_row_counts(gxc_batch *b, int sub, int *counts, size_t n);
/* Autotuner of a batch (input and delays of every subband set): times whole
 * subintegrations for each F-stage workgroup size (256/512 threads) and X
 * unit count (1024..8192, which sets every subband's FFT splits), with and
 * without launches of their own for the subbands of 2048 points and more
 * (gxc_batch_config.fstage_mfma), then with and without graph replay, and
 * keeps the fastest.  In the result, fblock is the multi-subband workgroup
 * size, or GXC_FBLOCK_MFMA / GXC_FBLOCK_ONEPOL when the own launches (on
 * the matrix cores where they cover the length / on the register kernel)
 * were chosen (the choice is kept in the batch); xsplits holds the chosen
 * X unit count and fstage_ms/xstage_ms are 0. */
gxc_status gxc_batch_autotune(gxc_batch *b, gxc_tuning *out);

/* Timing over all subbands (launch counts are per-stage launches, one per
 * chunk for F and X). */
gxc_status gxc_batch_set_profiling(gxc_batch *b, int enable);
gxc_status gxc_batch_get_timings(const gxc_batch *b, gxc_timings *t);

/* .conf reader, fxkernel/src/common.cpp:54-114 (same keywords, same
 * defaults, unknown keywords reported and skipped).  The strings and arrays
 * are owned by the struct and released by gxc_conffile_free. */
typedef struct {
  gxc_config cfg;
  int nant_read;     /* antenna lines actually read */
  char **names;      /* [nant] */
  char **files;      /* [nant] */
  double *delays;    /* [nant][3] */
  double *offsets;   /* [nant] */
} gxc_conffile;

gxc_status gxc_parse_config(const char *path, gxc_conffile *out);
void gxc_conffile_free(gxc_conffile *cf);

#ifdef __cplusplus
}
#endif

#endif /* GXCORR_H */
