/*
 * fxref.c -- IPP-free CPU restatement of fxkernel (fxkernel/src/fxkernel.cpp).
 * TEST INFRASTRUCTURE ONLY: see fxref.h for who may use it.
 *
 * Compile with -ffp-contract=off: the FP64 delay and phase model must be
 * evaluated with one rounding per operation, in fxkernel's operation order,
 * so that the integer sample shift and validity decisions are bit-exact
 * against any other implementation that does the same.
 */
#define _GNU_SOURCE
#include "fxref.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define FXREF_PI 3.14159265358979323846 /* IPP_PI; TWO_PI is 2*IPP_PI (vectordefs.h:8) */

static int ilog2_exact(int x) {
  if (x <= 0 || (x & (x - 1))) return -1;
  int k = 0;
  while ((1 << k) < x) k++;
  return k;
}

int fxref_derive(const fxref_config *c, fxref_derived *d) {
  memset(d, 0, sizeof(*d));
  if (!c || c->nant < 1 || c->numffts < 1 || !(c->bw > 0)) return 1;
  const int lg = ilog2_exact(c->nchan);
  if (lg < 1) return 2; /* FFT length must be a power of two */
  if (c->is_complex ? (c->nbit != 8) : (c->nbit != 2)) return 3;
  /* fxkernel.cpp:30-43: fftchannels = 2*nchan real / nchan complex;
   * sampletime = 1/(2 bw), doubled for complex. */
  d->fftchannels = c->is_complex ? c->nchan : 2 * c->nchan;
  d->sampletime = 1.0 / (2.0 * c->bw);
  if (c->is_complex) d->sampletime *= 2;
  /* fxkernel.cpp:58-64: stridesize = sqrt(nchan) for square nchan; the
   * generalisation s = 2^floor(log2(nchan)/2) agrees on every square power
   * of two and covers the others. */
  d->stride = 1 << (lg / 2);
  d->substride = c->is_complex ? d->stride : 2 * d->stride;
  d->nsteps = d->fftchannels / d->substride;
  d->fracsteps = c->nchan / d->stride;
  d->maxoffset = (c->numffts - 1) * d->fftchannels; /* :272 */
  d->nbaselines = c->nant * (c->nant - 1) / 2;       /* :147 */
  d->nrows = d->nbaselines + (c->autos ? c->nant : 0);
  /* common.cpp:16 sizes a station buffer as numffts+1 FFTs.  Complex
   * 8-bit takes 4 bytes per time sample (2 pols x re,im). */
  const size_t samples = (size_t)(c->numffts + 1) * (size_t)d->fftchannels;
  d->bytes_per_station = c->is_complex ? 4 * samples : samples / 2;
  return 0;
}

void fxref_lut2bit(float lut[256][4]) {
  /* fxkernel.cpp:15-28 */
  static const float HiMag = 3.3359f;
  const float lv[4] = {-HiMag, -1.0f, 1.0f, HiMag};
  for (int b = 0; b < 256; b++)
    for (int i = 0; i < 4; i++) lut[b][i] = lv[(b >> (2 * i)) & 0x03];
}

void fxref_station_delay(const double poly[3], int fftindex, double *mean,
                         double *a, double *b) {
  /* fxkernel.cpp:425-440, same operation order. */
  const double t0 = (double)fftindex;
  const double th = fftindex + 0.5;
  const double t1 = fftindex + 1.0;
  const double d0 = poly[0] * t0 * t0 + poly[1] * t0 + poly[2];
  const double d1 = poly[0] * th * th + poly[1] * th + poly[2];
  const double d2 = poly[0] * t1 * t1 + poly[1] * t1 + poly[2];
  *a = d2 - d0;
  *b = d0 + (d1 - (*a * 0.5 + d0)) / 3.0;
  *mean = *a * 0.5 + *b;
}

int fxref_delay_split(const fxref_config *c, const double poly[3],
                      double fileoffset, int fftindex, int *offset,
                      double *fracdelay, double *a, double *b) {
  fxref_derived d;
  if (fxref_derive(c, &d)) return 0;
  double mean;
  fxref_station_delay(poly, fftindex, &mean, a, b);
  /* fxkernel.cpp:289-309 */
  const double ndf = (mean - fileoffset) / d.sampletime;
  if (!(fabs(ndf) < 1.0e9)) { *offset = 0; *fracdelay = 0; return 0; }
  const int nd = (int)(ndf + 0.5); /* truncation toward zero */
  double frac = -(ndf - nd) * d.sampletime;
  long long off = (long long)fftindex * d.fftchannels - nd;
  if (off == -1) { ++off; frac += d.sampletime; }
  if (off == (long long)d.maxoffset + 1) { --off; frac -= d.sampletime; }
  *offset = (int)off;
  *fracdelay = frac;
  return !(off < 0 || off > d.maxoffset);
}

/* Missing-frame rule (an extension, see fxref.h): station st's FFT input
 * samples [offset, offset + L) must lie in frames flagged present. */
static int frames_present(const fxref_config *c, int st, long long offset, int L) {
  if (!c->frame_ok || st < 0) return 1;
  const uint8_t *ok = c->frame_ok[st];
  const long long spf = c->frame_samples[st];
  for (long long f = offset / spf; f <= (offset + L - 1) / spf; f++)
    if (!ok[f]) return 0;
  return 1;
}

static uint32_t le32(const uint8_t *p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

long long fxref_vdif_place(const fxref_config *c, const uint8_t *stream,
                           size_t nbytes, const fxref_vdif_spec *spec,
                           uint8_t *station, uint8_t *frame_ok,
                           int *frame_samples, fxref_vdif_stats *stats) {
  fxref_derived d;
  if (fxref_derive(c, &d) || nbytes < 16 || spec->frames_per_second <= 0) return -1;
  memset(stats, 0, sizeof(*stats));
  const uint32_t legacy0 = spec->header_bytes ? (spec->header_bytes == 16 ? 1u : 0u) : (le32(stream) >> 30) & 1;
  const long long hdr = legacy0 ? 16 : 32;
  const long long fb = spec->frame_bytes ? spec->frame_bytes
                                         : 8LL * (le32(stream + 8) & 0xffffff);
  const long long payload = fb - hdr;
  const long long bytes_per_time = (long long)c->nbit * 2 * (c->is_complex ? 2 : 1); /* bits */
  if (payload <= 0 || payload % 8 || (payload * 8) % bytes_per_time) return -1;
  *frame_samples = (int)(payload * 8 / bytes_per_time);
  const long long sb = (long long)d.bytes_per_station;
  const long long nslots = (sb + payload - 1) / payload;
  const long long nframes = (long long)(nbytes / (size_t)fb); /* whole frames only, ingest.c:304-305 */
  stats->frames = nframes;
  long long *src = (long long *)malloc(sizeof(long long) * (size_t)nslots);
  if (!src) return -1;
  for (long long s = 0; s < nslots; s++) src[s] = -1;
  for (long long i = 0; i < nframes; i++) {
    const uint8_t *f = stream + i * fb;
    const uint32_t w0 = le32(f), w1 = le32(f + 4), w2 = le32(f + 8), w3 = le32(f + 12);
    const int thread = (int)((w3 >> 16) & 0x3ff);
    if (spec->thread_id >= 0 && thread != spec->thread_id) { stats->other_thread++; continue; }
    const long long frame = w1 & 0xffffff;
    if (8LL * (w2 & 0xffffff) != fb || ((w0 >> 30) & 1) != legacy0 || ((w2 >> 24) & 0x1f) != 1 ||
        (int)((w3 >> 26) & 0x1f) + 1 != c->nbit || (int)(w3 >> 31) != (c->is_complex ? 1 : 0) ||
        frame >= spec->frames_per_second) {
      stats->malformed++;
      continue;
    }
    if (w0 >> 31) { stats->invalid++; continue; }
    const long long seconds = w0 & 0x3fffffff;
    const long long rel = (seconds - spec->ref_second) * spec->frames_per_second + frame - spec->ref_frame;
    if ((int)((w1 >> 24) & 0x3f) != spec->ref_epoch || rel < 0 || rel >= nslots) {
      stats->out_of_window++;
      continue;
    }
    if (src[rel] >= 0) { stats->duplicate++; continue; }
    src[rel] = i;
  }
  const int flip = c->nbit == 8 && spec->offset_binary8;
  for (long long s = 0; s < nslots; s++) {
    const long long n = s * payload + payload <= sb ? payload : sb - s * payload;
    uint8_t *dst = station + s * payload;
    if (src[s] >= 0) {
      const uint8_t *p = stream + src[s] * fb + hdr;
      for (long long b = 0; b < n; b++) dst[b] = flip ? (uint8_t)(p[b] ^ 0x80) : p[b];
      frame_ok[s] = 1;
      stats->placed++;
    } else {
      memset(dst, 0, (size_t)n);
      frame_ok[s] = 0;
      stats->missing++;
    }
  }
  free(src);
  return nslots;
}

#define REAL float
#define SUF f32
#define SINCOS_ARG(x, s, co) sincosf((float)(x), (s), (co))
#define SINCOS_REAL(x, s, co) sincosf((x), (s), (co))
#include "fxref_core.h"
#undef REAL
#undef SUF
#undef SINCOS_ARG
#undef SINCOS_REAL

#define REAL double
#define SUF f64
#define SINCOS_ARG(x, s, co) sincos((x), (s), (co))
#define SINCOS_REAL(x, s, co) sincos((x), (s), (co))
#include "fxref_core.h"
#undef REAL
#undef SUF
#undef SINCOS_ARG
#undef SINCOS_REAL

int fxref_accumulate_range_f32(const fxref_config *c,
                               const uint8_t *const *data,
                               const double *delays, const double *offsets,
                               int t0, int t1, float *vis, int *counts) {
  ws_f32 w;
  if (ws_init_f32(&w, c)) { ws_free_f32(&w); return -1; }
  if (t0 < 0) t0 = 0;
  if (t1 > c->numffts) t1 = c->numffts;
  process_range_f32(&w, c, data, delays, offsets, t0, t1, vis, counts);
  ws_free_f32(&w);
  return 0;
}

void fxref_xmac_f32(const float *spectra, int nant, int nchan, int nfft,
                    int autos, float *vis) {
  const int nbl = nant * (nant - 1) / 2;
  const int nrows = nbl + (autos ? nant : 0);
  memset(vis, 0, sizeof(float) * 8 * (size_t)nchan * nrows);
  for (int t = 0; t < nfft; t++) {
    const float *X = spectra + 2 * (size_t)t * 2 * nant * nchan;
    int row = 0;
    for (int j = 0; j < nant; j++)
      for (int k = j + 1; k < nant; k++, row++)
        for (int l = 0; l < 2; l++)
          for (int m = 0; m < 2; m++) {
            const float *x = X + 2 * (size_t)(2 * j + l) * nchan;
            const float *y = X + 2 * (size_t)(2 * k + m) * nchan;
            float *v = vis + 2 * (size_t)nchan * (4 * (size_t)row + 2 * l + m);
            for (int q = 0; q < nchan; q++) {
              v[2 * q] += x[2 * q] * y[2 * q] + x[2 * q + 1] * y[2 * q + 1];
              v[2 * q + 1] += x[2 * q + 1] * y[2 * q] - x[2 * q] * y[2 * q + 1];
            }
          }
    if (autos)
      for (int a = 0; a < nant; a++, row++)
        for (int l = 0; l < 2; l++)
          for (int m = 0; m < 2; m++) {
            const float *x = X + 2 * (size_t)(2 * a + l) * nchan;
            const float *y = X + 2 * (size_t)(2 * a + m) * nchan;
            float *v = vis + 2 * (size_t)nchan * (4 * (size_t)row + 2 * l + m);
            for (int q = 0; q < nchan; q++) {
              v[2 * q] += x[2 * q] * y[2 * q] + x[2 * q + 1] * y[2 * q + 1];
              v[2 * q + 1] += x[2 * q + 1] * y[2 * q] - x[2 * q] * y[2 * q + 1];
            }
          }
  }
}

/* ---- CPU baseline harness (bench_fxkernel analogue) ---- */

typedef struct {
  const fxref_config *c;
  const uint8_t *const *src;
  const double *delays, *offsets;
  int nloop;
  pthread_barrier_t *start;
  int status;
} bench_arg;

static void *bench_thread(void *p) {
  bench_arg *ba = (bench_arg *)p;
  const fxref_config *c = ba->c;
  fxref_derived d;
  fxref_derive(c, &d);
  /* Private copy of the input per thread (bench_fxkernel.cpp:222-231). */
  uint8_t **mine = (uint8_t **)calloc((size_t)c->nant, sizeof(uint8_t *));
  float *vis = (float *)malloc(sizeof(float) * 8 * (size_t)c->nchan * d.nrows);
  int *counts = (int *)malloc(sizeof(int) * (size_t)d.nrows);
  int ok = mine && vis && counts;
  for (int a = 0; ok && a < c->nant; a++) {
    mine[a] = (uint8_t *)malloc(d.bytes_per_station);
    if (!mine[a]) { ok = 0; break; }
    memcpy(mine[a], ba->src[a], d.bytes_per_station);
  }
  pthread_barrier_wait(ba->start); /* start barrier (:239-244) */
  for (int l = 0; ok && l < ba->nloop; l++)
    if (fxref_process_f32(c, (const uint8_t *const *)mine, ba->delays,
                          ba->offsets, vis, counts))
      ok = 0;
  for (int a = 0; mine && a < c->nant; a++) free(mine[a]);
  free(mine); free(vis); free(counts);
  ba->status = ok ? 0 : 1;
  return NULL;
}

double fxref_bench_f32(const fxref_config *c, const uint8_t *const *data,
                       const double *delays, const double *offsets,
                       int nthreads, int nloop) {
  fxref_derived d;
  if (fxref_derive(c, &d) || nthreads < 1 || nloop < 1) return -1.0;
  pthread_barrier_t start;
  pthread_barrier_init(&start, NULL, (unsigned)nthreads + 1);
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  bench_arg *args = (bench_arg *)calloc((size_t)nthreads, sizeof(bench_arg));
  for (int t = 0; t < nthreads; t++) {
    args[t].c = c; args[t].src = data; args[t].delays = delays;
    args[t].offsets = offsets; args[t].nloop = nloop; args[t].start = &start;
    pthread_create(&th[t], NULL, bench_thread, &args[t]);
  }
  /* Threads copy their input before the barrier; time starts at release. */
  pthread_barrier_wait(&start);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int bad = 0;
  for (int t = 0; t < nthreads; t++) {
    pthread_join(th[t], NULL);
    bad |= args[t].status;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  pthread_barrier_destroy(&start);
  free(th); free(args);
  if (bad) return -2.0;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
