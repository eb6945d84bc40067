/*
 * fxref_core.h -- type-generic body of the fxkernel restatement.
 * Included twice by fxref.c: REAL=float/SUF=f32 and REAL=double/SUF=f64.
 * TEST INFRASTRUCTURE ONLY (see fxref.h).
 *
 * Complex values are stored interleaved (re, im) in REAL arrays.
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* Forward, unnormalised, radix-4/radix-2 Stockham FFT.  The twiddle table is
 * computed in double and rounded once to REAL.  Stands in for IPP's
 * ippsFFTFwd_CToC_32fc (fxkernel.cpp:139-143,574). */
typedef struct {
  int L;
  REAL *tw; /* [L] complex: exp(-2 pi i m / L) */
} FN(fftplan);

static int FN(fftplan_init)(FN(fftplan) * p, int L) {
  p->L = L;
  p->tw = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)L);
  if (!p->tw) return 1;
  for (int m = 0; m < L; m++) {
    double ang = -2.0 * FXREF_PI * (double)m / (double)L;
    p->tw[2 * m] = (REAL)cos(ang);
    p->tw[2 * m + 1] = (REAL)sin(ang);
  }
  return 0;
}

static void FN(fftplan_free)(FN(fftplan) * p) { free(p->tw); p->tw = NULL; }

/* out must not alias in; work is [L] complex scratch. */
static void FN(fft_exec)(const FN(fftplan) * p, const REAL *in, REAL *out,
                         REAL *work) {
  const int L = p->L;
  int nstage = 0;
  while ((1 << nstage) < L) nstage++;
  /* Number of radix-4 passes and whether a final radix-2 pass is needed. */
  const int n4 = nstage / 2, n2 = nstage % 2;
  const int npass = n4 + n2;
  /* Arrange ping-pong so that the last pass writes to out. */
  const REAL *src = in;
  REAL *bufs[2];
  if (npass % 2) { bufs[0] = out; bufs[1] = work; }
  else { bufs[0] = work; bufs[1] = out; }
  if (npass == 0) { /* L == 1 */
    out[0] = in[0]; out[1] = in[1];
    return;
  }
  int Ns = 1, pass = 0;
  for (int q = 0; q < n4; q++, pass++) {
    REAL *dst = bufs[pass % 2];
    const int Q = L / 4;
    const int twstep = L / (4 * Ns);
    for (int j = 0; j < Q; j++) {
      const int k = j & (Ns - 1);
      const int base = (j / Ns) * 4 * Ns + k;
      REAL ar[4], ai[4];
      for (int m = 0; m < 4; m++) {
        REAL xr = src[2 * (j + m * Q)], xi = src[2 * (j + m * Q) + 1];
        if (m == 0 || k == 0) { ar[m] = xr; ai[m] = xi; }
        else {
          const int ti = k * m * twstep;
          REAL wr = p->tw[2 * ti], wi = p->tw[2 * ti + 1];
          ar[m] = xr * wr - xi * wi;
          ai[m] = xr * wi + xi * wr;
        }
      }
      REAL t0r = ar[0] + ar[2], t0i = ai[0] + ai[2];
      REAL t1r = ar[0] - ar[2], t1i = ai[0] - ai[2];
      REAL t2r = ar[1] + ar[3], t2i = ai[1] + ai[3];
      REAL t3r = ar[1] - ar[3], t3i = ai[1] - ai[3];
      dst[2 * base] = t0r + t2r;               dst[2 * base + 1] = t0i + t2i;
      dst[2 * (base + Ns)] = t1r + t3i;        dst[2 * (base + Ns) + 1] = t1i - t3r;
      dst[2 * (base + 2 * Ns)] = t0r - t2r;    dst[2 * (base + 2 * Ns) + 1] = t0i - t2i;
      dst[2 * (base + 3 * Ns)] = t1r - t3i;    dst[2 * (base + 3 * Ns) + 1] = t1i + t3r;
    }
    src = dst;
    Ns *= 4;
  }
  if (n2) {
    REAL *dst = bufs[pass % 2];
    const int H = L / 2;
    const int twstep = L / (2 * Ns);
    for (int j = 0; j < H; j++) {
      const int k = j & (Ns - 1);
      const int base = (j / Ns) * 2 * Ns + k;
      REAL ar = src[2 * j], ai = src[2 * j + 1];
      REAL xr = src[2 * (j + H)], xi = src[2 * (j + H) + 1];
      REAL br, bi;
      if (k == 0) { br = xr; bi = xi; }
      else {
        const int ti = k * twstep;
        REAL wr = p->tw[2 * ti], wi = p->tw[2 * ti + 1];
        br = xr * wr - xi * wi;
        bi = xr * wi + xi * wr;
      }
      dst[2 * base] = ar + br;        dst[2 * base + 1] = ai + bi;
      dst[2 * (base + Ns)] = ar - br; dst[2 * (base + Ns) + 1] = ai - bi;
    }
  }
}

void FN(fxref_fft)(int L, const REAL *in, REAL *out) {
  FN(fftplan) p;
  if (FN(fftplan_init)(&p, L)) return;
  REAL *work = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)L);
  FN(fft_exec)(&p, in, out, work);
  free(work);
  FN(fftplan_free)(&p);
}

/* 2-bit unpack (fxkernel.cpp:443-476) / complex 8-bit unpack (layout
 * [Are,Aim,Bre,Bim], doc/gCorr Design.pdf p.2; byte offset 4*offset, the
 * corrected form of fxkernel.cpp:478-487). out: [2][L] complex. */
void FN(fxref_unpack)(const fxref_config *c, const uint8_t *data, int offset,
                      REAL *out) {
  fxref_derived d;
  fxref_derive(c, &d);
  const int L = d.fftchannels;
  REAL *p0 = out, *p1 = out + 2 * (size_t)L;
  if (!c->is_complex) {
    float lut[256][4];
    fxref_lut2bit(lut);
    /* Sample n of the window is packed sample offset+n: byte (offset+n)/2,
     * time slot (offset+n)%2 holding pols at LUT entries 2*slot, 2*slot+1. */
    for (int n = 0; n < L; n++) {
      const long long s = (long long)offset + n;
      const uint8_t byte = data[s >> 1];
      const int slot = (int)(s & 1);
      p0[2 * n] = (REAL)lut[byte][2 * slot];
      p0[2 * n + 1] = 0;
      p1[2 * n] = (REAL)lut[byte][2 * slot + 1];
      p1[2 * n + 1] = 0;
    }
  } else {
    const int8_t *b = (const int8_t *)data + 4 * (size_t)offset;
    for (int n = 0; n < L; n++) {
      p0[2 * n] = (REAL)b[4 * n];
      p0[2 * n + 1] = (REAL)b[4 * n + 1];
      p1[2 * n] = (REAL)b[4 * n + 2];
      p1[2 * n + 1] = (REAL)b[4 * n + 3];
    }
  }
}

/* Fringe rotator, fxkernel.cpp:498-567.  Sub table covers substride
 * consecutive samples; the step table advances by stride (quirk) or by
 * substride (exact) per block of substride samples.  rot[i*S2+k] =
 * sub[k]*step[i], with rot[k] = sub[k] for i = 0 (fxkernel.cpp:553-558). */
void FN(fxref_rotator)(const fxref_config *c, double a, double b, REAL *out) {
  fxref_derived d;
  fxref_derive(c, &d);
  const int L = d.fftchannels, S2 = d.substride, ns = d.nsteps;
  const int stepmul = c->quirk ? d.stride : d.substride;
  const double lo = c->lo;
  const double integerdelay = trunc(b); /* static_cast<int>(b), :504 */
  b -= integerdelay;
  const int fraclo = (lo - trunc(lo)) > 1e-11; /* TINY, :66-71 */
  REAL *sub = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)S2);
  REAL *step = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)ns);
  for (int k = 0; k < S2; k++) {
    const double xoff = (double)k / (double)L;         /* subxoff, :107 */
    double ph = xoff * a;                              /* :508 */
    ph = ph + b;                                       /* :514 */
    ph = ph * lo;                                      /* :519 */
    if (fraclo) ph = ph + (lo - trunc(lo)) * integerdelay; /* :525-530 */
    const double arg = -2.0 * FXREF_PI * (ph - trunc(ph)); /* :535 */
    SINCOS_ARG(arg, &sub[2 * k + 1], &sub[2 * k]);
  }
  for (int i = 0; i < ns; i++) {
    const double xoff = (double)(i * stepmul) / (double)L; /* stepxoff, :112 */
    double ph = xoff * a;                                  /* :511 */
    ph = ph * lo;                                          /* :522 */
    const double arg = -2.0 * FXREF_PI * (ph - trunc(ph)); /* :539 */
    SINCOS_ARG(arg, &step[2 * i + 1], &step[2 * i]);
  }
  for (int k = 0; k < S2; k++) {
    out[2 * k] = sub[2 * k];
    out[2 * k + 1] = sub[2 * k + 1];
  }
  for (int i = 1; i < ns; i++) {
    const REAL wr = step[2 * i], wi = step[2 * i + 1];
    REAL *o = out + 2 * (size_t)i * S2;
    for (int k = 0; k < S2; k++) {
      const REAL sr = sub[2 * k], si = sub[2 * k + 1];
      o[2 * k] = sr * wr - si * wi;
      o[2 * k + 1] = sr * wi + si * wr;
    }
  }
  free(sub);
  free(step);
}

/* Fractional-sample rotator, fxkernel.cpp:159-175 (channel frequency tables)
 * and :595-634.  fr[i*s+k] = fsub[k]*fstep[i], fr[k] = fsub[k]. */
void FN(fxref_fracrot)(const fxref_config *c, double fracdelay, REAL *out) {
  fxref_derived d;
  fxref_derive(c, &d);
  const int s = d.stride, nsf = d.fracsteps, nchan = c->nchan;
  const double bw = c->bw;
  REAL *fsub = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)s);
  REAL *fstep = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)nsf);
  const REAL frac = (REAL)fracdelay; /* ippsMulC_32f takes an Ipp32f value, :600 */
  for (int k = 0; k < s; k++) {
    const REAL freq = (REAL)((2 * FXREF_PI * k * bw) / (double)nchan); /* :173 */
    const REAL arg = freq * frac;                                       /* :600 */
    SINCOS_REAL(arg, &fsub[2 * k + 1], &fsub[2 * k]);
  }
  for (int i = 0; i < nsf; i++) {
    const REAL freq = (REAL)((2 * FXREF_PI * i * s * bw) / (double)nchan); /* :174 */
    const REAL arg = freq * frac;                                           /* :603 */
    SINCOS_REAL(arg, &fstep[2 * i + 1], &fstep[2 * i]);
  }
  for (int k = 0; k < s; k++) {
    out[2 * k] = fsub[2 * k];
    out[2 * k + 1] = fsub[2 * k + 1];
  }
  for (int i = 1; i < nsf; i++) {
    const REAL wr = fstep[2 * i], wi = fstep[2 * i + 1];
    REAL *o = out + 2 * (size_t)i * s;
    for (int k = 0; k < s; k++) {
      const REAL sr = fsub[2 * k], si = fsub[2 * k + 1];
      o[2 * k] = sr * wr - si * wi;
      o[2 * k + 1] = sr * wi + si * wr;
    }
  }
  free(fsub);
  free(fstep);
}

/* Per-call workspace (fxkernel.cpp:73-180 allocations). */
typedef struct {
  fxref_derived d;
  FN(fftplan) plan;
  REAL *unpacked;   /* [2][L] */
  REAL *rot;        /* [L] */
  REAL *chan;       /* [nant][2][L] */
  REAL *fwork;      /* [L] */
  REAL *fr;         /* [nchan] */
  int *valid;       /* [nant] */
} FN(ws);

static int FN(ws_init)(FN(ws) * w, const fxref_config *c) {
  memset(w, 0, sizeof(*w));
  if (fxref_derive(c, &w->d)) return 1;
  const size_t L = (size_t)w->d.fftchannels;
  if (FN(fftplan_init)(&w->plan, (int)L)) return 1;
  w->unpacked = (REAL *)malloc(sizeof(REAL) * 4 * L);
  w->rot = (REAL *)malloc(sizeof(REAL) * 2 * L);
  w->chan = (REAL *)malloc(sizeof(REAL) * 4 * L * (size_t)c->nant);
  w->fwork = (REAL *)malloc(sizeof(REAL) * 2 * L);
  w->fr = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)c->nchan);
  w->valid = (int *)malloc(sizeof(int) * (size_t)c->nant);
  return !(w->unpacked && w->rot && w->chan && w->fwork && w->fr && w->valid);
}

static void FN(ws_free)(FN(ws) * w) {
  FN(fftplan_free)(&w->plan);
  free(w->unpacked); free(w->rot); free(w->chan); free(w->fwork); free(w->fr);
  free(w->valid);
}

/* Station-based processing of antenna j for FFT i, fxkernel.cpp:288-324.
 * Leaves nchan channelised bins of both pols in w->chan[j]. */
static int FN(station)(FN(ws) * w, const fxref_config *c, const uint8_t *data,
                       const double poly[3], double fileoffset, int i,
                       REAL *chan_j, int st) {
  const size_t L = (size_t)w->d.fftchannels;
  int offset;
  double frac, a, b;
  if (!fxref_delay_split(c, poly, fileoffset, i, &offset, &frac, &a, &b))
    return 0;
  if (!frames_present(c, st, offset, (int)L)) return 0; /* missing VDIF frames */
  FN(fxref_unpack)(c, data, offset, w->unpacked);      /* :310 */
  FN(fxref_rotator)(c, a, b, w->rot);                  /* :313 */
  for (int p = 0; p < 2; p++) {                        /* :561-566 */
    REAL *u = w->unpacked + 2 * L * p;
    for (size_t n = 0; n < L; n++) {
      const REAL rr = w->rot[2 * n], ri = w->rot[2 * n + 1];
      const REAL ur = u[2 * n], ui = u[2 * n + 1];
      u[2 * n] = rr * ur - ri * ui;
      u[2 * n + 1] = rr * ui + ri * ur;
    }
    FN(fft_exec)(&w->plan, u, chan_j + 2 * L * p, w->fwork); /* :316 */
  }
  FN(fxref_fracrot)(c, frac, w->fr);                   /* :321 */
  for (int p = 0; p < 2; p++) {
    REAL *x = chan_j + 2 * L * p;
    for (int q = 0; q < c->nchan; q++) {
      const REAL fr = w->fr[2 * q], fi = w->fr[2 * q + 1];
      const REAL xr = x[2 * q], xi = x[2 * q + 1];
      x[2 * q] = fr * xr - fi * xi;
      x[2 * q + 1] = fr * xi + fi * xr;
    }
  }
  if (c->polgains && st >= 0) { /* gpubench.cu:156-166 */
    const float *g0 = c->polgains + 2 * (size_t)c->nchan * (2 * (size_t)st);
    const float *g1 = g0 + 2 * (size_t)c->nchan;
    REAL *x0 = chan_j, *x1 = chan_j + 2 * L;
    for (int q = 0; q < c->nchan; q++) {
      const REAL ar = (REAL)g0[2 * q], ai = (REAL)g0[2 * q + 1];
      const REAL br = (REAL)g1[2 * q], bi = (REAL)g1[2 * q + 1];
      const REAL ur = ar * x0[2 * q] - ai * x0[2 * q + 1], ui = ar * x0[2 * q + 1] + ai * x0[2 * q];
      const REAL vr = br * x1[2 * q] - bi * x1[2 * q + 1], vi = br * x1[2 * q + 1] + bi * x1[2 * q];
      x0[2 * q] = ur - vr;
      x0[2 * q + 1] = ui - vi;
      x1[2 * q] = ur + vr;
      x1[2 * q + 1] = ui + vi;
    }
  }
  return 1;
}

int FN(fxref_station_spectrum)(const fxref_config *c, const uint8_t *data,
                               const double poly[3], double fileoffset,
                               int fftindex, REAL *out) {
  FN(ws) w;
  if (FN(ws_init)(&w, c)) { FN(ws_free)(&w); return -1; }
  const size_t L = (size_t)w.d.fftchannels;
  int ok = FN(station)(&w, c, data, poly, fileoffset, fftindex, w.chan, -1);
  for (int p = 0; p < 2; p++)
    for (int q = 0; q < c->nchan; q++) {
      out[2 * ((size_t)p * c->nchan + q)] = ok ? w.chan[2 * (L * p + q)] : 0;
      out[2 * ((size_t)p * c->nchan + q) + 1] = ok ? w.chan[2 * (L * p + q) + 1] : 0;
    }
  FN(ws_free)(&w);
  return ok;
}

/* Cross (and optional auto) multiply-accumulate for one FFT,
 * fxkernel.cpp:327-356: vis[b][2l+m] += X_j,l * conj(X_k,m). */
static void FN(xmac)(const FN(ws) * w, const fxref_config *c, REAL *vis,
                     int *counts) {
  const size_t L = (size_t)w->d.fftchannels;
  const int nchan = c->nchan, nant = c->nant;
  int b = 0;
  for (int j = 0; j < nant - 1; j++) {
    if (!w->valid[j]) { b += nant - (j + 1); continue; }
    for (int k = j + 1; k < nant; k++) {
      if (!w->valid[k]) { b++; continue; }
      for (int l = 0; l < 2; l++)
        for (int m = 0; m < 2; m++) {
          const REAL *x = w->chan + 2 * L * (2 * (size_t)j + l);
          const REAL *y = w->chan + 2 * L * (2 * (size_t)k + m);
          REAL *v = vis + 2 * (size_t)nchan * (4 * (size_t)b + 2 * l + m);
          for (int q = 0; q < nchan; q++) {
            const REAL xr = x[2 * q], xi = x[2 * q + 1];
            const REAL yr = y[2 * q], yi = -y[2 * q + 1]; /* conjChannels, :587 */
            v[2 * q] += xr * yr - xi * yi;
            v[2 * q + 1] += xr * yi + xi * yr;
          }
        }
      counts[b]++;
      b++;
    }
  }
  if (c->autos) {
    for (int a = 0; a < nant; a++) {
      if (!w->valid[a]) continue;
      const int row = w->d.nbaselines + a;
      for (int l = 0; l < 2; l++)
        for (int m = 0; m < 2; m++) {
          const REAL *x = w->chan + 2 * L * (2 * (size_t)a + l);
          const REAL *y = w->chan + 2 * L * (2 * (size_t)a + m);
          REAL *v = vis + 2 * (size_t)nchan * (4 * (size_t)row + 2 * l + m);
          for (int q = 0; q < nchan; q++) {
            const REAL xr = x[2 * q], xi = x[2 * q + 1];
            const REAL yr = y[2 * q], yi = -y[2 * q + 1];
            v[2 * q] += xr * yr - xi * yi;
            v[2 * q + 1] += xr * yi + xi * yr;
          }
        }
      counts[row]++;
    }
  }
}

static int FN(process_range)(FN(ws) * w, const fxref_config *c,
                             const uint8_t *const *data, const double *delays,
                             const double *offsets, int t0, int t1, REAL *vis,
                             int *counts) {
  const size_t L = (size_t)w->d.fftchannels;
  for (int i = t0; i < t1; i++) {                      /* :276 */
    for (int j = 0; j < c->nant; j++)                  /* :279 */
      w->valid[j] = FN(station)(w, c, data[j], delays + 3 * (size_t)j,
                                offsets[j], i, w->chan + 4 * L * (size_t)j, j);
    FN(xmac)(w, c, vis, counts);
  }
  return 0;
}

int FN(fxref_process)(const fxref_config *c, const uint8_t *const *data,
                      const double *delays, const double *offsets, REAL *vis,
                      int *counts) {
  FN(ws) w;
  if (FN(ws_init)(&w, c)) { FN(ws_free)(&w); return -1; }
  const int nrows = w.d.nrows, nchan = c->nchan;
  memset(vis, 0, sizeof(REAL) * 2 * 4 * (size_t)nchan * nrows); /* :264-270 */
  memset(counts, 0, sizeof(int) * (size_t)nrows);                /* :273 */
  FN(process_range)(&w, c, data, delays, offsets, 0, c->numffts, vis, counts);
  for (int r = 0; r < nrows; r++) {                               /* :358-367 */
    if (counts[r] == 0) continue;
    const REAL norm = (REAL)counts[r];
    REAL *v = vis + 2 * 4 * (size_t)nchan * r;
    for (size_t q = 0; q < 2 * 4 * (size_t)nchan; q++) v[q] = v[q] / norm;
  }
  FN(ws_free)(&w);
  return 0;
}

#undef FN
#undef CAT
#undef CAT_
