// C-ABI of the FX-correlation hot path (include/gxcorr.h): FxKernel's
// processing interface (fxkernel/src/fxkernel.h:18-183) over HIP kernels.
#include "../../include/gxcorr.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "gxc_internal.hpp"
#include "gxc_mixed.hpp"

using gxc::Geometry;

namespace {

thread_local std::string g_thread_error;

struct Derived {
  Geometry g;
  int L = 0, M = 0;
  int maxoffset = 0, nbl = 0, nrows = 0;
  double sampletime = 0;
  size_t station_bytes = 0;
};

gxc_status derive(const gxc_config *c, Derived *d, std::string *why) {
  if (!c) {
    *why = "null config";
    return GXC_ERR_INVALID_ARG;
  }
  if (c->nant < 1) {
    *why = "nant must be >= 1";
    return GXC_ERR_INVALID_ARG;
  }
  if (c->numffts < 1) {
    *why = "numffts must be >= 1";
    return GXC_ERR_INVALID_ARG;
  }
  if (c->nchan < 2 || (c->nchan & (c->nchan - 1))) {
    *why = "nchan must be a power of two";
    return GXC_ERR_INVALID_ARG;
  }
  if (!(c->bw > 0)) {
    *why = "bandwidth must be positive";
    return GXC_ERR_INVALID_ARG;
  }
  if (c->npol != 2) {
    *why = "only dual-polarisation data is supported (npol = 2)";
    return GXC_ERR_UNSUPPORTED;
  }
  if (c->is_complex ? c->nbit != 8 : c->nbit != 2) {
    // fxkernel.cpp:46-55,489-496: 2-bit real is the only supported input;
    // complex 8-bit is this build's extension.
    *why = c->is_complex ? "complex data must be 8-bit" : "real data must be 2-bit";
    return GXC_ERR_UNSUPPORTED;
  }
  if (c->precision < GXC_PREC_FP32 || c->precision > GXC_PREC_FP8) {
    *why = "unknown precision";
    return GXC_ERR_INVALID_ARG;
  }
  d->g = Geometry::make(c->nchan, c->is_complex != 0);
  d->L = d->g.L;
  d->M = 2 * c->nant;
  if (!gxc::fstage_supported(d->L, c->is_complex != 0)) {
    *why = "FFT length " + std::to_string(d->L) + " not supported (64..8192)";
    return GXC_ERR_UNSUPPORTED;
  }
  const long long maxoff = (long long)(c->numffts - 1) * d->L;
  if ((long long)c->numffts * d->L + d->L > 0x7fffffffLL) {
    *why = "subintegration too long for 32-bit sample offsets";
    return GXC_ERR_UNSUPPORTED;
  }
  d->maxoffset = (int)maxoff;
  d->nbl = c->nant * (c->nant - 1) / 2;
  d->nrows = d->nbl + (c->autos ? c->nant : 0);
  // fxkernel.cpp:31,37-38
  d->sampletime = 1.0 / (2.0 * c->bw);
  if (c->is_complex) d->sampletime *= 2;
  const size_t samples = (size_t)(c->numffts + 1) * (size_t)d->L;  // common.cpp:152
  d->station_bytes = c->is_complex ? 4 * samples : samples / 2;
  return GXC_OK;
}

}  // namespace

struct gxc_handle {
  gxc_config cfg{};
  Derived d;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // input
  bool input_set = false, input_device = false;
  std::vector<const uint8_t *> host_in;
  std::vector<const void *> dev_in;
  uint8_t *d_input_owned = nullptr;
  const uint8_t **d_station_ptrs = nullptr;

  // delays
  bool delays_set = false;
  std::vector<double> h_delays;  // [nant][4]
  double *d_delays = nullptr;

  // constant tables
  float2 *d_twiddle = nullptr;
  float *d_subfreq = nullptr, *d_stepfreq = nullptr;

  // work buffers
  int chunk = 0, sb = 1, nb = 0, ntiles = 0, splits = 1;
  float2 *d_spectra = nullptr;
  uint8_t *d_valid = nullptr;
  float2 *d_partial = nullptr;
  int *d_counts = nullptr;
  int2 *d_rowmap = nullptr;
  float2 *d_vis = nullptr;
  gxc::MixedState *mixed = nullptr;

  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> events;
  gxc_timings timings{};
};

namespace {

gxc_status set_err(gxc_handle *h, gxc_status s, const std::string &msg) {
  if (h) h->err = msg;
  g_thread_error = msg;
  return s;
}

#define GXC_HIP(h, call)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_err((h), e_ == hipErrorOutOfMemory ? GXC_ERR_OUT_OF_MEMORY : GXC_ERR_HIP,   \
                     std::string(#call) + ": " + hipGetErrorString(e_));                     \
  } while (0)

gxc_status check_device(gxc_handle *h, int *dev) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) return set_err(h, GXC_ERR_NO_DEVICE, "no HIP device available");
  GXC_HIP(h, hipGetDevice(dev));
  hipDeviceProp_t prop;
  GXC_HIP(h, hipGetDeviceProperties(&prop, *dev));
  if (!strstr(prop.gcnArchName, "gfx950"))
    return set_err(h, GXC_ERR_NO_DEVICE,
                   std::string("library is built for gfx950, device is ") + prop.gcnArchName);
  return GXC_OK;
}

template <typename T>
gxc_status dmalloc(gxc_handle *h, T **p, size_t nelem) {
  *p = nullptr;
  if (nelem == 0) nelem = 1;
  hipError_t e = hipMalloc((void **)p, nelem * sizeof(T));
  if (e != hipSuccess)
    return set_err(h, GXC_ERR_OUT_OF_MEMORY,
                   "hipMalloc of " + std::to_string(nelem * sizeof(T)) + " bytes failed: " +
                       hipGetErrorString(e));
  return GXC_OK;
}

#define GXC_TRY(expr)            \
  do {                           \
    gxc_status s_ = (expr);      \
    if (s_ != GXC_OK) return s_; \
  } while (0)

// Rows: cross baselines in fxkernel order (fxkernel.cpp:328-355), then autos.
std::vector<int2> make_rowmap(int nant, bool autos) {
  std::vector<int2> rm;
  for (int j = 0; j < nant - 1; j++)
    for (int k = j + 1; k < nant; k++) rm.push_back(make_int2(j, k));
  if (autos)
    for (int a = 0; a < nant; a++) rm.push_back(make_int2(a, a));
  return rm;
}

int choose_splits(int nchan, int ntiles, int nt) {
  const long long waves = (long long)((nchan + 63) / 64) * ntiles;
  long long s = (4096 + waves - 1) / waves;
  const long long maxs = nt / 16 > 0 ? nt / 16 : 1;
  if (s > maxs) s = maxs;
  if (s > 512) s = 512;
  if (s < 1) s = 1;
  return (int)s;
}

gxc_status upload_tables(gxc_handle *h) {
  const Derived &d = h->d;
  const int L = d.L, s = d.g.s, nsf = d.g.fracsteps, nchan = h->cfg.nchan;
  const double bw = h->cfg.bw;
  std::vector<float2> tw(L);
  for (int m = 0; m < L; m++) {
    const double ang = -2.0 * gxc::kPi * (double)m / (double)L;
    tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  // fxkernel.cpp:171-175 channel-frequency tables, identical expressions.
  std::vector<float> subf(s), stepf(nsf);
  for (int i = 0; i < s; i++) subf[i] = (float)((2 * gxc::kPi * i * bw) / (double)nchan);
  for (int i = 0; i < nsf; i++) stepf[i] = (float)((2 * gxc::kPi * i * s * bw) / (double)nchan);
  GXC_TRY(dmalloc(h, &h->d_twiddle, L));
  GXC_TRY(dmalloc(h, &h->d_subfreq, s));
  GXC_TRY(dmalloc(h, &h->d_stepfreq, nsf));
  GXC_HIP(h, hipMemcpy(h->d_twiddle, tw.data(), L * sizeof(float2), hipMemcpyHostToDevice));
  GXC_HIP(h, hipMemcpy(h->d_subfreq, subf.data(), s * sizeof(float), hipMemcpyHostToDevice));
  GXC_HIP(h, hipMemcpy(h->d_stepfreq, stepf.data(), nsf * sizeof(float), hipMemcpyHostToDevice));
  return GXC_OK;
}

gxc_status ensure_spectra(gxc_handle *h, int nt) {
  const size_t per_fft = (size_t)h->d.M * h->cfg.nchan;
  if (h->d_spectra && nt <= h->chunk) return GXC_OK;
  if (h->d_spectra) hipFree(h->d_spectra);
  h->d_spectra = nullptr;
  GXC_TRY(dmalloc(h, &h->d_spectra, per_fft * (size_t)nt));
  h->chunk = nt;
  return GXC_OK;
}

void record(gxc_handle *h, size_t idx) {
  if (!h->profiling) return;
  while (h->events.size() <= idx) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    h->events.push_back(e);
  }
  hipEventRecord(h->events[idx], h->stream);
}

gxc_status run_fstage(gxc_handle *h, int t0, int nt) {
  gxc::FParams p{};
  p.station = h->d_station_ptrs;
  p.delays = h->d_delays;
  p.subchanfreq = h->d_subfreq;
  p.stepchanfreq = h->d_stepfreq;
  p.twiddle = h->d_twiddle;
  p.spectra = h->d_spectra;
  p.valid = h->d_valid;
  p.t0 = t0;
  p.nt = nt;
  p.nant = h->cfg.nant;
  p.numffts = h->cfg.numffts;
  p.maxoffset = h->d.maxoffset;
  p.stepmul = h->cfg.exact_phase ? h->d.g.S2 : h->d.g.s;
  const double lo = h->cfg.lo;
  p.lo = lo;
  p.lofrac = lo - std::trunc(lo);
  p.fraclo = (lo - std::trunc(lo)) > 0.00000000001 ? 1 : 0;  // TINY, fxkernel.cpp:66-71
  p.sampletime = h->d.sampletime;
  GXC_HIP(h, gxc::launch_fstage(p, h->d.L, h->cfg.is_complex != 0, h->stream));
  return GXC_OK;
}

gxc_status run_xstage(gxc_handle *h, int nt, int accumulate) {
  gxc::XParams x{};
  x.spectra = h->d_spectra;
  x.partial = h->d_partial;
  x.nt = nt;
  x.nant = h->cfg.nant;
  x.nchan = h->cfg.nchan;
  x.nb = h->nb;
  x.ntiles = h->ntiles;
  x.tile0 = 0;
  x.splits = h->splits;
  x.accumulate = accumulate;
  GXC_HIP(h, gxc::launch_xstage(x, h->sb, h->stream));
  return GXC_OK;
}

gxc_status run_finalise(gxc_handle *h) {
  gxc::CountParams cp{};
  cp.valid = h->d_valid;
  cp.rowmap = h->d_rowmap;
  cp.counts = h->d_counts;
  cp.numffts = h->cfg.numffts;
  cp.nrows = h->d.nrows;
  GXC_HIP(h, gxc::launch_counts(cp, h->stream));
  gxc::FinParams f{};
  f.partial = h->d_partial;
  f.rowmap = h->d_rowmap;
  f.counts = h->d_counts;
  f.vis = h->d_vis;
  f.nrows = h->d.nrows;
  f.nchan = h->cfg.nchan;
  f.nb = h->nb;
  f.ntiles = h->ntiles;
  f.splits = h->splits;
  f.sb = h->sb;
  GXC_HIP(h, gxc::launch_finalise(f, h->stream));
  return GXC_OK;
}

gxc_status enqueue_process(gxc_handle *h) {
  if (!h->input_set) return set_err(h, GXC_ERR_STATE, "input data not set");
  if (!h->delays_set) return set_err(h, GXC_ERR_STATE, "delays not set");
  const gxc_config &c = h->cfg;
  size_t ev = 0;
  record(h, ev++);  // 0: start
  if (!h->input_device) {
    if (!h->d_input_owned) {
      GXC_TRY(dmalloc(h, &h->d_input_owned, h->d.station_bytes * (size_t)c.nant));
      std::vector<const uint8_t *> ptrs(c.nant);
      for (int a = 0; a < c.nant; a++) ptrs[a] = h->d_input_owned + (size_t)a * h->d.station_bytes;
      GXC_HIP(h, hipMemcpy(h->d_station_ptrs, ptrs.data(), c.nant * sizeof(void *),
                           hipMemcpyHostToDevice));
    }
    for (int a = 0; a < c.nant; a++)
      GXC_HIP(h, hipMemcpyAsync(h->d_input_owned + (size_t)a * h->d.station_bytes, h->host_in[a],
                                h->d.station_bytes, hipMemcpyHostToDevice, h->stream));
  }
  record(h, ev++);  // 1: input resident
  if (c.precision != GXC_PREC_FP32)
    return set_err(h, GXC_ERR_UNSUPPORTED, "mixed-precision process not available in this build");
  int nchunks = 0;
  for (int t0 = 0; t0 < c.numffts; t0 += h->chunk, nchunks++) {
    const int nt = std::min(h->chunk, c.numffts - t0);
    record(h, ev++);
    GXC_TRY(run_fstage(h, t0, nt));
    record(h, ev++);
    GXC_TRY(run_xstage(h, nt, t0 > 0 ? 1 : 0));
    record(h, ev++);
  }
  GXC_TRY(run_finalise(h));
  record(h, ev++);
  h->timings.fstage_launches = nchunks;
  h->timings.xstage_launches = nchunks;
  return GXC_OK;
}

float elapsed(gxc_handle *h, size_t a, size_t b) {
  float ms = 0.f;
  if (b < h->events.size() && hipEventElapsedTime(&ms, h->events[a], h->events[b]) == hipSuccess)
    return ms;
  return 0.f;
}

void collect_timings(gxc_handle *h) {
  if (!h->profiling || h->events.size() < 3) return;
  gxc_timings &t = h->timings;
  const int n = t.fstage_launches;
  t.h2d_ms = elapsed(h, 0, 1);
  t.fstage_ms = t.xstage_ms = 0;
  for (int k = 0; k < n; k++) {
    t.fstage_ms += elapsed(h, 2 + 3 * k, 3 + 3 * k);
    t.xstage_ms += elapsed(h, 3 + 3 * k, 4 + 3 * k);
  }
  t.finalise_ms = elapsed(h, 1 + 3 * n, 2 + 3 * n);
  t.total_ms = elapsed(h, 0, 2 + 3 * n);
}

}  // namespace

extern "C" {

void gxc_config_init(gxc_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->nbit = 2;  // common.cpp:200-202 defaults
  cfg->npol = 2;
  cfg->is_complex = 0;
  cfg->precision = GXC_PREC_FP32;
}

const char *gxc_status_string(gxc_status s) {
  switch (s) {
    case GXC_OK: return "ok";
    case GXC_ERR_INVALID_ARG: return "invalid argument";
    case GXC_ERR_UNSUPPORTED: return "unsupported";
    case GXC_ERR_NO_DEVICE: return "no device";
    case GXC_ERR_HIP: return "HIP error";
    case GXC_ERR_OUT_OF_MEMORY: return "out of memory";
    case GXC_ERR_STATE: return "bad state";
    case GXC_ERR_IO: return "I/O error";
  }
  return "unknown";
}

const char *gxc_last_error(const gxc_handle *h) {
  return h ? h->err.c_str() : g_thread_error.c_str();
}

size_t gxc_station_bytes(const gxc_config *cfg) {
  Derived d;
  std::string why;
  if (derive(cfg, &d, &why) != GXC_OK) return 0;
  return d.station_bytes;
}

int gxc_num_rows(const gxc_config *cfg) {
  Derived d;
  std::string why;
  if (derive(cfg, &d, &why) != GXC_OK) return 0;
  return d.nrows;
}

gxc_status gxc_create(const gxc_config *cfg, gxc_handle **out) {
  if (!out) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null output handle pointer");
  *out = nullptr;
  Derived d;
  std::string why;
  gxc_status s = derive(cfg, &d, &why);
  if (s != GXC_OK) return set_err(nullptr, s, why);
  int dev = 0;
  s = check_device(nullptr, &dev);
  if (s != GXC_OK) return s;
  gxc_handle *h = new gxc_handle();
  h->cfg = *cfg;
  h->d = d;
  h->device = dev;
  auto fail = [&](gxc_status st) {
    g_thread_error = h->err;
    gxc_destroy(h);
    return st;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(h, GXC_ERR_HIP, "hipStreamCreate failed"));
  if ((s = upload_tables(h)) != GXC_OK) return fail(s);
  const int nant = cfg->nant, nchan = cfg->nchan;
  h->h_delays.assign((size_t)nant * 4, 0.0);
  if ((s = dmalloc(h, &h->d_delays, (size_t)nant * 4)) != GXC_OK) return fail(s);
  if ((s = dmalloc(h, &h->d_station_ptrs, (size_t)nant)) != GXC_OK) return fail(s);
  if ((s = dmalloc(h, &h->d_valid, (size_t)nant * cfg->numffts)) != GXC_OK) return fail(s);
  // Rows and their station pairs.
  std::vector<int2> rm = make_rowmap(nant, cfg->autos != 0);
  if ((s = dmalloc(h, &h->d_rowmap, rm.size())) != GXC_OK) return fail(s);
  if (!rm.empty() &&
      hipMemcpy(h->d_rowmap, rm.data(), rm.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_err(h, GXC_ERR_HIP, "rowmap upload failed"));
  if ((s = dmalloc(h, &h->d_counts, (size_t)d.nrows)) != GXC_OK) return fail(s);
  if ((s = dmalloc(h, &h->d_vis, (size_t)d.nrows * 4 * nchan)) != GXC_OK) return fail(s);
  if (hipMemset(h->d_vis, 0, (size_t)d.nrows * 4 * nchan * sizeof(float2)) != hipSuccess)
    return fail(set_err(h, GXC_ERR_HIP, "hipMemset failed"));
  // Chunk of FFTs per device pass: bounded spectra buffer (default 8 GiB).
  const size_t per_fft = (size_t)d.M * nchan * sizeof(float2);
  long long chunk = cfg->chunk_ffts > 0 ? cfg->chunk_ffts : (long long)((8ull << 30) / per_fft);
  if (chunk < 1) chunk = 1;
  if (chunk > cfg->numffts) chunk = cfg->numffts;
  if (cfg->precision == GXC_PREC_FP32) {
    if ((s = ensure_spectra(h, (int)chunk)) != GXC_OK) return fail(s);
    h->sb = gxc::xstage_block_stations(nant);
    h->nb = (nant + h->sb - 1) / h->sb;
    h->ntiles = h->nb * (h->nb + 1) / 2;
    h->splits = choose_splits(nchan, h->ntiles, (int)chunk);
    const size_t ni = 2 * (size_t)h->sb;
    if ((s = dmalloc(h, &h->d_partial, (size_t)h->splits * h->ntiles * ni * ni * nchan)) != GXC_OK)
      return fail(s);
  } else {
    h->chunk = (int)chunk;
    std::string merr;
    s = gxc::mixed_create(&h->mixed, *cfg, d.M, d.nrows, d.nbl, (int)chunk, h->d_valid, h->d_rowmap,
                          h->d_counts, h->d_vis, &merr);
    if (s != GXC_OK) return fail(set_err(h, s, merr));
  }
  *out = h;
  return GXC_OK;
}

void gxc_destroy(gxc_handle *h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (hipEvent_t e : h->events) hipEventDestroy(e);
  gxc::mixed_destroy(h->mixed);
  if (h->cfg.precision == GXC_PREC_FP32 && h->d_spectra) hipFree(h->d_spectra);
  hipFree(h->d_input_owned);
  hipFree(h->d_station_ptrs);
  hipFree(h->d_delays);
  hipFree(h->d_twiddle);
  hipFree(h->d_subfreq);
  hipFree(h->d_stepfreq);
  hipFree(h->d_valid);
  hipFree(h->d_partial);
  hipFree(h->d_counts);
  hipFree(h->d_rowmap);
  hipFree(h->d_vis);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

gxc_status gxc_set_input(gxc_handle *h, const uint8_t *const *per_station) {
  if (!h || !per_station) return set_err(h, GXC_ERR_INVALID_ARG, "null input");
  h->host_in.assign(per_station, per_station + h->cfg.nant);
  for (const uint8_t *p : h->host_in)
    if (!p) return set_err(h, GXC_ERR_INVALID_ARG, "null station buffer");
  if (h->input_device && h->d_input_owned) {
    // switching back from device input: restore pointers into the owned copy
    std::vector<const uint8_t *> ptrs(h->cfg.nant);
    for (int a = 0; a < h->cfg.nant; a++) ptrs[a] = h->d_input_owned + (size_t)a * h->d.station_bytes;
    GXC_HIP(h, hipMemcpy(h->d_station_ptrs, ptrs.data(), h->cfg.nant * sizeof(void *),
                         hipMemcpyHostToDevice));
  }
  h->input_device = false;
  h->input_set = true;
  return GXC_OK;
}

gxc_status gxc_set_input_device(gxc_handle *h, const void *const *per_station) {
  if (!h || !per_station) return set_err(h, GXC_ERR_INVALID_ARG, "null input");
  h->dev_in.assign(per_station, per_station + h->cfg.nant);
  for (const void *p : h->dev_in)
    if (!p) return set_err(h, GXC_ERR_INVALID_ARG, "null station buffer");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  GXC_HIP(h, hipMemcpy(h->d_station_ptrs, h->dev_in.data(), h->cfg.nant * sizeof(void *),
                       hipMemcpyHostToDevice));
  h->input_device = true;
  h->input_set = true;
  return GXC_OK;
}

gxc_status gxc_set_delays(gxc_handle *h, const double *const *poly3, const double *fileoffsets) {
  if (!h || !poly3 || !fileoffsets) return set_err(h, GXC_ERR_INVALID_ARG, "null delays");
  for (int a = 0; a < h->cfg.nant; a++) {
    if (!poly3[a]) return set_err(h, GXC_ERR_INVALID_ARG, "null delay polynomial");
    for (int i = 0; i < 3; i++) h->h_delays[4 * a + i] = poly3[a][i];
    h->h_delays[4 * a + 3] = fileoffsets[a];
  }
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  GXC_HIP(h, hipMemcpy(h->d_delays, h->h_delays.data(), h->h_delays.size() * sizeof(double),
                       hipMemcpyHostToDevice));
  h->delays_set = true;
  return GXC_OK;
}

gxc_status gxc_process_async(gxc_handle *h) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  return enqueue_process(h);
}

gxc_status gxc_synchronize(gxc_handle *h) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  collect_timings(h);
  return GXC_OK;
}

gxc_status gxc_process(gxc_handle *h) {
  GXC_TRY(gxc_process_async(h));
  return gxc_synchronize(h);
}

gxc_status gxc_get_visibilities(gxc_handle *h, gxc_cf32 *out, size_t nelem) {
  if (!h || !out) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  const size_t n = (size_t)h->d.nrows * 4 * h->cfg.nchan;
  if (nelem < n) return set_err(h, GXC_ERR_INVALID_ARG, "output buffer too small");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  GXC_HIP(h, hipMemcpy(out, h->d_vis, n * sizeof(gxc_cf32), hipMemcpyDeviceToHost));
  return GXC_OK;
}

gxc_status gxc_get_row_counts(gxc_handle *h, int *counts, size_t n) {
  if (!h || !counts) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  if (n < (size_t)h->d.nrows) return set_err(h, GXC_ERR_INVALID_ARG, "output buffer too small");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  GXC_HIP(h, hipMemcpy(counts, h->d_counts, h->d.nrows * sizeof(int), hipMemcpyDeviceToHost));
  return GXC_OK;
}

const void *gxc_visibilities_device(const gxc_handle *h) { return h ? h->d_vis : nullptr; }

gxc_status gxc_accumulate(gxc_handle *h, gxc_cf32 *const *const *odata) {
  if (!h || !odata) return set_err(h, GXC_ERR_INVALID_ARG, "null output");
  const int nchan = h->cfg.nchan;
  std::vector<gxc_cf32> v((size_t)h->d.nrows * 4 * nchan);
  GXC_TRY(gxc_get_visibilities(h, v.data(), v.size()));
  // fxkernel.cpp:414-423: odata[b][p][c] += vis[b][p][c]
  for (int r = 0; r < h->d.nrows; r++)
    for (int p = 0; p < 4; p++) {
      gxc_cf32 *o = odata[r][p];
      const gxc_cf32 *s = &v[((size_t)r * 4 + p) * nchan];
      for (int c = 0; c < nchan; c++) {
        o[c].re += s[c].re;
        o[c].im += s[c].im;
      }
    }
  return GXC_OK;
}

gxc_status gxc_save_visibilities(gxc_handle *h, const char *path) {
  if (!h || !path) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  const int nchan = h->cfg.nchan, nrows = h->d.nrows;
  std::vector<gxc_cf32> v((size_t)nrows * 4 * nchan);
  GXC_TRY(gxc_get_visibilities(h, v.data(), v.size()));
  std::ofstream fvis(path);
  if (!fvis) return set_err(h, GXC_ERR_IO, std::string("cannot open ") + path);
  // fxkernel.cpp:389-399: channel, centre frequency (MHz), then per row and
  // product: re, im, amplitude, phase.
  for (int c = 0; c < nchan; c++) {
    fvis << std::setw(5) << c << " " << std::setw(11) << std::fixed << std::setprecision(6)
         << (c + 0.5) / nchan * h->cfg.bw / 1e6;
    fvis << std::setprecision(5);
    for (int r = 0; r < nrows; r++)
      for (int p = 0; p < 4; p++) {
        const gxc_cf32 x = v[((size_t)r * 4 + p) * nchan + c];
        const float amp = std::sqrt(x.re * x.re + x.im * x.im);
        const float ph = std::atan2(x.im, x.re);
        fvis << " " << std::setw(11) << x.re << " " << std::setw(11) << x.im;
        fvis << " " << std::setw(11) << amp << " " << std::setw(10) << ph;
      }
    fvis << std::endl;
  }
  if (!fvis) return set_err(h, GXC_ERR_IO, std::string("write failed: ") + path);
  return GXC_OK;
}

gxc_status gxc_fstage(gxc_handle *h, int t0, int nfft, gxc_cf32 *spectra, uint8_t *valid) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (h->cfg.precision != GXC_PREC_FP32)
    return set_err(h, GXC_ERR_UNSUPPORTED, "gxc_fstage runs the FP32 F-stage only");
  if (t0 < 0 || nfft < 1 || t0 + nfft > h->cfg.numffts)
    return set_err(h, GXC_ERR_INVALID_ARG, "FFT range outside the subintegration");
  if (!h->input_set || !h->delays_set) return set_err(h, GXC_ERR_STATE, "input or delays not set");
  if (!h->input_device) {
    if (!h->d_input_owned) {
      GXC_TRY(dmalloc(h, &h->d_input_owned, h->d.station_bytes * (size_t)h->cfg.nant));
      std::vector<const uint8_t *> ptrs(h->cfg.nant);
      for (int a = 0; a < h->cfg.nant; a++)
        ptrs[a] = h->d_input_owned + (size_t)a * h->d.station_bytes;
      GXC_HIP(h, hipMemcpy(h->d_station_ptrs, ptrs.data(), h->cfg.nant * sizeof(void *),
                           hipMemcpyHostToDevice));
    }
    for (int a = 0; a < h->cfg.nant; a++)
      GXC_HIP(h, hipMemcpyAsync(h->d_input_owned + (size_t)a * h->d.station_bytes, h->host_in[a],
                                h->d.station_bytes, hipMemcpyHostToDevice, h->stream));
  }
  GXC_TRY(ensure_spectra(h, nfft));
  GXC_TRY(run_fstage(h, t0, nfft));
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  const size_t per_fft = (size_t)h->d.M * h->cfg.nchan;
  if (spectra)
    GXC_HIP(h, hipMemcpy(spectra, h->d_spectra, per_fft * nfft * sizeof(float2), hipMemcpyDeviceToHost));
  if (valid)
    for (int a = 0; a < h->cfg.nant; a++)
      GXC_HIP(h, hipMemcpy(valid + (size_t)a * nfft, h->d_valid + (size_t)a * h->cfg.numffts + t0, nfft,
                           hipMemcpyDeviceToHost));
  return GXC_OK;
}

gxc_status gxc_xengine(const gxc_cf32 *spectra, int nant, int nchan, int nfft, int autos, int precision,
                       gxc_cf32 *out) {
  if (!spectra || !out || nant < 1 || nchan < 1 || nfft < 1)
    return set_err(nullptr, GXC_ERR_INVALID_ARG, "bad xengine arguments");
  int dev = 0;
  GXC_TRY(check_device(nullptr, &dev));
  if (precision != GXC_PREC_FP32)
    return gxc::mixed_xengine(spectra, nant, nchan, nfft, autos, precision, out, &g_thread_error);
  gxc_handle tmp;
  gxc_handle *h = &tmp;
  h->cfg.nant = nant;
  h->cfg.nchan = nchan;
  h->cfg.numffts = nfft;
  h->cfg.autos = autos;
  h->d.M = 2 * nant;
  h->d.nbl = nant * (nant - 1) / 2;
  h->d.nrows = h->d.nbl + (autos ? nant : 0);
  struct Cleanup {
    gxc_handle *h;
    ~Cleanup() {
      if (h->stream) hipStreamSynchronize(h->stream);
      hipFree(h->d_spectra);
      hipFree(h->d_partial);
      hipFree(h->d_valid);
      hipFree(h->d_counts);
      hipFree(h->d_rowmap);
      hipFree(h->d_vis);
      if (h->stream) hipStreamDestroy(h->stream);
    }
  } cleanup{h};
  GXC_HIP(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  GXC_TRY(ensure_spectra(h, nfft));
  const size_t nspec = (size_t)nfft * h->d.M * nchan;
  GXC_HIP(h, hipMemcpy(h->d_spectra, spectra, nspec * sizeof(float2), hipMemcpyHostToDevice));
  GXC_TRY(dmalloc(h, &h->d_valid, (size_t)nant * nfft));
  GXC_HIP(h, hipMemset(h->d_valid, 1, (size_t)nant * nfft));
  std::vector<int2> rm = make_rowmap(nant, autos != 0);
  GXC_TRY(dmalloc(h, &h->d_rowmap, rm.size()));
  if (!rm.empty()) GXC_HIP(h, hipMemcpy(h->d_rowmap, rm.data(), rm.size() * sizeof(int2), hipMemcpyHostToDevice));
  GXC_TRY(dmalloc(h, &h->d_counts, (size_t)h->d.nrows));
  GXC_TRY(dmalloc(h, &h->d_vis, (size_t)h->d.nrows * 4 * nchan));
  h->sb = gxc::xstage_block_stations(nant);
  h->nb = (nant + h->sb - 1) / h->sb;
  h->ntiles = h->nb * (h->nb + 1) / 2;
  h->splits = choose_splits(nchan, h->ntiles, nfft);
  const size_t ni = 2 * (size_t)h->sb;
  GXC_TRY(dmalloc(h, &h->d_partial, (size_t)h->splits * h->ntiles * ni * ni * nchan));
  GXC_TRY(run_xstage(h, nfft, 0));
  GXC_TRY(run_finalise(h));
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  GXC_HIP(h, hipMemcpy(out, h->d_vis, (size_t)h->d.nrows * 4 * nchan * sizeof(float2), hipMemcpyDeviceToHost));
  return GXC_OK;
}

gxc_status gxc_set_profiling(gxc_handle *h, int enable) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  h->profiling = enable != 0;
  return GXC_OK;
}

gxc_status gxc_get_timings(const gxc_handle *h, gxc_timings *t) {
  if (!h || !t) return GXC_ERR_INVALID_ARG;
  *t = h->timings;
  return GXC_OK;
}

}  // extern "C"
