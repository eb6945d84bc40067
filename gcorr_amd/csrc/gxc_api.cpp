// C-ABI of the FX-correlation hot path (include/gxcorr.h): FxKernel's
// processing interface (fxkernel/src/fxkernel.h:18-183) over HIP kernels.
//
// One subintegration on the device:
//   [H2D copy of host input]
//   [reduced precision: F-stage power pass on 64 sampled FFTs -> scales]
//   for each chunk of FFTs: F-stage -> spectra buffer (two buffers when
//     overlapped), X-stage (MFMA) accumulating split partials
//   row counts from the validity flags, finalise (sum partials, normalise)
// With overlap on, F runs on the handle's stream and X on a second stream,
// ordered by events per spectra buffer.  With graph on, everything after the
// input copy is captured once as a hipGraph and replayed on later calls.
#include "../../include/gxcorr.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "gxc_host.hpp"
#include "gxc_internal.hpp"
#include "gxc_lags.hpp"
#include "gxc_requant.hpp"
#include "gxc_vdif.hpp"
#include "gxc_xengine.hpp"

using gxc::Geometry;
using gxc::host::Derived;
using gxc::host::derive;
using gxc::host::make_rowmap;
using gxc::host::out_mode_of;

namespace {

thread_local std::string g_thread_error;

}  // namespace

struct gxc_handle {
  gxc_config cfg{};
  Derived d;
  int device = 0;
  hipStream_t stream = nullptr;   // F-stage, input copies, final join
  hipStream_t xstream = nullptr;  // X-stage when overlapped
  std::string err;

  // input
  bool input_set = false, input_device = false;
  std::vector<const uint8_t *> host_in;
  std::vector<const void *> dev_in;
  uint8_t *d_input_owned = nullptr;
  const uint8_t **d_station_ptrs = nullptr;

  // VDIF ingest (gxc_set_input_vdif): payloads go into d_input_owned, and
  // the per-station frame flags make FFTs over missing frames invalid
  bool vdif = false;                   // input comes from gxc_set_input_vdif
  std::vector<char> vdif_done;         // stations ingested since the switch
  std::vector<uint8_t *> d_flags;      // [nant] frame flags, capacity flag_cap
  std::vector<long long> flag_cap;
  const uint8_t **d_flag_ptrs = nullptr;  // [nant] = d_flags
  int *d_spf = nullptr;                   // [nant] time samples per frame (set by the ingest)
  long long *d_nslots = nullptr;          // [nant] slots of the last ingest (set by the ingest)
  unsigned *d_slot_src = nullptr;         // [slot_cap] ingest scratch
  long long slot_cap = 0;
  unsigned long long *d_vstats = nullptr;  // [kVdifCounters]

  // polarisation conversion of the visibilities (gxc_set_pol_conversion)
  bool polconv = false;
  float2 *d_polgains = nullptr;  // [nant][2][nchan]

  // delays
  bool delays_set = false;
  std::vector<double> h_delays;  // [nant][4]
  double *d_delays = nullptr;

  // constant tables
  float2 *d_twiddle = nullptr;
  float *d_subfreq = nullptr, *d_stepfreq = nullptr;

  // spectra, X-stage and output
  int out_mode = gxc::OUT_F32;
  int chunk = 0, nbuf = 1, splits = 1;
  int fblock = 0, xtile = 0;  // launch shapes (0 = default)
  int partial_cap = 0;        // splits the partial buffer holds
  void *d_spec[2] = {nullptr, nullptr};  // [chunk][M][nchan] packed complex
  float *d_partial = nullptr;             // [splits][nchan][2][M][M]
  float *d_scales = nullptr;              // [M][nchan] (reduced precision)
  float *d_power = nullptr;               // [kPowerRows][M][nchan]
  uint8_t *d_valid = nullptr;             // [nant][numffts]
  int *d_counts = nullptr;
  int2 *d_rowmap = nullptr;
  float2 *d_vis = nullptr;

  // ordering between the two streams
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_f[2] = {nullptr, nullptr}, ev_x[2] = {nullptr, nullptr};

  // graph replay
  hipGraphExec_t graph_exec = nullptr;

  // streaming integration (gxc_integrate_begin / gxc_submit / gxc_integrate_end):
  // two input slots, so the copy of subintegration k+1 (copy stream) runs
  // while subintegration k is correlated
  hipStream_t cstream = nullptr;
  struct Slot {
    uint8_t *d_in = nullptr;           // [nant][station_bytes]
    const uint8_t **d_ptrs = nullptr;  // [nant] into d_in
    double *d_delays = nullptr;        // [nant][4]
    double *h_delays = nullptr;        // pinned staging of the same
    hipEvent_t copied = nullptr, done = nullptr;
    bool used = false;
  } slot[2];
  float2 *d_acc = nullptr;  // sum of the subintegrations' visibilities
  long long nsubmitted = 0;
  bool integrating = false;

  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> tev;
  struct Mark {
    int kind;  // 0 F start, 1 F end, 2 X start, 3 X end, 4 finalise start, 5 finalise end, 6 H2D end
    size_t ev;
  };
  std::vector<Mark> marks;
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

#define GXC_TRY(expr)            \
  do {                           \
    gxc_status s_ = (expr);      \
    if (s_ != GXC_OK) return s_; \
  } while (0)

gxc_status check_device(gxc_handle *h, int *dev) {
  std::string why;
  const gxc_status st = gxc::host::check_device(dev, &why);
  return st == GXC_OK ? GXC_OK : set_err(h, st, why);
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

gxc_status upload_tables(gxc_handle *h) {
  std::string why;
  const gxc_status st = gxc::host::upload_tables(h->d, h->cfg.bw, h->cfg.nchan, &h->d_twiddle, &h->d_subfreq,
                                                 &h->d_stepfreq, &why);
  return st == GXC_OK ? GXC_OK : set_err(h, st, why);
}

void mark(gxc_handle *h, int kind, hipStream_t s) {
  if (!h->profiling) return;
  const size_t idx = h->marks.size();
  while (h->tev.size() <= idx) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    h->tev.push_back(e);
  }
  hipEventRecord(h->tev[idx], s);
  h->marks.push_back({kind, idx});
}

gxc_status run_fstage(gxc_handle *h, int t0, int nt, int mode, void *out, int tstride, hipStream_t s) {
  gxc::FParams p{};
  p.station = h->d_station_ptrs;
  p.delays = h->d_delays;
  p.subchanfreq = h->d_subfreq;
  p.stepchanfreq = h->d_stepfreq;
  p.twiddle = h->d_twiddle;
  p.spectra = static_cast<float2 *>(out);
  p.out = out;
  p.scales = h->d_scales;
  p.power = h->d_power;
  p.valid = h->d_valid;
  p.frame_ok = h->vdif ? h->d_flag_ptrs : nullptr;
  p.frame_samples = h->d_spf;
  p.out_mode = mode;
  p.t0 = t0;
  p.nt = nt;
  p.tstride = tstride;
  p.nant = h->cfg.nant;
  p.numffts = h->cfg.numffts;
  p.maxoffset = h->d.maxoffset;
  p.stepmul = h->cfg.exact_phase ? h->d.g.S2 : h->d.g.s;
  const double lo = h->cfg.lo;
  p.lo = lo;
  p.lofrac = lo - std::trunc(lo);
  p.fraclo = (lo - std::trunc(lo)) > 0.00000000001 ? 1 : 0;  // TINY, fxkernel.cpp:66-71
  p.sampletime = h->d.sampletime;
  GXC_HIP(h, gxc::launch_fstage(p, h->d.L, h->cfg.is_complex != 0, h->fblock, s));
  return GXC_OK;
}

gxc_status run_xstage(gxc_handle *h, const void *spec, int nt, int accumulate, hipStream_t s) {
  gxc::XMParams x{};
  x.spectra = spec;
  x.partial = h->d_partial;
  x.nt = nt;
  x.M = h->d.M;
  x.nchan = h->cfg.nchan;
  x.splits = h->splits;
  x.accumulate = accumulate;
  GXC_HIP(h, gxc::launch_xmfma(x, h->out_mode, h->xtile, s));
  return GXC_OK;
}

gxc_status run_finalise(gxc_handle *h, hipStream_t s) {
  gxc::CountParams cp{};
  cp.valid = h->d_valid;
  cp.rowmap = h->d_rowmap;
  cp.counts = h->d_counts;
  cp.numffts = h->cfg.numffts;
  cp.nrows = h->d.nrows;
  GXC_HIP(h, gxc::launch_counts(cp, s));
  gxc::FinMParams f{};
  f.partial = h->d_partial;
  f.scales = h->out_mode == gxc::OUT_F32 ? nullptr : h->d_scales;
  f.rowmap = h->d_rowmap;
  f.counts = h->d_counts;
  f.vis = h->d_vis;
  f.M = h->d.M;
  f.nchan = h->cfg.nchan;
  f.nrows = h->d.nrows;
  f.splits = h->splits;
  GXC_HIP(h, gxc::launch_finalise_mixed(f, s));
  return GXC_OK;
}

// Device work of one subintegration once the input is resident.
gxc_status enqueue_device_work(gxc_handle *h) {
  const gxc_config &c = h->cfg;
  const bool overlap = h->nbuf == 2;
  hipStream_t fs = h->stream, xs = overlap ? h->xstream : h->stream;
  if (h->out_mode != gxc::OUT_F32) {
    // Per-(input, channel) scales from |X|^2 of FFTs sampled across the subint.
    const int npow = std::min(c.numffts, gxc::kPowerRows);
    GXC_TRY(run_fstage(h, 0, npow, gxc::OUT_POWER, nullptr, c.numffts / npow, fs));
    GXC_HIP(h, gxc::launch_scales(h->d_scales, h->d_power, h->d.M * c.nchan, npow,
                                  gxc::xm_target_rms(h->out_mode), fs));
  }
  if (overlap) {
    GXC_HIP(h, hipEventRecord(h->ev_fork, fs));
    GXC_HIP(h, hipStreamWaitEvent(xs, h->ev_fork, 0));
  }
  int nchunks = 0;
  for (int t0 = 0; t0 < c.numffts; t0 += h->chunk, nchunks++) {
    const int nt = std::min(h->chunk, c.numffts - t0);
    const int b = overlap ? (nchunks & 1) : 0;
    if (overlap && nchunks >= 2) GXC_HIP(h, hipStreamWaitEvent(fs, h->ev_x[b], 0));
    mark(h, 0, fs);
    GXC_TRY(run_fstage(h, t0, nt, h->out_mode, h->d_spec[b], 1, fs));
    mark(h, 1, fs);
    if (overlap) {
      GXC_HIP(h, hipEventRecord(h->ev_f[b], fs));
      GXC_HIP(h, hipStreamWaitEvent(xs, h->ev_f[b], 0));
    }
    mark(h, 2, xs);
    GXC_TRY(run_xstage(h, h->d_spec[b], nt, t0 > 0 ? 1 : 0, xs));
    mark(h, 3, xs);
    if (overlap) GXC_HIP(h, hipEventRecord(h->ev_x[b], xs));
  }
  mark(h, 4, xs);
  GXC_TRY(run_finalise(h, xs));
  if (h->polconv)
    GXC_HIP(h, gxc::launch_polconv(h->d_vis, h->d_polgains, h->d_rowmap, h->d.nrows, c.nchan, xs));
  mark(h, 5, xs);
  if (overlap) {
    GXC_HIP(h, hipEventRecord(h->ev_join, xs));
    GXC_HIP(h, hipStreamWaitEvent(fs, h->ev_join, 0));
  }
  h->timings.fstage_launches = nchunks;
  h->timings.xstage_launches = nchunks;
  return GXC_OK;
}

gxc_status ensure_owned_input(gxc_handle *h) {
  if (h->d_input_owned) return GXC_OK;
  const int nant = h->cfg.nant;
  GXC_TRY(dmalloc(h, &h->d_input_owned, h->d.station_bytes * (size_t)nant));
  std::vector<const uint8_t *> ptrs(nant);
  for (int a = 0; a < nant; a++) ptrs[a] = h->d_input_owned + (size_t)a * h->d.station_bytes;
  GXC_HIP(h, hipMemcpy(h->d_station_ptrs, ptrs.data(), nant * sizeof(void *), hipMemcpyHostToDevice));
  return GXC_OK;
}

gxc_status enqueue_input_copy(gxc_handle *h) {
  if (h->input_device || h->vdif) return GXC_OK;
  GXC_TRY(ensure_owned_input(h));
  for (int a = 0; a < h->cfg.nant; a++)
    GXC_HIP(h, hipMemcpyAsync(h->d_input_owned + (size_t)a * h->d.station_bytes, h->host_in[a],
                              h->d.station_bytes, hipMemcpyHostToDevice, h->stream));
  return GXC_OK;
}

gxc_status enqueue_process(gxc_handle *h) {
  if (!h->input_set) return set_err(h, GXC_ERR_STATE, "input data not set");
  if (!h->delays_set) return set_err(h, GXC_ERR_STATE, "delays not set");
  h->marks.clear();
  mark(h, 7, h->stream);
  GXC_TRY(enqueue_input_copy(h));
  mark(h, 6, h->stream);
  const bool use_graph = h->cfg.graph && !h->profiling;
  if (!use_graph) return enqueue_device_work(h);
  if (!h->graph_exec) {
    hipGraph_t graph = nullptr;
    GXC_HIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    gxc_status s = enqueue_device_work(h);
    hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (s != GXC_OK) {
      if (graph) hipGraphDestroy(graph);
      return s;
    }
    if (e != hipSuccess) return set_err(h, GXC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&h->graph_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      h->graph_exec = nullptr;
      return set_err(h, GXC_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
  }
  GXC_HIP(h, hipGraphLaunch(h->graph_exec, h->stream));
  return GXC_OK;
}

float elapsed(gxc_handle *h, size_t a, size_t b) {
  float ms = 0.f;
  if (a < h->tev.size() && b < h->tev.size() && hipEventElapsedTime(&ms, h->tev[a], h->tev[b]) == hipSuccess)
    return ms;
  return 0.f;
}

void collect_timings(gxc_handle *h) {
  if (!h->profiling || h->marks.empty()) return;
  gxc_timings &t = h->timings;
  t.h2d_ms = t.fstage_ms = t.xstage_ms = t.finalise_ms = t.total_ms = 0;
  size_t first = h->marks.front().ev, last = h->marks.back().ev;
  size_t open[8] = {0};
  for (const auto &m : h->marks) {
    switch (m.kind) {
      case 7: open[7] = m.ev; break;
      case 6: t.h2d_ms += elapsed(h, open[7], m.ev); break;
      case 0: open[0] = m.ev; break;
      case 1: t.fstage_ms += elapsed(h, open[0], m.ev); break;
      case 2: open[2] = m.ev; break;
      case 3: t.xstage_ms += elapsed(h, open[2], m.ev); break;
      case 4: open[4] = m.ev; break;
      case 5: t.finalise_ms += elapsed(h, open[4], m.ev); break;
    }
  }
  t.total_ms = elapsed(h, first, last);
}

void destroy_graph(gxc_handle *h) {
  if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
  h->graph_exec = nullptr;
}

// Leave VDIF mode: the frame flags stop applying (a captured graph holds them).
void leave_vdif(gxc_handle *h) {
  if (!h->vdif) return;
  h->vdif = false;
  destroy_graph(h);
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
  cfg->overlap = 0;  // both stages fill the device: overlap measured no faster
  cfg->graph = 1;
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

const char *gxc_last_error(const gxc_handle *h) { return h ? h->err.c_str() : g_thread_error.c_str(); }

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
  const int out_mode = out_mode_of(cfg->precision);
  const int ct = gxc::xm_channels_per_group(d.M, out_mode);
  if (cfg->nchan % ct != 0)
    return set_err(nullptr, GXC_ERR_UNSUPPORTED,
                   "nchan must be a multiple of " + std::to_string(ct) + " for " + std::to_string(cfg->nant) +
                       " stations at this precision");
  int dev = 0;
  s = check_device(nullptr, &dev);
  if (s != GXC_OK) return s;
  gxc_handle *h = new gxc_handle();
  h->cfg = *cfg;
  h->d = d;
  h->device = dev;
  h->out_mode = out_mode;
  auto fail = [&](gxc_status st) {
    g_thread_error = h->err;
    gxc_destroy(h);
    return st;
  };
  // Default priority streams only.
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->xstream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(h, GXC_ERR_HIP, "hipStreamCreate failed"));
  hipEvent_t *evs[6] = {&h->ev_fork, &h->ev_join, &h->ev_f[0], &h->ev_f[1], &h->ev_x[0], &h->ev_x[1]};
  for (hipEvent_t *e : evs)
    if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
      return fail(set_err(h, GXC_ERR_HIP, "hipEventCreate failed"));
  if ((s = upload_tables(h)) != GXC_OK) return fail(s);
  const int nant = cfg->nant, nchan = cfg->nchan;
  h->h_delays.assign((size_t)nant * 4, 0.0);
  if ((s = dmalloc(h, &h->d_delays, (size_t)nant * 4)) != GXC_OK) return fail(s);
  if ((s = dmalloc(h, &h->d_station_ptrs, (size_t)nant)) != GXC_OK) return fail(s);
  if ((s = dmalloc(h, &h->d_valid, (size_t)nant * cfg->numffts)) != GXC_OK) return fail(s);
  std::vector<int2> rm = make_rowmap(nant, cfg->autos != 0);
  if ((s = dmalloc(h, &h->d_rowmap, rm.size())) != GXC_OK) return fail(s);
  if (!rm.empty() &&
      hipMemcpy(h->d_rowmap, rm.data(), rm.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_err(h, GXC_ERR_HIP, "rowmap upload failed"));
  if ((s = dmalloc(h, &h->d_counts, (size_t)d.nrows)) != GXC_OK) return fail(s);
  if ((s = dmalloc(h, &h->d_vis, (size_t)d.nrows * 4 * nchan)) != GXC_OK) return fail(s);
  if (hipMemset(h->d_vis, 0, (size_t)d.nrows * 4 * nchan * sizeof(float2)) != hipSuccess)
    return fail(set_err(h, GXC_ERR_HIP, "hipMemset failed"));
  // Spectra buffers: one chunk each, two when overlapped.  Automatic chunks:
  // the whole subint in one pass without overlap (bounded by 8 GiB), eight
  // passes with overlap (bounded by 4 GiB per buffer).
  h->nbuf = cfg->overlap ? 2 : 1;
  const size_t eb = gxc::xm_element_bytes(out_mode);
  const size_t per_fft = (size_t)d.M * nchan * eb;
  long long chunk = cfg->chunk_ffts;
  if (chunk <= 0) {
    const size_t budget = cfg->overlap ? (4ull << 30) : (8ull << 30);
    chunk = cfg->overlap ? (cfg->numffts + 7) / 8 : cfg->numffts;
    chunk = std::min<long long>(chunk, (long long)(budget / per_fft));
  }
  chunk = std::max<long long>(1, std::min<long long>(chunk, cfg->numffts));
  if (chunk >= cfg->numffts) h->nbuf = 1;  // one pass: nothing to overlap
  h->chunk = (int)chunk;
  for (int b = 0; b < h->nbuf; b++)
    if ((s = dmalloc(h, reinterpret_cast<unsigned char **>(&h->d_spec[b]), (size_t)h->chunk * per_fft)) !=
        GXC_OK)
      return fail(s);
  // launch shapes: given, or defaults
  {
    int fb[4], nfb = gxc::fstage_blocks(d.L, cfg->is_complex != 0, fb);
    bool ok = cfg->fblock == 0 ||
              (cfg->fblock == gxc::kFBlockMfma && out_mode != gxc::OUT_F32 &&
               gxc::fstage_mfma_supported(d.L, cfg->is_complex != 0));
    for (int i = 0; i < nfb; i++) ok = ok || cfg->fblock == fb[i];
    if (!ok) return fail(set_err(h, GXC_ERR_INVALID_ARG, "fblock is not an F-stage workgroup size for this FFT length"));
    int xt[3], nxt = gxc::xm_tiles(d.M, out_mode, xt);
    ok = cfg->xtile == 0;
    for (int i = 0; i < nxt; i++) ok = ok || (cfg->xtile == xt[i] && nchan % xt[i] == 0);
    if (!ok) return fail(set_err(h, GXC_ERR_INVALID_ARG, "xtile is not an X-stage channel tile for this shape"));
    if (cfg->xsplits < 0 || cfg->xsplits > 1024) return fail(set_err(h, GXC_ERR_INVALID_ARG, "xsplits out of range"));
    h->fblock = cfg->fblock;
    h->xtile = cfg->xtile;
  }
  h->splits = cfg->xsplits > 0 ? std::min(cfg->xsplits, h->chunk) : gxc::xm_choose_splits(d.M, nchan, h->chunk, out_mode);
  h->partial_cap = h->splits;
  if ((s = dmalloc(h, reinterpret_cast<unsigned char **>(&h->d_partial),
                   gxc::xm_partial_bytes(d.M, nchan, h->splits))) != GXC_OK)
    return fail(s);
  if (out_mode != gxc::OUT_F32) {
    if ((s = dmalloc(h, &h->d_scales, (size_t)d.M * nchan)) != GXC_OK) return fail(s);
    if ((s = dmalloc(h, &h->d_power, (size_t)gxc::kPowerRows * d.M * nchan)) != GXC_OK) return fail(s);
  }
  *out = h;
  return GXC_OK;
}

void gxc_destroy(gxc_handle *h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->xstream) hipStreamSynchronize(h->xstream);
  destroy_graph(h);
  for (hipEvent_t e : h->tev) hipEventDestroy(e);
  hipEvent_t evs[6] = {h->ev_fork, h->ev_join, h->ev_f[0], h->ev_f[1], h->ev_x[0], h->ev_x[1]};
  for (hipEvent_t e : evs)
    if (e) hipEventDestroy(e);
  for (uint8_t *f : h->d_flags) hipFree(f);
  hipFree(h->d_flag_ptrs);
  hipFree(h->d_polgains);
  hipFree(h->d_spf);
  hipFree(h->d_nslots);
  hipFree(h->d_slot_src);
  hipFree(h->d_vstats);
  if (h->cstream) hipStreamSynchronize(h->cstream);
  for (auto &sl : h->slot) {
    hipFree(sl.d_in);
    hipFree(sl.d_ptrs);
    hipFree(sl.d_delays);
    if (sl.h_delays) hipHostFree(sl.h_delays);
    if (sl.copied) hipEventDestroy(sl.copied);
    if (sl.done) hipEventDestroy(sl.done);
  }
  hipFree(h->d_acc);
  if (h->cstream) hipStreamDestroy(h->cstream);
  for (void *p : h->d_spec) hipFree(p);
  hipFree(h->d_partial);
  hipFree(h->d_scales);
  hipFree(h->d_power);
  hipFree(h->d_input_owned);
  hipFree(h->d_station_ptrs);
  hipFree(h->d_delays);
  hipFree(h->d_twiddle);
  hipFree(h->d_subfreq);
  hipFree(h->d_stepfreq);
  hipFree(h->d_valid);
  hipFree(h->d_counts);
  hipFree(h->d_rowmap);
  hipFree(h->d_vis);
  if (h->xstream) hipStreamDestroy(h->xstream);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

gxc_status gxc_set_input(gxc_handle *h, const uint8_t *const *per_station) {
  if (!h || !per_station) return set_err(h, GXC_ERR_INVALID_ARG, "null input");
  std::vector<const uint8_t *> in(per_station, per_station + h->cfg.nant);
  for (const uint8_t *p : in)
    if (!p) return set_err(h, GXC_ERR_INVALID_ARG, "null station buffer");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  leave_vdif(h);
  h->host_in = in;
  if (h->input_device || !h->d_input_owned) {
    h->input_device = false;
    GXC_TRY(ensure_owned_input(h));
    // restore the station pointers to the owned copy
    std::vector<const uint8_t *> ptrs(h->cfg.nant);
    for (int a = 0; a < h->cfg.nant; a++) ptrs[a] = h->d_input_owned + (size_t)a * h->d.station_bytes;
    GXC_HIP(h, hipMemcpy(h->d_station_ptrs, ptrs.data(), h->cfg.nant * sizeof(void *), hipMemcpyHostToDevice));
  }
  h->input_device = false;
  h->input_set = true;
  return GXC_OK;
}

gxc_status gxc_set_input_device(gxc_handle *h, const void *const *per_station) {
  if (!h || !per_station) return set_err(h, GXC_ERR_INVALID_ARG, "null input");
  std::vector<const void *> in(per_station, per_station + h->cfg.nant);
  for (const void *p : in) {
    if (!p) return set_err(h, GXC_ERR_INVALID_ARG, "null station buffer");
    if (reinterpret_cast<uintptr_t>(p) % 4)
      return set_err(h, GXC_ERR_INVALID_ARG, "device station buffers must be 4-byte aligned");
  }
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  leave_vdif(h);
  h->dev_in = in;
  GXC_HIP(h, hipMemcpy(h->d_station_ptrs, h->dev_in.data(), h->cfg.nant * sizeof(void *), hipMemcpyHostToDevice));
  h->input_device = true;
  h->input_set = true;
  return GXC_OK;
}

void gxc_vdif_spec_init(gxc_vdif_spec *spec) {
  if (!spec) return;
  std::memset(spec, 0, sizeof(*spec));
  spec->thread_id = 0;  // dump_vdif.py keeps thread 0 unless told otherwise
  spec->offset_binary8 = 1;
}

long long gxc_vdif_slots(const gxc_config *cfg, int frame_bytes, int legacy_header) {
  Derived d;
  std::string why;
  if (derive(cfg, &d, &why) != GXC_OK) return 0;
  const long long payload = (long long)frame_bytes - (legacy_header ? 16 : 32);
  const long long bits = (long long)cfg->nbit * 2 * (cfg->is_complex ? 2 : 1);
  if (payload <= 0 || payload % 8 || (payload * 8) % bits) return 0;
  return ((long long)d.station_bytes + payload - 1) / payload;
}

gxc_status gxc_set_input_vdif(gxc_handle *h, int station, const void *d_frames, size_t nbytes,
                              const gxc_vdif_spec *spec, gxc_vdif_stats *stats) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  const int nant = h->cfg.nant;
  if (station < 0 || station >= nant) return set_err(h, GXC_ERR_INVALID_ARG, "station out of range");
  if (!d_frames || !spec) return set_err(h, GXC_ERR_INVALID_ARG, "null frames or spec");
  if (reinterpret_cast<uintptr_t>(d_frames) % 8)
    return set_err(h, GXC_ERR_INVALID_ARG, "VDIF frames must be 8-byte aligned in device memory");
  if (spec->frames_per_second <= 0) return set_err(h, GXC_ERR_INVALID_ARG, "frames_per_second must be positive");
  if (spec->frame_bytes < 0) return set_err(h, GXC_ERR_INVALID_ARG, "negative frame_bytes");
  if (nbytes < 32) return set_err(h, GXC_ERR_INVALID_ARG, "VDIF stream shorter than one header");
  long long fb = spec->frame_bytes;
  if (fb == 0) {  // frame size from the first header: the one host read of the stream
    uint32_t w[4];
    GXC_HIP(h, hipStreamSynchronize(h->stream));
    GXC_HIP(h, hipMemcpy(w, d_frames, sizeof(w), hipMemcpyDeviceToHost));
    fb = 8LL * (w[2] & 0xffffffu);
  }
  // The header size (16 legacy, 32 otherwise) is read from the first frame on
  // the device; buffers are sized for the larger slot count of the two.
  const long long nslots_max = gxc_vdif_slots(&h->cfg, (int)std::min<long long>(fb, INT32_MAX), 0);
  if (fb <= 32 || fb > INT32_MAX || fb % 8 || nslots_max <= 0 ||
      gxc_vdif_slots(&h->cfg, (int)fb, 1) <= 0)
    return set_err(h, GXC_ERR_INVALID_ARG,
                   "VDIF frame size " + std::to_string(fb) + " does not hold a whole number of samples");
  const long long nframes = (long long)(nbytes / (size_t)fb);
  if (nframes >= (long long)UINT32_MAX) return set_err(h, GXC_ERR_INVALID_ARG, "too many VDIF frames in one call");
  if (h->d.station_bytes / 8 >= (size_t)UINT32_MAX)
    return set_err(h, GXC_ERR_UNSUPPORTED, "station buffers of 32 GiB or more are not supported by the VDIF ingest");

  const bool grow = !h->d_flag_ptrs || h->flag_cap[station] < nslots_max || h->slot_cap < nslots_max;
  if (grow || !h->vdif || !h->d_input_owned) {
    // buffers change under work that may still be queued: wait for it
    GXC_HIP(h, hipStreamSynchronize(h->stream));
  }
  GXC_TRY(ensure_owned_input(h));
  if (!h->d_flag_ptrs) {
    GXC_TRY(dmalloc(h, &h->d_flag_ptrs, (size_t)nant));
    GXC_TRY(dmalloc(h, &h->d_spf, (size_t)nant));
    GXC_TRY(dmalloc(h, &h->d_nslots, (size_t)nant));
    GXC_TRY(dmalloc(h, &h->d_vstats, (size_t)gxc::kVdifCounters));
    h->d_flags.assign(nant, nullptr);
    h->flag_cap.assign(nant, 0);
  }
  if (h->flag_cap[station] < nslots_max) {
    hipFree(h->d_flags[station]);
    h->d_flags[station] = nullptr;
    h->flag_cap[station] = 0;
    GXC_TRY(dmalloc(h, &h->d_flags[station], (size_t)nslots_max));
    h->flag_cap[station] = nslots_max;
    GXC_HIP(h, hipMemcpy(h->d_flag_ptrs + station, &h->d_flags[station], sizeof(void *), hipMemcpyHostToDevice));
  }
  if (h->slot_cap < nslots_max) {
    hipFree(h->d_slot_src);
    h->d_slot_src = nullptr;
    h->slot_cap = 0;
    GXC_TRY(dmalloc(h, &h->d_slot_src, (size_t)nslots_max));
    h->slot_cap = nslots_max;
  }
  if (!h->vdif) {
    // enter VDIF mode: station pointers to the owned buffer, flags on
    std::vector<const uint8_t *> ptrs(nant);
    for (int a = 0; a < nant; a++) ptrs[a] = h->d_input_owned + (size_t)a * h->d.station_bytes;
    GXC_HIP(h, hipMemcpy(h->d_station_ptrs, ptrs.data(), nant * sizeof(void *), hipMemcpyHostToDevice));
    destroy_graph(h);
    h->vdif = true;
    h->input_device = false;
    h->input_set = false;
    h->vdif_done.assign(nant, 0);
  }

  gxc::VdifParams vp{};
  vp.frames = static_cast<const uint8_t *>(d_frames);
  vp.nframes = nframes;
  vp.frame_bytes = (int)fb;
  vp.thread_id = spec->thread_id;
  vp.fps = spec->frames_per_second;
  vp.ref_epoch = spec->ref_epoch;
  vp.ref_frame = spec->ref_frame;
  vp.ref_second = spec->ref_second;
  vp.nbit = h->cfg.nbit;
  vp.is_complex = h->cfg.is_complex ? 1 : 0;
  vp.flip8 = h->cfg.nbit == 8 && spec->offset_binary8 ? 1 : 0;
  vp.slot_src = h->d_slot_src;
  vp.nslots_max = nslots_max;
  vp.station = h->d_input_owned + (size_t)station * h->d.station_bytes;
  vp.station_bytes = h->d.station_bytes;
  vp.frame_ok = h->d_flags[station];
  vp.spf_out = h->d_spf + station;
  vp.nslots_out = h->d_nslots + station;
  vp.stats = h->d_vstats;
  GXC_HIP(h, gxc::launch_vdif(vp, stats != nullptr, h->stream));
  h->vdif_done[station] = 1;
  h->input_set = std::all_of(h->vdif_done.begin(), h->vdif_done.end(), [](char c) { return c != 0; });
  if (stats) {
    unsigned long long c[gxc::kVdifCounters];
    GXC_HIP(h, hipStreamSynchronize(h->stream));
    GXC_HIP(h, hipMemcpy(c, h->d_vstats, sizeof(c), hipMemcpyDeviceToHost));
    stats->frames = nframes;
    stats->placed = (long long)c[gxc::kVdifPlaced];
    stats->missing = (long long)c[gxc::kVdifMissing];
    stats->duplicate = (long long)(c[gxc::kVdifCandidates] - c[gxc::kVdifPlaced]);
    stats->invalid = (long long)c[gxc::kVdifInvalid];
    stats->other_thread = (long long)c[gxc::kVdifOtherThread];
    stats->out_of_window = (long long)c[gxc::kVdifOutOfWindow];
    stats->malformed = (long long)c[gxc::kVdifMalformed];
  }
  return GXC_OK;
}

gxc_status gxc_get_frame_flags(gxc_handle *h, int station, uint8_t *out, size_t n, long long *nslots) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (station < 0 || station >= h->cfg.nant) return set_err(h, GXC_ERR_INVALID_ARG, "station out of range");
  if (!h->vdif || !h->vdif_done[station]) return set_err(h, GXC_ERR_STATE, "station has no VDIF ingest");
  long long ns = 0;
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  GXC_HIP(h, hipMemcpy(&ns, h->d_nslots + station, sizeof(ns), hipMemcpyDeviceToHost));
  if (nslots) *nslots = ns;
  if (!out) return GXC_OK;
  if (n < (size_t)ns) return set_err(h, GXC_ERR_INVALID_ARG, "flag buffer too small");
  GXC_HIP(h, hipMemcpy(out, h->d_flags[station], (size_t)ns, hipMemcpyDeviceToHost));
  return GXC_OK;
}

namespace {
gxc_status requant_args(const float *a, const float *b, const void *out, size_t outalign) {
  if (!a || !b || !out) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null buffer");
  if (reinterpret_cast<uintptr_t>(a) % 16 || reinterpret_cast<uintptr_t>(b) % 16 ||
      reinterpret_cast<uintptr_t>(out) % outalign)
    return set_err(nullptr, GXC_ERR_INVALID_ARG, "misaligned requantiser buffer");
  return GXC_OK;
}
}  // namespace

gxc_status gxc_requantise_2bit(const float *d_pol0, const float *d_pol1, size_t nsamples, float threshold,
                               uint8_t *d_out, void *stream) {
  GXC_TRY(requant_args(d_pol0, d_pol1, d_out, 8));
  if (nsamples % 2) return set_err(nullptr, GXC_ERR_INVALID_ARG, "2-bit requantiser needs an even sample count");
  if (!(threshold > 0.f)) return set_err(nullptr, GXC_ERR_INVALID_ARG, "threshold must be positive");
  GXC_HIP(nullptr, gxc::launch_requant2(d_pol0, d_pol1, (long long)nsamples, threshold, d_out,
                                        static_cast<hipStream_t>(stream)));
  return GXC_OK;
}

gxc_status gxc_requantise_8bit(const float *d_pol0, const float *d_pol1, size_t nsamples, float scale,
                               uint8_t *d_out, void *stream) {
  GXC_TRY(requant_args(d_pol0, d_pol1, d_out, 16));
  GXC_HIP(nullptr, gxc::launch_requant8(d_pol0, d_pol1, (long long)nsamples, scale, d_out,
                                        static_cast<hipStream_t>(stream)));
  return GXC_OK;
}

gxc_status gxc_set_pol_conversion(gxc_handle *h, const gxc_cf32 *gains) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  const bool on = gains != nullptr;
  if (on) {
    const size_t n = (size_t)h->cfg.nant * 2 * h->cfg.nchan;
    if (!h->d_polgains) GXC_TRY(dmalloc(h, &h->d_polgains, n));
    GXC_HIP(h, hipMemcpy(h->d_polgains, gains, n * sizeof(float2), hipMemcpyHostToDevice));
  }
  if (on != h->polconv) destroy_graph(h);  // the captured work changes
  h->polconv = on;
  return GXC_OK;
}

gxc_status gxc_set_delays(gxc_handle *h, const double *const *poly3, const double *fileoffsets) {
  if (!h || !poly3 || !fileoffsets) return set_err(h, GXC_ERR_INVALID_ARG, "null delays");
  for (int a = 0; a < h->cfg.nant; a++)
    if (!poly3[a]) return set_err(h, GXC_ERR_INVALID_ARG, "null delay polynomial");
  for (int a = 0; a < h->cfg.nant; a++) {
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

gxc_status gxc_save_visibilities_binary(gxc_handle *h, const char *path) {
  if (!h || !path) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  const int nchan = h->cfg.nchan, nrows = h->d.nrows;
  std::vector<gxc_cf32> v((size_t)nrows * 4 * nchan);
  std::vector<int> counts(nrows);
  GXC_TRY(gxc_get_visibilities(h, v.data(), v.size()));
  GXC_TRY(gxc_get_row_counts(h, counts.data(), counts.size()));
  const std::vector<int2> rows = gxc::host::make_rowmap(h->cfg.nant, h->cfg.autos != 0);
  std::vector<int32_t> st(2 * (size_t)nrows);
  for (int r = 0; r < nrows; r++) {
    st[2 * r] = rows[r].x;
    st[2 * r + 1] = rows[r].y;
  }
  std::ofstream f(path, std::ios::binary);
  if (!f) return set_err(h, GXC_ERR_IO, std::string("cannot open ") + path);
  const int32_t hdr[8] = {1, h->cfg.nant, nchan, nrows, 4, h->cfg.autos ? 1 : 0, h->cfg.numffts, h->cfg.precision};
  const double freq[2] = {h->cfg.bw, h->cfg.lo};
  f.write("GXCVIS01", 8);
  f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
  f.write(reinterpret_cast<const char *>(freq), sizeof(freq));
  f.write(reinterpret_cast<const char *>(st.data()), st.size() * sizeof(int32_t));
  f.write(reinterpret_cast<const char *>(counts.data()), counts.size() * sizeof(int32_t));
  f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(gxc_cf32));
  if (!f) return set_err(h, GXC_ERR_IO, std::string("write failed: ") + path);
  return GXC_OK;
}

gxc_status gxc_save_difx(gxc_handle *h, const char *path, const gxc_difx_meta *meta) {
  if (!h || !path || !meta) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  const int nchan = h->cfg.nchan, nrows = h->d.nrows;
  std::vector<gxc_cf32> v((size_t)nrows * 4 * nchan);
  std::vector<int> counts(nrows);
  GXC_TRY(gxc_get_visibilities(h, v.data(), v.size()));
  GXC_TRY(gxc_get_row_counts(h, counts.data(), counts.size()));
  const std::vector<int2> rows = gxc::host::make_rowmap(h->cfg.nant, h->cfg.autos != 0);
  std::ofstream f(path, std::ios::binary);
  if (!f) return set_err(h, GXC_ERR_IO, std::string("cannot open ") + path);
  for (int r = 0; r < nrows; r++)
    for (int p = 0; p < 4; p++) {
      // DiFX binary record header, 74 bytes, packed
      char rec[74];
      char *q = rec;
      auto put = [&q](const void *x, size_t n) {
        std::memcpy(q, x, n);
        q += n;
      };
      const int32_t sync = (int32_t)0xFF00FF00u, version = 1;
      const int32_t baseline = 256 * (rows[r].x + 1) + (rows[r].y + 1);
      const int32_t mjd = meta->mjd, cfgi = meta->config_index, src = meta->source_index, fq = meta->freq_index;
      const int32_t bin = 0;
      const double sec = meta->seconds;
      const double weight = h->cfg.numffts ? (double)counts[r] / h->cfg.numffts : 0.0;
      const double uvw[3] = {meta->uvw ? meta->uvw[3 * r] : 0.0, meta->uvw ? meta->uvw[3 * r + 1] : 0.0,
                             meta->uvw ? meta->uvw[3 * r + 2] : 0.0};
      const char pol[2] = {meta->pols[p >> 1], meta->pols[p & 1]};  // p = 2l + m
      put(&sync, 4);
      put(&version, 4);
      put(&baseline, 4);
      put(&mjd, 4);
      put(&sec, 8);
      put(&cfgi, 4);
      put(&src, 4);
      put(&fq, 4);
      put(pol, 2);
      put(&bin, 4);
      put(&weight, 8);
      put(uvw, 24);
      f.write(rec, sizeof(rec));
      f.write(reinterpret_cast<const char *>(&v[((size_t)r * 4 + p) * nchan]), (size_t)nchan * sizeof(gxc_cf32));
    }
  if (!f) return set_err(h, GXC_ERR_IO, std::string("write failed: ") + path);
  return GXC_OK;
}

namespace {
gxc_status lag_snr_on(gxc_handle *h, const void *src, hipMemcpyKind kind, int nseries, int nchan, float *snr,
                      int *peak, hipStream_t s) {
  if (nchan < 2 || (nchan & (nchan - 1))) return set_err(h, GXC_ERR_INVALID_ARG, "nchan must be a power of two >= 2");
  if (nseries < 1) return set_err(h, GXC_ERR_INVALID_ARG, "no series");
  const size_t n = (size_t)nseries * nchan;
  float2 *a = nullptr, *b = nullptr;
  float *dsnr = nullptr;
  int *dpeak = nullptr;
  gxc_status st = dmalloc(h, &a, n);
  if (st == GXC_OK) st = dmalloc(h, &b, n);
  if (st == GXC_OK) st = dmalloc(h, &dsnr, (size_t)nseries);
  if (st == GXC_OK) st = dmalloc(h, &dpeak, (size_t)nseries);
  hipError_t e = hipSuccess;
  if (st == GXC_OK) e = hipMemcpyAsync(a, src, n * sizeof(float2), kind, s);
  if (st == GXC_OK && e == hipSuccess) e = gxc::launch_lag_snr(a, b, nchan, nseries, dsnr, dpeak, s);
  if (st == GXC_OK && e == hipSuccess) e = hipMemcpyAsync(snr, dsnr, nseries * sizeof(float), hipMemcpyDeviceToHost, s);
  if (st == GXC_OK && e == hipSuccess) e = hipMemcpyAsync(peak, dpeak, nseries * sizeof(int), hipMemcpyDeviceToHost, s);
  if (st == GXC_OK && e == hipSuccess) e = hipStreamSynchronize(s);
  if (st == GXC_OK && e != hipSuccess) st = set_err(h, GXC_ERR_HIP, std::string("lag S/N: ") + hipGetErrorString(e));
  hipFree(a);
  hipFree(b);
  hipFree(dsnr);
  hipFree(dpeak);
  return st;
}
}  // namespace

gxc_status gxc_lag_snr(gxc_handle *h, float *snr, int *peak, size_t n) {
  if (!h || !snr || !peak) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  const int nseries = h->d.nrows * 4;
  if (n < (size_t)nseries) return set_err(h, GXC_ERR_INVALID_ARG, "output arrays too small");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  return lag_snr_on(h, h->d_vis, hipMemcpyDeviceToDevice, nseries, h->cfg.nchan, snr, peak, h->stream);
}

gxc_status gxc_lag_snr_series(const gxc_cf32 *vis, int nseries, int nchan, float *snr, int *peak) {
  if (!vis || !snr || !peak) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null argument");
  int dev = 0;
  GXC_TRY(check_device(nullptr, &dev));
  return lag_snr_on(nullptr, vis, hipMemcpyHostToDevice, nseries, nchan, snr, peak, nullptr);
}

gxc_status gxc_fstage(gxc_handle *h, int t0, int nfft, gxc_cf32 *spectra, uint8_t *valid) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (h->cfg.precision != GXC_PREC_FP32)
    return set_err(h, GXC_ERR_UNSUPPORTED, "gxc_fstage runs the FP32 F-stage only");
  if (t0 < 0 || nfft < 1 || t0 + nfft > h->cfg.numffts)
    return set_err(h, GXC_ERR_INVALID_ARG, "FFT range outside the subintegration");
  if (!h->input_set || !h->delays_set) return set_err(h, GXC_ERR_STATE, "input or delays not set");
  GXC_TRY(enqueue_input_copy(h));
  const size_t per_fft = (size_t)h->d.M * h->cfg.nchan;
  float2 *tmp = nullptr;
  GXC_TRY(dmalloc(h, &tmp, per_fft * nfft));
  gxc_status s = run_fstage(h, t0, nfft, gxc::OUT_F32, tmp, 1, h->stream);
  if (s == GXC_OK && hipStreamSynchronize(h->stream) != hipSuccess) s = set_err(h, GXC_ERR_HIP, "F-stage failed");
  if (s == GXC_OK && spectra &&
      hipMemcpy(spectra, tmp, per_fft * nfft * sizeof(float2), hipMemcpyDeviceToHost) != hipSuccess)
    s = set_err(h, GXC_ERR_HIP, "copy of spectra failed");
  hipFree(tmp);
  GXC_TRY(s);
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
  if (precision < GXC_PREC_FP32 || precision > GXC_PREC_FP8)
    return set_err(nullptr, GXC_ERR_INVALID_ARG, "unknown precision");
  const int M = 2 * nant, mode = out_mode_of(precision);
  if (gxc::xm_blocks(M) == 0 || nchan % gxc::xm_channels_per_group(M, mode) != 0)
    return set_err(nullptr, GXC_ERR_UNSUPPORTED, "X-engine shape not supported");
  int dev = 0;
  GXC_TRY(check_device(nullptr, &dev));
  gxc_handle tmp;
  gxc_handle *h = &tmp;
  h->cfg.nant = nant;
  h->cfg.nchan = nchan;
  h->cfg.numffts = nfft;
  h->cfg.autos = autos;
  h->out_mode = mode;
  h->d.M = M;
  h->d.nbl = nant * (nant - 1) / 2;
  h->d.nrows = h->d.nbl + (autos ? nant : 0);
  float2 *d_in = nullptr;
  struct Cleanup {
    gxc_handle *h;
    float2 **in;
    ~Cleanup() {
      if (h->stream) hipStreamSynchronize(h->stream);
      hipFree(*in);
      hipFree(h->d_spec[0]);
      hipFree(h->d_partial);
      hipFree(h->d_valid);
      hipFree(h->d_counts);
      hipFree(h->d_rowmap);
      hipFree(h->d_vis);
      if (h->stream) hipStreamDestroy(h->stream);
    }
  } cleanup{h, &d_in};
  GXC_HIP(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  const size_t nspec = (size_t)nfft * M * nchan;
  GXC_TRY(dmalloc(h, &d_in, nspec));
  GXC_HIP(h, hipMemcpy(d_in, spectra, nspec * sizeof(float2), hipMemcpyHostToDevice));
  GXC_TRY(dmalloc(h, reinterpret_cast<unsigned char **>(&h->d_spec[0]), nspec * gxc::xm_element_bytes(mode)));
  GXC_HIP(h, gxc::launch_quantise(h->d_spec[0], d_in, nspec, mode, h->stream));
  GXC_TRY(dmalloc(h, &h->d_valid, (size_t)nant * nfft));
  GXC_HIP(h, hipMemset(h->d_valid, 1, (size_t)nant * nfft));
  std::vector<int2> rm = make_rowmap(nant, autos != 0);
  GXC_TRY(dmalloc(h, &h->d_rowmap, rm.size()));
  if (!rm.empty()) GXC_HIP(h, hipMemcpy(h->d_rowmap, rm.data(), rm.size() * sizeof(int2), hipMemcpyHostToDevice));
  GXC_TRY(dmalloc(h, &h->d_counts, (size_t)h->d.nrows));
  GXC_TRY(dmalloc(h, &h->d_vis, (size_t)h->d.nrows * 4 * nchan));
  h->splits = gxc::xm_choose_splits(M, nchan, nfft, mode);
  GXC_TRY(dmalloc(h, reinterpret_cast<unsigned char **>(&h->d_partial), gxc::xm_partial_bytes(M, nchan, h->splits)));
  GXC_TRY(run_xstage(h, h->d_spec[0], nfft, 0, h->stream));
  // unit scales: the spectra were packed without scaling
  h->out_mode = gxc::OUT_F32;
  GXC_TRY(run_finalise(h, h->stream));
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  GXC_HIP(h, hipMemcpy(out, h->d_vis, (size_t)h->d.nrows * 4 * nchan * sizeof(float2), hipMemcpyDeviceToHost));
  return GXC_OK;
}

}  // extern "C"

namespace {

// Median of `reps` timings (ms) of fn() on the handle's stream, after one
// untimed run.
template <typename F>
gxc_status time_ms(gxc_handle *h, F fn, int reps, double *out) {
  hipEvent_t a = nullptr, b = nullptr;
  GXC_HIP(h, hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return set_err(h, GXC_ERR_HIP, "hipEventCreate failed");
  }
  gxc_status st = fn();
  std::vector<float> ts;
  for (int r = 0; r < reps && st == GXC_OK; r++) {
    float ms = 0.f;
    if (hipEventRecord(a, h->stream) != hipSuccess) st = set_err(h, GXC_ERR_HIP, "hipEventRecord failed");
    if (st == GXC_OK) st = fn();
    if (st == GXC_OK && (hipEventRecord(b, h->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                         hipEventElapsedTime(&ms, a, b) != hipSuccess))
      st = set_err(h, GXC_ERR_HIP, "event timing failed");
    ts.push_back(ms);
  }
  if (st == GXC_OK) st = hipStreamSynchronize(h->stream) == hipSuccess ? GXC_OK : set_err(h, GXC_ERR_HIP, "sync failed");
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (st != GXC_OK) return st;
  std::sort(ts.begin(), ts.end());
  *out = ts[ts.size() / 2];
  return GXC_OK;
}

}  // namespace

extern "C" {

gxc_status gxc_autotune(gxc_handle *h, gxc_tuning *out) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (!h->input_set || !h->delays_set) return set_err(h, GXC_ERR_STATE, "input and delays must be set");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  const bool prof = h->profiling;
  h->profiling = false;
  destroy_graph(h);
  const int graph_cfg = h->cfg.graph;
  const int nt = std::min(h->chunk, h->cfg.numffts);
  const int nchunks = (h->cfg.numffts + h->chunk - 1) / h->chunk;
  gxc_tuning t{};
  int cand = 0;
  auto process = [&]() { return enqueue_process(h); };

  // whole subintegration at the current shapes, eager (this also leaves the
  // reduced-precision scales and chunk 0's input in place for the stages)
  h->cfg.graph = 0;
  double def_ms = 0;
  gxc_status st = time_ms(h, process, 3, &def_ms);
  cand++;

  // F-stage: FFT blocks per workgroup
  double best_f = 1e30;
  int best_fb = h->fblock;
  int fb[5], nfb = gxc::fstage_blocks(h->d.L, h->cfg.is_complex != 0, fb);
  if (h->out_mode != gxc::OUT_F32 && gxc::fstage_mfma_supported(h->d.L, h->cfg.is_complex != 0))
    fb[nfb++] = gxc::kFBlockMfma;
  for (int i = 0; i < nfb && st == GXC_OK; i++) {
    h->fblock = fb[i];
    double ms = 0;
    st = time_ms(h, [&]() { return run_fstage(h, 0, nt, h->out_mode, h->d_spec[0], 1, h->stream); }, 3, &ms);
    cand++;
    if (st == GXC_OK && ms < best_f) best_f = ms, best_fb = fb[i];
  }
  if (st == GXC_OK) h->fblock = best_fb;
  // leave chunk 0's spectra in place for the X trials
  if (st == GXC_OK) st = run_fstage(h, 0, nt, h->out_mode, h->d_spec[0], 1, h->stream);

  // X-stage: channel tile x FFT splits
  double best_x = 1e30;
  int best_xt = h->xtile, best_sp = h->splits;
  if (st == GXC_OK) {
    const int tb = h->out_mode == gxc::OUT_F32 ? 16 : 32;
    const int s0 = gxc::xm_choose_splits(h->d.M, h->cfg.nchan, nt, h->out_mode);
    const int smax = std::max(1, std::min(1024, nt / (2 * tb)));
    std::vector<int> sps;
    for (int f : {1, 2, 4})
      for (int v : {s0 / f, s0 * f}) {
        const int c = std::max(1, std::min(v, smax));
        if (std::find(sps.begin(), sps.end(), c) == sps.end()) sps.push_back(c);
      }
    const int need = *std::max_element(sps.begin(), sps.end());
    if (need > h->partial_cap) {
      float *p = nullptr;
      if (hipMalloc((void **)&p, gxc::xm_partial_bytes(h->d.M, h->cfg.nchan, need)) == hipSuccess) {
        hipFree(h->d_partial);
        h->d_partial = p;
        h->partial_cap = need;
      } else {
        (void)hipGetLastError();
        sps.erase(std::remove_if(sps.begin(), sps.end(), [&](int v) { return v > h->partial_cap; }), sps.end());
      }
    }
    int xt[3], nxt = gxc::xm_tiles(h->d.M, h->out_mode, xt);
    for (int i = 0; i < nxt && st == GXC_OK; i++) {
      if (h->cfg.nchan % xt[i]) continue;
      h->xtile = xt[i];
      for (int sp : sps) {
        h->splits = sp;
        double ms = 0;
        st = time_ms(h, [&]() { return run_xstage(h, h->d_spec[0], nt, 0, h->stream); }, 3, &ms);
        cand++;
        if (st != GXC_OK) break;
        if (ms < best_x) best_x = ms, best_xt = xt[i], best_sp = sp;
      }
    }
    h->xtile = best_xt;
    h->splits = best_sp;
  }

  // pipeline: eager launches or graph replay
  double eager_ms = 0, graph_ms = 1e30;
  if (st == GXC_OK) {
    h->cfg.graph = 0;
    st = time_ms(h, process, 3, &eager_ms);
    cand++;
  }
  if (st == GXC_OK) {
    h->cfg.graph = 1;
    destroy_graph(h);
    st = time_ms(h, process, 3, &graph_ms);  // the untimed first run captures
    cand++;
  }
  h->cfg.graph = st == GXC_OK ? (graph_ms <= eager_ms ? 1 : 0) : graph_cfg;
  destroy_graph(h);
  h->profiling = prof;
  if (st != GXC_OK) return st;
  t.fblock = h->fblock ? h->fblock : fb[0];
  t.xtile = h->xtile ? h->xtile : gxc::xm_channels_per_group(h->d.M, h->out_mode);
  t.xsplits = h->splits;
  t.graph = h->cfg.graph;
  t.fstage_ms = best_f * nchunks;
  t.xstage_ms = best_x * nchunks;
  t.process_ms = std::min(eager_ms, graph_ms);
  t.default_ms = def_ms;
  t.candidates = cand;
  if (out) *out = t;
  return GXC_OK;
}

void *gxc_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void gxc_host_free(void *p) {
  if (p) hipHostFree(p);
}

gxc_status gxc_integrate_begin(gxc_handle *h) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  if (h->vdif) {  // submitted input has no frame flags
    leave_vdif(h);
    h->input_set = false;
  }
  if (!h->cstream) GXC_HIP(h, hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking));
  const size_t nvis = (size_t)h->d.nrows * 4 * h->cfg.nchan;
  if (!h->d_acc) GXC_TRY(dmalloc(h, &h->d_acc, nvis));
  for (auto &sl : h->slot) {
    if (sl.d_in) continue;
    const int nant = h->cfg.nant;
    GXC_TRY(dmalloc(h, &sl.d_in, h->d.station_bytes * (size_t)nant));
    GXC_TRY(dmalloc(h, &sl.d_ptrs, (size_t)nant));
    GXC_TRY(dmalloc(h, &sl.d_delays, (size_t)nant * 4));
    GXC_HIP(h, hipHostMalloc((void **)&sl.h_delays, (size_t)nant * 4 * sizeof(double), hipHostMallocDefault));
    GXC_HIP(h, hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    GXC_HIP(h, hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    std::vector<const uint8_t *> ptrs(nant);
    for (int a = 0; a < nant; a++) ptrs[a] = sl.d_in + (size_t)a * h->d.station_bytes;
    GXC_HIP(h, hipMemcpy(sl.d_ptrs, ptrs.data(), nant * sizeof(void *), hipMemcpyHostToDevice));
  }
  GXC_HIP(h, gxc::launch_fill(reinterpret_cast<float *>(h->d_acc), 2 * nvis, 0.f, h->stream));
  for (auto &sl : h->slot) sl.used = false;
  h->nsubmitted = 0;
  h->integrating = true;
  return GXC_OK;
}

gxc_status gxc_submit(gxc_handle *h, const uint8_t *const *per_station, const double *const *poly3,
                      const double *fileoffsets) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (!h->integrating) return set_err(h, GXC_ERR_STATE, "gxc_integrate_begin first");
  if (!per_station || !poly3 || !fileoffsets) return set_err(h, GXC_ERR_INVALID_ARG, "null input or delays");
  const int nant = h->cfg.nant;
  for (int a = 0; a < nant; a++)
    if (!per_station[a] || !poly3[a]) return set_err(h, GXC_ERR_INVALID_ARG, "null station buffer or polynomial");
  auto &sl = h->slot[h->nsubmitted & 1];
  h->marks.clear();
  // the slot's previous copy (subintegration k-2, already waited on by the
  // previous call) has finished reading the delay staging before it is rewritten
  if (sl.used) GXC_HIP(h, hipEventSynchronize(sl.copied));
  for (int a = 0; a < nant; a++) {
    for (int i = 0; i < 3; i++) sl.h_delays[4 * a + i] = poly3[a][i];
    sl.h_delays[4 * a + 3] = fileoffsets[a];
  }
  // copy stream: after subintegration k-2 released the slot
  if (sl.used) GXC_HIP(h, hipStreamWaitEvent(h->cstream, sl.done, 0));
  for (int a = 0; a < nant; a++)
    GXC_HIP(h, hipMemcpyAsync(sl.d_in + (size_t)a * h->d.station_bytes, per_station[a], h->d.station_bytes,
                              hipMemcpyHostToDevice, h->cstream));
  GXC_HIP(h, hipMemcpyAsync(sl.d_delays, sl.h_delays, (size_t)nant * 4 * sizeof(double), hipMemcpyHostToDevice,
                            h->cstream));
  GXC_HIP(h, hipEventRecord(sl.copied, h->cstream));
  // compute stream: correlate from the slot, then add into the sum
  GXC_HIP(h, hipStreamWaitEvent(h->stream, sl.copied, 0));
  const uint8_t **ptrs = h->d_station_ptrs;
  double *dl = h->d_delays;
  h->d_station_ptrs = sl.d_ptrs;
  h->d_delays = sl.d_delays;
  gxc_status st = enqueue_device_work(h);
  h->d_station_ptrs = ptrs;
  h->d_delays = dl;
  GXC_TRY(st);
  const size_t nvis = (size_t)h->d.nrows * 4 * h->cfg.nchan;
  GXC_HIP(h, gxc::launch_add(reinterpret_cast<float *>(h->d_acc), reinterpret_cast<const float *>(h->d_vis),
                             2 * nvis, h->stream));
  GXC_HIP(h, hipEventRecord(sl.done, h->stream));
  sl.used = true;
  // the previous subintegration's host buffers have been read once this
  // returns, so a caller may refill them while this one copies
  if (h->nsubmitted > 0) GXC_HIP(h, hipEventSynchronize(h->slot[(h->nsubmitted - 1) & 1].copied));
  h->nsubmitted++;
  return GXC_OK;
}

gxc_status gxc_integrate_end(gxc_handle *h, gxc_cf32 *out, size_t nelem, int *nsubints) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (!h->integrating) return set_err(h, GXC_ERR_STATE, "no integration in progress");
  const size_t nvis = (size_t)h->d.nrows * 4 * h->cfg.nchan;
  if (out && nelem < nvis) return set_err(h, GXC_ERR_INVALID_ARG, "output buffer too small");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  if (out) GXC_HIP(h, hipMemcpy(out, h->d_acc, nvis * sizeof(gxc_cf32), hipMemcpyDeviceToHost));
  if (nsubints) *nsubints = (int)h->nsubmitted;
  h->integrating = false;
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
