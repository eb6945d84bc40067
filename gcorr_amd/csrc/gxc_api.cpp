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
#include "gxc_handle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

using gxc::Geometry;
using gxc::host::Derived;
using gxc::host::derive;
using gxc::host::make_rowmap;
using gxc::host::out_mode_of;

namespace {

thread_local std::string g_thread_error;

}  // namespace

namespace gxc {
namespace api {

gxc_status set_err(gxc_handle *h, gxc_status s, const std::string &msg) {
  if (h) h->err = msg;
  g_thread_error = msg;
  return s;
}

const std::string &thread_error() { return g_thread_error; }

gxc_status check_device(gxc_handle *h, int *dev) {
  std::string why;
  const gxc_status st = gxc::host::check_device(dev, &why);
  return st == GXC_OK ? GXC_OK : set_err(h, st, why);
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
  p.frame_ok = h->vdif ? h->vs.d_flag_ptrs : nullptr;
  p.frame_samples = h->vs.d_spf;
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
  x.comp = h->d_comp;
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
  f.comp = h->d_comp;
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
}  // namespace api
}  // namespace gxc

using namespace gxc::api;

extern "C" {

void gxc_config_init(gxc_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->nbit = 2;  // common.cpp:64-66 defaults
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
  // the whole subint in one pass without overlap (bounded by 32 GiB of the
  // 288 GB of HBM; one pass saves the launch tails of the others: 16
  // stations x 4096 channels, FP16, 15.80 -> 15.34 ms), eight passes with
  // overlap (bounded by 4 GiB per buffer).  Either budget is also held to
  // half of the device memory free at creation, and an automatic chunk that
  // still does not fit is halved until it does.
  const size_t eb = gxc::xm_element_bytes(out_mode);
  const size_t per_fft = (size_t)d.M * nchan * eb;
  const bool auto_chunk = cfg->chunk_ffts <= 0;
  long long chunk = cfg->chunk_ffts;
  if (auto_chunk) {
    size_t budget = cfg->overlap ? (4ull << 30) : (32ull << 30);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
      budget = std::min(budget, free_b / 2 / (cfg->overlap ? 2 : 1));
    chunk = cfg->overlap ? (cfg->numffts + 7) / 8 : cfg->numffts;
    chunk = std::min<long long>(chunk, (long long)(budget / per_fft));
  }
  chunk = std::max<long long>(1, std::min<long long>(chunk, cfg->numffts));
  for (;;) {
    h->chunk = (int)chunk;
    h->nbuf = (cfg->overlap && chunk < cfg->numffts) ? 2 : 1;  // one pass: nothing to overlap
    int b = 0;
    for (; b < h->nbuf; b++)
      if ((s = dmalloc(h, reinterpret_cast<unsigned char **>(&h->d_spec[b]), (size_t)h->chunk * per_fft)) !=
          GXC_OK)
        break;
    if (b == h->nbuf) break;
    if (!auto_chunk || chunk == 1) return fail(s);
    for (int i = 0; i < b; i++) {
      hipFree(h->d_spec[i]);
      h->d_spec[i] = nullptr;
    }
    (void)hipGetLastError();  // the failed allocation's error is handled here
    chunk = (chunk + 1) / 2;
  }
  // launch shapes: given, or defaults
  {
    int fb[4], nfb = gxc::fstage_blocks(d.L, cfg->is_complex != 0, fb);
    bool ok = cfg->fblock == 0 ||
              (cfg->fblock == gxc::kFBlockMfma &&
               gxc::fstage_mfma_supported(d.L, cfg->is_complex != 0, out_mode == gxc::OUT_F32));
    for (int i = 0; i < nfb; i++) ok = ok || cfg->fblock == fb[i];
    if (!ok) return fail(set_err(h, GXC_ERR_INVALID_ARG, "fblock is not an F-stage workgroup size for this FFT length"));
    int xt[4], nxt = gxc::xm_tiles(d.M, out_mode, xt);
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
  // Several chunks per subintegration: later chunks add into the partials
  // with Neumaier compensation, kept in a buffer of the same shape.
  if (h->chunk < cfg->numffts &&
      (s = dmalloc(h, reinterpret_cast<unsigned char **>(&h->d_comp),
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
  h->vs.release();
  hipFree(h->d_polgains);
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
  hipFree(h->d_comp);
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
