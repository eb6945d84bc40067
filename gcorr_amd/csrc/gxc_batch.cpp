// Multi-subband batch (include/gxcorr.h, gxc_batch_*): every subband of a
// subintegration through one persistent launch per stage.
//
// Device work of one subintegration:
//   [H2D copies of host input, per subband]
//   [reduced precision: power pass over 64 sampled FFTs of every subband
//    (one launch), power-of-two scales (one launch)]
//   for each chunk c (1/chunks of every subband's FFTs):
//     F-stage of all subbands (one launch) -> spectra buffer c % 2
//     X-stage of all subbands (one launch), accumulating split partials
//   row counts (one launch), finalise (one launch)
// The launches read segment tables built once at gxc_batch_create: one
// entry per subband with its parameters and its first unit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gxcorr.h"
#include "gxc_host.hpp"
#include "gxc_internal.hpp"
#include "gxc_vdif_state.hpp"
#include "gxc_xengine.hpp"

namespace {

thread_local std::string g_batch_error;

struct Sub {
  gxc_config cfg{};
  gxc::host::Derived d;
  int ntc = 0;     // FFTs of the largest chunk
  int splits = 1;  // X-stage FFT splits
  int partial_cap = 0;  // splits the partial buffer holds
  int npow = 0;    // FFTs sampled for the scales
  // input
  bool input_set = false, input_device = false;
  std::vector<const uint8_t *> host_in;
  uint8_t *d_input_owned = nullptr;
  const uint8_t **d_station_ptrs = nullptr;
  // delays
  bool delays_set = false;
  std::vector<double> h_delays;
  double *d_delays = nullptr;
  // tables and outputs
  float2 *d_twiddle = nullptr;
  float *d_subfreq = nullptr, *d_stepfreq = nullptr;
  uint8_t *d_valid = nullptr;
  int *d_counts = nullptr;
  float2 *d_vis = nullptr;
  float *d_partial = nullptr;
  float *d_comp = nullptr;  // compensation of d_partial across chunks (several chunks only)
  // polarisation conversion of the visibilities (gxc_batch_set_pol_conversion)
  bool polconv = false;
  float2 *d_polgains = nullptr;  // [nant][2][nchan]
  // VDIF ingest (gxc_batch_set_input_vdif): payloads go into d_input_owned,
  // and the frame flags make FFTs over missing frames invalid
  bool vdif = false;
  gxc::host::VdifState vs;
  float *d_scales = nullptr, *d_power = nullptr;
  size_t spec_off = 0;  // byte offset of this subband in a spectra buffer
};

}  // namespace

struct gxc_batch {
  gxc_batch_config cfg{};
  std::vector<Sub> sub;
  std::string err;
  int device = 0, out_mode = 0, M = 0, nrows = 0, nchunks = 1, nbuf = 1, fblock = 256;
  int maxL = 0, xunits = 2048;  // longest FFT; X units per launch the splits aim at
  hipStream_t stream = nullptr, xstream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_f[2] = {nullptr, nullptr}, ev_x[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec = nullptr;
  int2 *d_rowmap = nullptr;
  void *d_spec[2] = {nullptr, nullptr};
  // segment tables and unit totals
  gxc::FSeg *d_fseg = nullptr;  // [nchunks][nmulti]: subbands of the multi-subband F launch
  int nmulti = 0;
  std::vector<std::vector<gxc::FSeg>> mseg;  // [nchunks]: subbands on the matrix-core F-stage
  gxc::XSeg *d_xseg = nullptr;  // [nchunks][nsub]
  gxc::FSeg *d_powseg = nullptr;
  gxc::ScaleSeg *d_scaleseg = nullptr;
  gxc::CountSeg *d_countseg = nullptr;
  gxc::FinSeg *d_finseg = nullptr;
  std::vector<long long> ftotal, xtotal;
  long long powtotal = 0, scaletotal = 0, counttotal = 0, fintotal = 0;
  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> tev;
  std::vector<std::pair<int, size_t>> marks;
  gxc_timings timings{};
};

namespace {

gxc_status set_err(gxc_batch *b, gxc_status s, const std::string &msg) {
  if (b) b->err = msg;
  g_batch_error = msg;
  return s;
}

#define GXB_HIP(b, call)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_err((b), e_ == hipErrorOutOfMemory ? GXC_ERR_OUT_OF_MEMORY : GXC_ERR_HIP,   \
                     std::string(#call) + ": " + hipGetErrorString(e_));                     \
  } while (0)

#define GXB_TRY(expr)            \
  do {                           \
    gxc_status s_ = (expr);      \
    if (s_ != GXC_OK) return s_; \
  } while (0)

template <typename T>
gxc_status dmalloc(gxc_batch *b, T **p, size_t nelem) {
  *p = nullptr;
  if (nelem == 0) nelem = 1;
  hipError_t e = hipMalloc((void **)p, nelem * sizeof(T));
  if (e != hipSuccess)
    return set_err(b, GXC_ERR_OUT_OF_MEMORY,
                   "hipMalloc of " + std::to_string(nelem * sizeof(T)) + " bytes failed: " +
                       hipGetErrorString(e));
  return GXC_OK;
}

template <typename T>
gxc_status upload(gxc_batch *b, T **dst, const std::vector<T> &v) {
  GXB_TRY(dmalloc(b, dst, v.size()));
  if (!v.empty()) GXB_HIP(b, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return GXC_OK;
}

void mark(gxc_batch *b, int kind, hipStream_t s) {
  if (!b->profiling) return;
  const size_t idx = b->marks.size();
  while (b->tev.size() <= idx) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    b->tev.push_back(e);
  }
  hipEventRecord(b->tev[idx], s);
  b->marks.push_back({kind, idx});
}

// F-stage parameters of subband s, FFTs t0.. (stride tstride), output mode.
gxc::FParams fparams(const gxc_batch *b, const Sub &u, int t0, int nt, int tstride, int mode, void *out) {
  gxc::FParams p{};
  p.station = u.d_station_ptrs;
  p.delays = u.d_delays;
  p.subchanfreq = u.d_subfreq;
  p.stepchanfreq = u.d_stepfreq;
  p.twiddle = u.d_twiddle;
  p.spectra = static_cast<float2 *>(out);
  p.out = out;
  p.scales = u.d_scales;
  p.power = u.d_power;
  p.valid = u.d_valid;
  p.out_mode = mode;
  p.t0 = t0;
  p.nt = nt;
  p.tstride = tstride;
  p.nant = b->cfg.nant;
  p.numffts = u.cfg.numffts;
  p.maxoffset = u.d.maxoffset;
  p.stepmul = b->cfg.exact_phase ? u.d.g.S2 : u.d.g.s;
  const double lo = u.cfg.lo;
  p.lo = lo;
  p.lofrac = lo - std::trunc(lo);
  p.fraclo = (lo - std::trunc(lo)) > 0.00000000001 ? 1 : 0;  // TINY, fxkernel.cpp:66-71
  p.sampletime = u.d.sampletime;
  p.frame_ok = u.vdif ? u.vs.d_flag_ptrs : nullptr;
  p.frame_samples = u.vs.d_spf;
  return p;
}

int chunk_begin(const Sub &u, int c, int nchunks) {
  return (int)((long long)u.cfg.numffts * c / nchunks);
}

// Segment tables of every launch, from the handle's buffers (fixed for the
// handle's lifetime: inputs and delays are reached through device arrays).
gxc_status build_tables(gxc_batch *b) {
  for (void *p : {(void *)b->d_fseg, (void *)b->d_xseg, (void *)b->d_countseg, (void *)b->d_finseg,
                  (void *)b->d_powseg, (void *)b->d_scaleseg})
    hipFree(p);
  b->d_fseg = nullptr;
  b->d_xseg = nullptr;
  b->d_countseg = nullptr;
  b->d_finseg = nullptr;
  b->d_powseg = nullptr;
  b->d_scaleseg = nullptr;
  const int ns = (int)b->sub.size(), C = b->nchunks;
  const int ct = gxc::xm_channels_per_group(b->M, b->out_mode);
  // with fstage_mfma, the subbands of 2048 points and more get launches of
  // their own: on the matrix cores where those cover them, else on the
  // register kernel (fstage_mfma 1), or all on the register kernel (2)
  auto on_mfma = [&](const Sub &u) { return b->cfg.fstage_mfma != 0 && u.d.L >= 2048; };
  b->nmulti = 0;
  for (const Sub &u : b->sub) b->nmulti += on_mfma(u) ? 0 : 1;
  b->mseg.assign(C, {});
  std::vector<gxc::FSeg> fseg((size_t)C * std::max(1, b->nmulti));
  std::vector<gxc::XSeg> xseg((size_t)C * ns);
  b->ftotal.assign(C, 0);
  b->xtotal.assign(C, 0);
  for (int c = 0; c < C; c++) {
    const int buf = b->nbuf == 2 ? (c & 1) : 0;
    long long fu = 0, xu = 0;
    int m = 0;
    for (int s = 0; s < ns; s++) {
      const Sub &u = b->sub[s];
      const int t0 = chunk_begin(u, c, C), nt = chunk_begin(u, c + 1, C) - t0;
      void *spec = static_cast<unsigned char *>(b->d_spec[buf]) + u.spec_off;
      gxc::FSeg f;
      f.p = fparams(b, u, t0, nt, 1, b->out_mode, spec);
      f.L = u.d.L;
      if (on_mfma(u)) {
        f.unit0 = 0;
        b->mseg[c].push_back(f);
      } else {
        f.unit0 = fu;
        fu += gxc::fstage_multi_units(u.d.L, b->fblock, (long long)b->cfg.nant * nt);
        fseg[(size_t)c * b->nmulti + m++] = f;
      }
      gxc::XSeg &x = xseg[(size_t)c * ns + s];
      x.p.spectra = spec;
      x.p.partial = u.d_partial;
      x.p.comp = u.d_comp;
      x.p.nt = nt;
      x.p.M = b->M;
      x.p.nchan = u.cfg.nchan;
      x.p.splits = u.splits;
      x.p.accumulate = c > 0;
      x.unit0 = xu;
      xu += (long long)(u.cfg.nchan / ct) * u.splits;
    }
    b->ftotal[c] = fu;
    b->xtotal[c] = xu;
  }
  GXB_TRY(upload(b, &b->d_fseg, fseg));
  GXB_TRY(upload(b, &b->d_xseg, xseg));

  std::vector<gxc::CountSeg> cseg(ns);
  std::vector<gxc::FinSeg> finseg(ns);
  long long cu = 0, fin = 0;
  for (int s = 0; s < ns; s++) {
    const Sub &u = b->sub[s];
    cseg[s].p.valid = u.d_valid;
    cseg[s].p.rowmap = b->d_rowmap;
    cseg[s].p.counts = u.d_counts;
    cseg[s].p.numffts = u.cfg.numffts;
    cseg[s].p.nrows = b->nrows;
    cseg[s].unit0 = cu;
    cu += b->nrows;
    gxc::FinMParams &f = finseg[s].p;
    f.partial = u.d_partial;
    f.comp = u.d_comp;
    f.scales = b->out_mode == gxc::OUT_F32 ? nullptr : u.d_scales;
    f.rowmap = b->d_rowmap;
    f.counts = u.d_counts;
    f.vis = u.d_vis;
    f.M = b->M;
    f.nchan = u.cfg.nchan;
    f.nrows = b->nrows;
    f.splits = u.splits;
    finseg[s].unit0 = fin;
    fin += u.cfg.nchan;
  }
  b->counttotal = cu;
  b->fintotal = fin;
  GXB_TRY(upload(b, &b->d_countseg, cseg));
  GXB_TRY(upload(b, &b->d_finseg, finseg));

  if (b->out_mode != gxc::OUT_F32) {
    std::vector<gxc::FSeg> pseg(ns);
    std::vector<gxc::ScaleSeg> sseg(ns);
    long long pu = 0, su = 0;
    for (int s = 0; s < ns; s++) {
      const Sub &u = b->sub[s];
      pseg[s].p = fparams(b, u, 0, u.npow, u.cfg.numffts / u.npow, gxc::OUT_POWER, nullptr);
      pseg[s].L = u.d.L;
      pseg[s].unit0 = pu;
      pu += gxc::fstage_multi_units(u.d.L, b->fblock, (long long)b->cfg.nant * u.npow);
      sseg[s].scales = u.d_scales;
      sseg[s].power = u.d_power;
      sseg[s].n = b->M * u.cfg.nchan;
      sseg[s].nrows = u.npow;
      sseg[s].target = gxc::xm_target_rms(b->out_mode);
      sseg[s].unit0 = su;
      su += (sseg[s].n + 255) / 256;
    }
    b->powtotal = pu;
    b->scaletotal = su;
    GXB_TRY(upload(b, &b->d_powseg, pseg));
    GXB_TRY(upload(b, &b->d_scaleseg, sseg));
  }
  return GXC_OK;
}

// X splits per subband so that a launch has about b->xunits units, each
// covering a similar number of FFTs (a unit's X work is channels x FFTs x
// M^2); partial buffers grow when needed.
gxc_status assign_splits(gxc_batch *b) {
  const int ct = gxc::xm_channels_per_group(b->M, b->out_mode);
  const int tb = b->out_mode == gxc::OUT_F32 ? 16 : 32;
  double work = 0;
  for (const Sub &u : b->sub) work += (double)(u.cfg.nchan / ct) * u.ntc;
  const double per_unit = std::max(1.0, work / b->xunits);
  for (Sub &u : b->sub) {
    const long long sp = std::llround(u.ntc / per_unit);
    const long long maxs = std::max(1, u.ntc / (4 * tb));
    u.splits = (int)std::max(1LL, std::min(std::min(sp, maxs), 256LL));
    if (u.splits > u.partial_cap) {
      hipFree(u.d_partial);
      hipFree(u.d_comp);
      u.d_partial = u.d_comp = nullptr;
      u.partial_cap = 0;
      const size_t nb = gxc::xm_partial_bytes(b->M, u.cfg.nchan, u.splits);
      GXB_TRY(dmalloc(b, reinterpret_cast<unsigned char **>(&u.d_partial), nb));
      // several chunks: later chunks add with Neumaier compensation
      if (b->nchunks > 1) GXB_TRY(dmalloc(b, reinterpret_cast<unsigned char **>(&u.d_comp), nb));
      u.partial_cap = u.splits;
    }
  }
  return GXC_OK;
}

gxc_status enqueue_device_work(gxc_batch *b) {
  const int ns = (int)b->sub.size();
  const bool overlap = b->nbuf == 2;
  hipStream_t fs = b->stream, xs = overlap ? b->xstream : b->stream;
  if (b->out_mode != gxc::OUT_F32) {
    GXB_HIP(b, gxc::launch_fstage_multi(b->d_powseg, ns, b->powtotal, gxc::OUT_POWER, b->fblock, fs));
    GXB_HIP(b, gxc::launch_scales_multi(b->d_scaleseg, ns, b->scaletotal, fs));
  }
  if (overlap) {
    GXB_HIP(b, hipEventRecord(b->ev_fork, fs));
    GXB_HIP(b, hipStreamWaitEvent(xs, b->ev_fork, 0));
  }
  for (int c = 0; c < b->nchunks; c++) {
    const int buf = overlap ? (c & 1) : 0;
    if (overlap && c >= 2) GXB_HIP(b, hipStreamWaitEvent(fs, b->ev_x[buf], 0));
    mark(b, 0, fs);
    if (b->nmulti > 0 && b->ftotal[c] > 0)
      GXB_HIP(b, gxc::launch_fstage_multi(b->d_fseg + (size_t)c * b->nmulti, b->nmulti, b->ftotal[c], b->out_mode,
                                          b->fblock, fs));
    for (const gxc::FSeg &f : b->mseg[c])
      GXB_HIP(b, b->cfg.fstage_mfma == 1 &&
                         gxc::fstage_mfma_supported(f.L, b->cfg.is_complex != 0, b->out_mode == gxc::OUT_F32)
                     ? gxc::launch_fstage_mfma(f.p, f.L, fs)
                     : gxc::launch_fstage_big(f.p, f.L, b->cfg.is_complex != 0, fs));
    mark(b, 1, fs);
    if (overlap) {
      GXB_HIP(b, hipEventRecord(b->ev_f[buf], fs));
      GXB_HIP(b, hipStreamWaitEvent(xs, b->ev_f[buf], 0));
    }
    mark(b, 2, xs);
    GXB_HIP(b, gxc::launch_xmfma_multi(b->d_xseg + (size_t)c * ns, ns, b->xtotal[c], b->out_mode, b->M, xs));
    mark(b, 3, xs);
    if (overlap) GXB_HIP(b, hipEventRecord(b->ev_x[buf], xs));
  }
  mark(b, 4, xs);
  GXB_HIP(b, gxc::launch_counts_multi(b->d_countseg, ns, b->counttotal, xs));
  GXB_HIP(b, gxc::launch_finalise_multi(b->d_finseg, ns, b->fintotal, b->M, xs));
  for (const Sub &u : b->sub)
    if (u.polconv)
      GXB_HIP(b, gxc::launch_polconv(u.d_vis, u.d_polgains, b->d_rowmap, b->nrows, u.cfg.nchan, xs));
  mark(b, 5, xs);
  if (overlap) {
    GXB_HIP(b, hipEventRecord(b->ev_join, xs));
    GXB_HIP(b, hipStreamWaitEvent(fs, b->ev_join, 0));
  }
  b->timings.fstage_launches = b->nchunks;
  b->timings.xstage_launches = b->nchunks;
  return GXC_OK;
}

gxc_status ensure_owned_input(gxc_batch *b, Sub &u) {
  if (u.d_input_owned) return GXC_OK;
  const int nant = b->cfg.nant;
  GXB_TRY(dmalloc(b, &u.d_input_owned, u.d.station_bytes * (size_t)nant));
  return GXC_OK;
}

// The segment tables hold each subband's frame-flag pointers: rebuild them,
// and drop a captured graph, when a subband enters or leaves VDIF mode.
gxc_status vdif_mode(gxc_batch *b, Sub &u, bool on) {
  if (u.vdif == on) return GXC_OK;
  u.vdif = on;
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  b->graph_exec = nullptr;
  return build_tables(b);
}

gxc_status leave_vdif(gxc_batch *b, Sub &u) { return vdif_mode(b, u, false); }

gxc_status point_at_owned(gxc_batch *b, Sub &u) {
  std::vector<const uint8_t *> ptrs(b->cfg.nant);
  for (int a = 0; a < b->cfg.nant; a++) ptrs[a] = u.d_input_owned + (size_t)a * u.d.station_bytes;
  GXB_HIP(b, hipMemcpy(u.d_station_ptrs, ptrs.data(), ptrs.size() * sizeof(void *), hipMemcpyHostToDevice));
  return GXC_OK;
}

gxc_status enqueue_process(gxc_batch *b) {
  for (size_t s = 0; s < b->sub.size(); s++) {
    if (!b->sub[s].input_set) return set_err(b, GXC_ERR_STATE, "input not set for subband " + std::to_string(s));
    if (!b->sub[s].delays_set)
      return set_err(b, GXC_ERR_STATE, "delays not set for subband " + std::to_string(s));
  }
  b->marks.clear();
  mark(b, 7, b->stream);
  for (Sub &u : b->sub) {
    if (u.input_device || u.vdif) continue;
    for (int a = 0; a < b->cfg.nant; a++)
      GXB_HIP(b, hipMemcpyAsync(u.d_input_owned + (size_t)a * u.d.station_bytes, u.host_in[a], u.d.station_bytes,
                                hipMemcpyHostToDevice, b->stream));
  }
  mark(b, 6, b->stream);
  if (!b->cfg.graph || b->profiling) return enqueue_device_work(b);
  if (!b->graph_exec) {
    hipGraph_t graph = nullptr;
    GXB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    gxc_status s = enqueue_device_work(b);
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (s != GXC_OK) {
      if (graph) hipGraphDestroy(graph);
      return s;
    }
    if (e != hipSuccess) return set_err(b, GXC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      b->graph_exec = nullptr;
      return set_err(b, GXC_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
  }
  GXB_HIP(b, hipGraphLaunch(b->graph_exec, b->stream));
  return GXC_OK;
}

float elapsed(gxc_batch *b, size_t x, size_t y) {
  float ms = 0.f;
  if (x < b->tev.size() && y < b->tev.size() && hipEventElapsedTime(&ms, b->tev[x], b->tev[y]) == hipSuccess)
    return ms;
  return 0.f;
}

void collect_timings(gxc_batch *b) {
  if (!b->profiling || b->marks.empty()) return;
  gxc_timings &t = b->timings;
  t.h2d_ms = t.fstage_ms = t.xstage_ms = t.finalise_ms = t.total_ms = 0;
  size_t open[8] = {0};
  for (const auto &m : b->marks) {
    switch (m.first) {
      case 7: open[7] = m.second; break;
      case 6: t.h2d_ms += elapsed(b, open[7], m.second); break;
      case 0: open[0] = m.second; break;
      case 1: t.fstage_ms += elapsed(b, open[0], m.second); break;
      case 2: open[2] = m.second; break;
      case 3: t.xstage_ms += elapsed(b, open[2], m.second); break;
      case 4: open[4] = m.second; break;
      case 5: t.finalise_ms += elapsed(b, open[4], m.second); break;
    }
  }
  t.total_ms = elapsed(b, b->marks.front().second, b->marks.back().second);
}

Sub *get_sub(gxc_batch *b, int sub) {
  if (!b || sub < 0 || sub >= (int)b->sub.size()) return nullptr;
  return &b->sub[sub];
}

}  // namespace

extern "C" {

void gxc_batch_config_init(gxc_batch_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->nbit = 2;
  cfg->npol = 2;
  cfg->precision = GXC_PREC_FP32;
  cfg->overlap = 0;
  cfg->graph = 1;
}

const char *gxc_batch_last_error(const gxc_batch *b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

gxc_status gxc_batch_create(const gxc_batch_config *cfg, gxc_batch **out) {
  if (!out) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null output handle pointer");
  *out = nullptr;
  if (!cfg || cfg->nsub < 1 || !cfg->subbands) return set_err(nullptr, GXC_ERR_INVALID_ARG, "no subbands");
  if (cfg->nsub > 1024) return set_err(nullptr, GXC_ERR_UNSUPPORTED, "at most 1024 subbands");
  if (cfg->nbit != 2 || cfg->is_complex)
    return set_err(nullptr, GXC_ERR_UNSUPPORTED, "batches take real 2-bit input");
  if (cfg->fstage_mfma < 0 || cfg->fstage_mfma > 2)
    return set_err(nullptr, GXC_ERR_INVALID_ARG, "fstage_mfma is 0, 1 or 2");
  std::vector<Sub> subs(cfg->nsub);
  int maxL = 0, minfft = 1 << 30;
  const int out_mode = gxc::host::out_mode_of(cfg->precision < 0 || cfg->precision > 3 ? 0 : cfg->precision);
  for (int s = 0; s < cfg->nsub; s++) {
    gxc_config c;
    gxc_config_init(&c);
    c.nant = cfg->nant;
    c.nchan = cfg->subbands[s].nchan;
    c.numffts = cfg->subbands[s].numffts;
    c.nbit = cfg->nbit;
    c.npol = cfg->npol;
    c.is_complex = cfg->is_complex;
    c.lo = cfg->subbands[s].lo;
    c.bw = cfg->subbands[s].bw;
    c.precision = cfg->precision;
    c.autos = cfg->autos;
    c.exact_phase = cfg->exact_phase;
    std::string why;
    gxc_status st = gxc::host::derive(&c, &subs[s].d, &why);
    if (st != GXC_OK) return set_err(nullptr, st, "subband " + std::to_string(s) + ": " + why);
    if (!gxc::fstage_multi_supported(subs[s].d.L, false))
      return set_err(nullptr, GXC_ERR_UNSUPPORTED,
                     "subband " + std::to_string(s) + ": FFT length must be 256..8192 in a batch");
    const int ct = gxc::xm_channels_per_group(subs[s].d.M, out_mode);
    if (ct == 0 || c.nchan % ct)
      return set_err(nullptr, GXC_ERR_UNSUPPORTED,
                     "subband " + std::to_string(s) + ": channel count not a multiple of the X tile");
    subs[s].cfg = c;
    maxL = std::max(maxL, subs[s].d.L);
    minfft = std::min(minfft, c.numffts);
  }
  int dev = 0;
  {
    std::string why;
    gxc_status st = gxc::host::check_device(&dev, &why);
    if (st != GXC_OK) return set_err(nullptr, st, why);
  }
  gxc_batch *b = new gxc_batch();
  b->cfg = *cfg;
  b->cfg.subbands = nullptr;
  b->sub = std::move(subs);
  b->device = dev;
  b->out_mode = out_mode;
  b->M = 2 * cfg->nant;
  b->nrows = b->sub[0].d.nrows;
  b->fblock = gxc::fstage_multi_block(maxL);
  auto fail = [&](gxc_status st) {
    g_batch_error = b->err;
    gxc_batch_destroy(b);
    return st;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->xstream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(b, GXC_ERR_HIP, "hipStreamCreate failed"));
  hipEvent_t *evs[6] = {&b->ev_fork, &b->ev_join, &b->ev_f[0], &b->ev_f[1], &b->ev_x[0], &b->ev_x[1]};
  for (hipEvent_t *e : evs)
    if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
      return fail(set_err(b, GXC_ERR_HIP, "hipEventCreate failed"));

  // chunks: automatic = 8 passes with overlap, else as few as keep one
  // spectra buffer within 32 GiB and within half of the device memory free
  // now; at most the smallest subband's FFT count, so every chunk has FFTs
  // of every subband
  int C = cfg->chunks;
  if (C <= 0) {
    double total = 0;
    for (const Sub &u : b->sub)
      total += (double)u.cfg.numffts * 2 * cfg->nant * u.cfg.nchan * gxc::xm_element_bytes(out_mode);
    double budget = 32.0 * (1ull << 30);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
      budget = std::min(budget, (double)free_b / 2 / (cfg->overlap ? 2 : 1));
    C = std::max(cfg->overlap ? 8 : 1, (int)std::ceil(total / budget));
  }
  C = std::max(1, std::min(C, minfft));
  b->nchunks = C;
  b->nbuf = (cfg->overlap && C > 1) ? 2 : 1;

  const int nant = cfg->nant, M = b->M;
  const size_t eb = gxc::xm_element_bytes(out_mode);
  gxc_status s = GXC_OK;
  if ((s = upload(b, &b->d_rowmap, gxc::host::make_rowmap(nant, cfg->autos != 0))) != GXC_OK) return fail(s);
  size_t spec_bytes = 0;
  b->maxL = maxL;
  for (Sub &u : b->sub) u.ntc = (u.cfg.numffts + C - 1) / C;
  for (Sub &u : b->sub) {
    u.npow = std::min(u.cfg.numffts, gxc::kPowerRows);
    u.spec_off = spec_bytes;
    spec_bytes += ((size_t)u.ntc * M * u.cfg.nchan * eb + 255) / 256 * 256;
    std::string why;
    if ((s = gxc::host::upload_tables(u.d, u.cfg.bw, u.cfg.nchan, &u.d_twiddle, &u.d_subfreq, &u.d_stepfreq,
                                      &why)) != GXC_OK)
      return fail(set_err(b, s, why));
    u.h_delays.assign((size_t)nant * 4, 0.0);
    if ((s = dmalloc(b, &u.d_delays, (size_t)nant * 4)) != GXC_OK) return fail(s);
    if ((s = dmalloc(b, &u.d_station_ptrs, (size_t)nant)) != GXC_OK) return fail(s);
    if ((s = dmalloc(b, &u.d_valid, (size_t)nant * u.cfg.numffts)) != GXC_OK) return fail(s);
    if ((s = dmalloc(b, &u.d_counts, (size_t)b->nrows)) != GXC_OK) return fail(s);
    if ((s = dmalloc(b, &u.d_vis, (size_t)b->nrows * 4 * u.cfg.nchan)) != GXC_OK) return fail(s);
    if (hipMemset(u.d_vis, 0, (size_t)b->nrows * 4 * u.cfg.nchan * sizeof(float2)) != hipSuccess)
      return fail(set_err(b, GXC_ERR_HIP, "hipMemset failed"));
    if (out_mode != gxc::OUT_F32) {
      if ((s = dmalloc(b, &u.d_scales, (size_t)M * u.cfg.nchan)) != GXC_OK) return fail(s);
      if ((s = dmalloc(b, &u.d_power, (size_t)u.npow * M * u.cfg.nchan)) != GXC_OK) return fail(s);
    }
  }
  for (int k = 0; k < b->nbuf; k++)
    if ((s = dmalloc(b, reinterpret_cast<unsigned char **>(&b->d_spec[k]), spec_bytes)) != GXC_OK) return fail(s);
  if ((s = assign_splits(b)) != GXC_OK) return fail(s);
  if ((s = build_tables(b)) != GXC_OK) return fail(s);
  *out = b;
  return GXC_OK;
}

void gxc_batch_destroy(gxc_batch *b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->xstream) hipStreamSynchronize(b->xstream);
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  for (hipEvent_t e : b->tev) hipEventDestroy(e);
  hipEvent_t evs[6] = {b->ev_fork, b->ev_join, b->ev_f[0], b->ev_f[1], b->ev_x[0], b->ev_x[1]};
  for (hipEvent_t e : evs)
    if (e) hipEventDestroy(e);
  for (Sub &u : b->sub) {
    hipFree(u.d_input_owned);
    hipFree(u.d_station_ptrs);
    hipFree(u.d_delays);
    hipFree(u.d_twiddle);
    hipFree(u.d_subfreq);
    hipFree(u.d_stepfreq);
    hipFree(u.d_valid);
    hipFree(u.d_counts);
    hipFree(u.d_vis);
    hipFree(u.d_partial);
    hipFree(u.d_comp);
    hipFree(u.d_polgains);
    u.vs.release();
    hipFree(u.d_scales);
    hipFree(u.d_power);
  }
  for (void *p : b->d_spec) hipFree(p);
  hipFree(b->d_rowmap);
  hipFree(b->d_fseg);
  hipFree(b->d_xseg);
  hipFree(b->d_powseg);
  hipFree(b->d_scaleseg);
  hipFree(b->d_countseg);
  hipFree(b->d_finseg);
  if (b->xstream) hipStreamDestroy(b->xstream);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

gxc_status gxc_batch_set_input(gxc_batch *b, int sub, const uint8_t *const *per_station) {
  Sub *u = get_sub(b, sub);
  if (!u || !per_station) return set_err(b, GXC_ERR_INVALID_ARG, "bad subband or null input");
  std::vector<const uint8_t *> in(per_station, per_station + b->cfg.nant);
  for (const uint8_t *p : in)
    if (!p) return set_err(b, GXC_ERR_INVALID_ARG, "null station buffer");
  GXB_HIP(b, hipStreamSynchronize(b->stream));
  GXB_TRY(leave_vdif(b, *u));
  GXB_TRY(ensure_owned_input(b, *u));
  if (u->input_device || !u->input_set) GXB_TRY(point_at_owned(b, *u));
  u->host_in = in;
  u->input_device = false;
  u->input_set = true;
  return GXC_OK;
}

gxc_status gxc_batch_set_input_device(gxc_batch *b, int sub, const void *const *per_station) {
  Sub *u = get_sub(b, sub);
  if (!u || !per_station) return set_err(b, GXC_ERR_INVALID_ARG, "bad subband or null input");
  std::vector<const void *> in(per_station, per_station + b->cfg.nant);
  for (const void *p : in)
    if (!p) return set_err(b, GXC_ERR_INVALID_ARG, "null station buffer");
  GXB_HIP(b, hipStreamSynchronize(b->stream));
  GXB_TRY(leave_vdif(b, *u));
  GXB_HIP(b, hipMemcpy(u->d_station_ptrs, in.data(), in.size() * sizeof(void *), hipMemcpyHostToDevice));
  u->input_device = true;
  u->input_set = true;
  return GXC_OK;
}

gxc_status gxc_batch_set_delays(gxc_batch *b, int sub, const double *const *poly3, const double *fileoffsets) {
  Sub *u = get_sub(b, sub);
  if (!u || !poly3 || !fileoffsets) return set_err(b, GXC_ERR_INVALID_ARG, "bad subband or null delays");
  for (int a = 0; a < b->cfg.nant; a++) {
    if (!poly3[a]) return set_err(b, GXC_ERR_INVALID_ARG, "null delay polynomial");
    for (int i = 0; i < 3; i++) u->h_delays[4 * a + i] = poly3[a][i];
    u->h_delays[4 * a + 3] = fileoffsets[a];
  }
  GXB_HIP(b, hipStreamSynchronize(b->stream));
  GXB_HIP(b, hipMemcpy(u->d_delays, u->h_delays.data(), u->h_delays.size() * sizeof(double),
                       hipMemcpyHostToDevice));
  u->delays_set = true;
  return GXC_OK;
}

gxc_status gxc_batch_set_input_vdif(gxc_batch *b, int sub, int station, const void *d_frames, size_t nbytes,
                                    const gxc_vdif_spec *spec, gxc_vdif_stats *stats) {
  Sub *u = get_sub(b, sub);
  if (!u) return set_err(b, GXC_ERR_INVALID_ARG, "bad subband");
  const int nant = b->cfg.nant;
  if (station < 0 || station >= nant) return set_err(b, GXC_ERR_INVALID_ARG, "station out of range");
  gxc::host::VdifCall call;
  std::string why;
  gxc_status st = gxc::host::vdif_prepare(&u->cfg, u->d.station_bytes, d_frames, nbytes, spec, b->stream, &call, &why);
  if (st != GXC_OK) return set_err(b, st, why);
  if (gxc::host::vdif_needs_growth(u->vs, station, call) || !u->vdif || !u->d_input_owned) {
    // buffers change under work that may still be queued: wait for it
    GXB_HIP(b, hipStreamSynchronize(b->stream));
  }
  GXB_TRY(ensure_owned_input(b, *u));
  st = gxc::host::vdif_reserve(u->vs, nant, station, call, &why);
  if (st != GXC_OK) return set_err(b, st, why);
  if (!u->vdif) {
    // enter VDIF mode: station pointers to the owned buffer, flags on
    GXB_TRY(point_at_owned(b, *u));
    u->input_device = false;
    u->input_set = false;
    u->vs.done.assign(nant, 0);
    GXB_TRY(vdif_mode(b, *u, true));
  }
  st = gxc::host::vdif_ingest(u->vs, &u->cfg, station, u->d_input_owned + (size_t)station * u->d.station_bytes,
                              u->d.station_bytes, d_frames, call, spec, stats, b->stream, &why);
  if (st != GXC_OK) return set_err(b, st, why);
  u->input_set = std::all_of(u->vs.done.begin(), u->vs.done.end(), [](char c) { return c != 0; });
  return GXC_OK;
}

gxc_status gxc_batch_get_frame_flags(gxc_batch *b, int sub, int station, uint8_t *out, size_t n, long long *nslots) {
  Sub *u = get_sub(b, sub);
  if (!u) return set_err(b, GXC_ERR_INVALID_ARG, "bad subband");
  if (station < 0 || station >= b->cfg.nant) return set_err(b, GXC_ERR_INVALID_ARG, "station out of range");
  if (!u->vdif) return set_err(b, GXC_ERR_STATE, "station has no VDIF ingest");
  std::string why;
  const gxc_status st = gxc::host::vdif_flags(u->vs, station, out, n, nslots, b->stream, &why);
  return st == GXC_OK ? GXC_OK : set_err(b, st, why);
}

gxc_status gxc_batch_set_pol_conversion(gxc_batch *b, int sub, const gxc_cf32 *gains) {
  Sub *u = get_sub(b, sub);
  if (!u) return set_err(b, GXC_ERR_INVALID_ARG, "bad subband");
  GXB_HIP(b, hipStreamSynchronize(b->stream));
  const bool on = gains != nullptr;
  if (on) {
    const size_t n = (size_t)b->cfg.nant * 2 * u->cfg.nchan;
    if (!u->d_polgains) GXB_TRY(dmalloc(b, &u->d_polgains, n));
    GXB_HIP(b, hipMemcpy(u->d_polgains, gains, n * sizeof(float2), hipMemcpyHostToDevice));
  }
  if (on != u->polconv && b->graph_exec) {  // the captured work changes
    hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
  }
  u->polconv = on;
  return GXC_OK;
}

gxc_status gxc_batch_process_async(gxc_batch *b) {
  if (!b) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null batch");
  return enqueue_process(b);
}

gxc_status gxc_batch_synchronize(gxc_batch *b) {
  if (!b) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null batch");
  GXB_HIP(b, hipStreamSynchronize(b->stream));
  collect_timings(b);
  return GXC_OK;
}

gxc_status gxc_batch_process(gxc_batch *b) {
  GXB_TRY(gxc_batch_process_async(b));
  return gxc_batch_synchronize(b);
}

gxc_status gxc_batch_get_visibilities(gxc_batch *b, int sub, gxc_cf32 *out, size_t nelem) {
  Sub *u = get_sub(b, sub);
  if (!u || !out) return set_err(b, GXC_ERR_INVALID_ARG, "bad subband or null output");
  const size_t n = (size_t)b->nrows * 4 * u->cfg.nchan;
  if (nelem < n) return set_err(b, GXC_ERR_INVALID_ARG, "output buffer too small");
  GXB_HIP(b, hipStreamSynchronize(b->stream));
  GXB_HIP(b, hipMemcpy(out, u->d_vis, n * sizeof(gxc_cf32), hipMemcpyDeviceToHost));
  return GXC_OK;
}

gxc_status gxc_batch_get_row_counts(gxc_batch *b, int sub, int *counts, size_t n) {
  Sub *u = get_sub(b, sub);
  if (!u || !counts) return set_err(b, GXC_ERR_INVALID_ARG, "bad subband or null output");
  if (n < (size_t)b->nrows) return set_err(b, GXC_ERR_INVALID_ARG, "output buffer too small");
  GXB_HIP(b, hipStreamSynchronize(b->stream));
  GXB_HIP(b, hipMemcpy(counts, u->d_counts, b->nrows * sizeof(int), hipMemcpyDeviceToHost));
  return GXC_OK;
}

gxc_status gxc_batch_set_profiling(gxc_batch *b, int enable) {
  if (!b) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null batch");
  b->profiling = enable != 0;
  return GXC_OK;
}

}  // extern "C"

namespace {

// Median of 3 timed subintegrations (ms) after one untimed one.
gxc_status time_process(gxc_batch *b, double *out) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  GXB_HIP(b, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return set_err(b, GXC_ERR_HIP, "hipEventCreate failed");
  }
  gxc_status st = enqueue_process(b);
  std::vector<float> ts;
  for (int r = 0; r < 3 && st == GXC_OK; r++) {
    float ms = 0.f;
    hipEventRecord(e0, b->stream);
    st = enqueue_process(b);
    if (st == GXC_OK && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      st = set_err(b, GXC_ERR_HIP, "event timing failed");
    ts.push_back(ms);
  }
  if (st == GXC_OK && hipStreamSynchronize(b->stream) != hipSuccess) st = set_err(b, GXC_ERR_HIP, "sync failed");
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st != GXC_OK) return st;
  std::sort(ts.begin(), ts.end());
  *out = ts[ts.size() / 2];
  return GXC_OK;
}

}  // namespace

extern "C" {

gxc_status gxc_batch_autotune(gxc_batch *b, gxc_tuning *out) {
  if (!b) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null batch");
  for (const Sub &u : b->sub)
    if (!u.input_set || !u.delays_set) return set_err(b, GXC_ERR_STATE, "input and delays must be set");
  GXB_HIP(b, hipStreamSynchronize(b->stream));
  const bool prof = b->profiling;
  b->profiling = false;
  const int graph_cfg = b->cfg.graph;
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  b->graph_exec = nullptr;
  b->cfg.graph = 0;
  gxc_tuning t{};
  int cand = 0;
  double def_ms = 0, best = 1e30;
  gxc_status st = time_process(b, &def_ms);
  cand++;
  int best_fb = b->fblock, best_units = b->xunits, best_mfma = b->cfg.fstage_mfma;
  bool any_mfma = false;  // a subband that fstage_mfma would give its own launch
  for (const Sub &u : b->sub) any_mfma = any_mfma || u.d.L >= 2048;
  for (int mf = 0; mf <= (any_mfma ? 2 : 0); mf++)
    for (int fb : {256, 512}) {
      if (st != GXC_OK) continue;
      for (int units : {1024, 2048, 4096, 8192}) {
        b->cfg.fstage_mfma = mf;
        b->fblock = fb;
        b->xunits = units;
        st = assign_splits(b);
        if (st == GXC_OK) st = build_tables(b);
        double ms = 0;
        if (st == GXC_OK) st = time_process(b, &ms);
        cand++;
        if (st != GXC_OK) break;
        if (ms < best) best = ms, best_fb = fb, best_units = units, best_mfma = mf;
      }
    }
  if (st == GXC_OK) {
    b->cfg.fstage_mfma = best_mfma;
    b->fblock = best_fb;
    b->xunits = best_units;
    st = assign_splits(b);
    if (st == GXC_OK) st = build_tables(b);
  }
  double graph_ms = 1e30;
  if (st == GXC_OK) {
    b->cfg.graph = 1;
    st = time_process(b, &graph_ms);
    cand++;
  }
  b->cfg.graph = st == GXC_OK ? (graph_ms <= best ? 1 : 0) : graph_cfg;
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  b->graph_exec = nullptr;
  b->profiling = prof;
  if (st != GXC_OK) return st;
  t.fblock = b->cfg.fstage_mfma == 1 ? gxc::kFBlockMfma : b->cfg.fstage_mfma == 2 ? GXC_FBLOCK_ONEPOL : b->fblock;
  t.xtile = gxc::xm_channels_per_group(b->M, b->out_mode);
  t.xsplits = b->xunits;  // for a batch: the X units per launch the splits aim at
  t.graph = b->cfg.graph;
  t.process_ms = std::min(best, graph_ms);
  t.default_ms = def_ms;
  t.candidates = cand;
  if (out) *out = t;
  return GXC_OK;
}

gxc_status gxc_batch_get_timings(const gxc_batch *b, gxc_timings *t) {
  if (!b || !t) return GXC_ERR_INVALID_ARG;
  *t = b->timings;
  return GXC_OK;
}

}  // extern "C"
