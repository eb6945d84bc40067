// The autotuner through the C-ABI (gxc_autotune): times launch shapes on the
// handle's own input with HIP events and keeps the fastest.
#include "gxc_handle.hpp"

#include <algorithm>
#include <vector>

using namespace gxc::api;

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
  if (gxc::fstage_mfma_supported(h->d.L, h->cfg.is_complex != 0, h->out_mode == gxc::OUT_F32))
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
      // the partials and, with several chunks, their compensation buffer
      const size_t nb = gxc::xm_partial_bytes(h->d.M, h->cfg.nchan, need);
      float *p = nullptr, *pc = nullptr;
      bool ok = hipMalloc((void **)&p, nb) == hipSuccess;
      if (ok && h->d_comp) ok = hipMalloc((void **)&pc, nb) == hipSuccess;
      if (ok) {
        hipFree(h->d_partial);
        h->d_partial = p;
        if (h->d_comp) {
          hipFree(h->d_comp);
          h->d_comp = pc;
        }
        h->partial_cap = need;
      } else {
        hipFree(p);
        (void)hipGetLastError();
        sps.erase(std::remove_if(sps.begin(), sps.end(), [&](int v) { return v > h->partial_cap; }), sps.end());
      }
    }
    int xt[4], nxt = gxc::xm_tiles(h->d.M, h->out_mode, xt);
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

}  // extern "C"
