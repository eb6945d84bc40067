// VDIF ingest through the C-ABI (gxc_set_input_vdif and its helpers):
// gcorr's ingest (ingest/src/ingest.c:267-355) on the GPU, kernels in
// vdif.hip.
#include "gxc_handle.hpp"

#include <algorithm>
#include <cstring>

using namespace gxc::api;
using gxc::host::Derived;
using gxc::host::derive;

extern "C" {

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
  gxc::host::VdifCall call;
  std::string why;
  gxc_status st = gxc::host::vdif_prepare(&h->cfg, h->d.station_bytes, d_frames, nbytes, spec, h->stream, &call, &why);
  if (st != GXC_OK) return set_err(h, st, why);
  if (gxc::host::vdif_needs_growth(h->vs, station, call) || !h->vdif || !h->d_input_owned) {
    // buffers change under work that may still be queued: wait for it
    GXC_HIP(h, hipStreamSynchronize(h->stream));
  }
  GXC_TRY(ensure_owned_input(h));
  st = gxc::host::vdif_reserve(h->vs, nant, station, call, &why);
  if (st != GXC_OK) return set_err(h, st, why);
  if (!h->vdif) {
    // enter VDIF mode: station pointers to the owned buffer, flags on
    std::vector<const uint8_t *> ptrs(nant);
    for (int a = 0; a < nant; a++) ptrs[a] = h->d_input_owned + (size_t)a * h->d.station_bytes;
    GXC_HIP(h, hipMemcpy(h->d_station_ptrs, ptrs.data(), nant * sizeof(void *), hipMemcpyHostToDevice));
    destroy_graph(h);
    h->vdif = true;
    h->input_device = false;
    h->input_set = false;
    h->vs.done.assign(nant, 0);
  }
  st = gxc::host::vdif_ingest(h->vs, &h->cfg, station, h->d_input_owned + (size_t)station * h->d.station_bytes,
                              h->d.station_bytes, d_frames, call, spec, stats, h->stream, &why);
  if (st != GXC_OK) return set_err(h, st, why);
  h->input_set = std::all_of(h->vs.done.begin(), h->vs.done.end(), [](char c) { return c != 0; });
  return GXC_OK;
}

gxc_status gxc_get_frame_flags(gxc_handle *h, int station, uint8_t *out, size_t n, long long *nslots) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (station < 0 || station >= h->cfg.nant) return set_err(h, GXC_ERR_INVALID_ARG, "station out of range");
  if (!h->vdif) return set_err(h, GXC_ERR_STATE, "station has no VDIF ingest");
  std::string why;
  const gxc_status st = gxc::host::vdif_flags(h->vs, station, out, n, nslots, h->stream, &why);
  return st == GXC_OK ? GXC_OK : set_err(h, st, why);
}

}  // extern "C"
