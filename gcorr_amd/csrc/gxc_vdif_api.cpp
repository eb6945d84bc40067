// VDIF ingest through the C-ABI (gxc_set_input_vdif and its helpers):
// gcorr's ingest (ingest/src/ingest.c:267-355) on the GPU, kernels in
// vdif.hip.
#include "gxc_handle.hpp"

#include <algorithm>
#include <climits>
#include <cstring>

#include "gxc_vdif.hpp"

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

}  // extern "C"
