// Host steps of the VDIF ingest shared by gxc_set_input_vdif and
// gxc_batch_set_input_vdif: argument checks, buffer sizing and the launch of
// the ingest kernels (vdif.hip).
#include "gxc_vdif_state.hpp"

#include <algorithm>
#include <climits>

#include "gxc_vdif.hpp"

namespace gxc {
namespace host {

namespace {

template <typename T>
gxc_status alloc(T **p, size_t nelem, std::string *why) {
  *p = nullptr;
  hipError_t e = hipMalloc((void **)p, (nelem ? nelem : 1) * sizeof(T));
  if (e != hipSuccess) {
    *why = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    return GXC_ERR_OUT_OF_MEMORY;
  }
  return GXC_OK;
}

gxc_status hip_fail(hipError_t e, const char *what, std::string *why) {
  *why = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? GXC_ERR_OUT_OF_MEMORY : GXC_ERR_HIP;
}

}  // namespace

void VdifState::release() {
  for (uint8_t *f : d_flags) hipFree(f);
  hipFree(d_flag_ptrs);
  hipFree(d_spf);
  hipFree(d_nslots);
  hipFree(d_slot_src);
  hipFree(d_vstats);
  *this = VdifState();
}

gxc_status vdif_prepare(const gxc_config *cfg, size_t station_bytes, const void *d_frames, size_t nbytes,
                        const gxc_vdif_spec *spec, hipStream_t s, VdifCall *call, std::string *why) {
  if (!d_frames || !spec) return *why = "null frames or spec", GXC_ERR_INVALID_ARG;
  if (reinterpret_cast<uintptr_t>(d_frames) % 8)
    return *why = "VDIF frames must be 8-byte aligned in device memory", GXC_ERR_INVALID_ARG;
  if (spec->frames_per_second <= 0) return *why = "frames_per_second must be positive", GXC_ERR_INVALID_ARG;
  if (spec->frame_bytes < 0) return *why = "negative frame_bytes", GXC_ERR_INVALID_ARG;
  if (spec->header_bytes != 0 && spec->header_bytes != 16 && spec->header_bytes != 32)
    return *why = "header_bytes must be 0, 16 or 32", GXC_ERR_INVALID_ARG;
  if (nbytes < 32) return *why = "VDIF stream shorter than one header", GXC_ERR_INVALID_ARG;
  long long fb = spec->frame_bytes;
  if (fb == 0) {  // frame size from the first header: the one host read of the stream
    uint32_t w[4];
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(w, d_frames, sizeof(w), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "reading the first VDIF header", why);
    fb = 8LL * (w[2] & 0xffffffu);
    // The frame geometry (and, with header_bytes 0, the header size) comes
    // from this header: reject it here rather than let every frame of the
    // stream be judged by a malformed first one.
    const int legacy = (int)((w[0] >> 30) & 1u);
    if (((w[2] >> 24) & 0x1fu) != 1u || (int)((w[3] >> 26) & 0x1fu) + 1 != cfg->nbit ||
        (int)(w[3] >> 31) != (cfg->is_complex ? 1 : 0) ||
        (spec->header_bytes != 0 && legacy != (spec->header_bytes == 16 ? 1 : 0))) {
      *why = "the first VDIF frame's header does not match the configuration (two channels, " +
             std::to_string(cfg->nbit) + " bits, " + (cfg->is_complex ? "complex" : "real") +
             (spec->header_bytes ? ", " + std::to_string(spec->header_bytes) + "-byte header" : std::string()) +
             ")";
      return GXC_ERR_INVALID_ARG;
    }
  }
  // The header size (16 legacy, 32 otherwise) is read from the first frame on
  // the device; buffers are sized for the larger slot count of the two.
  const long long nslots_max = gxc_vdif_slots(cfg, (int)std::min<long long>(fb, INT32_MAX), 0);
  if (fb <= 32 || fb > INT32_MAX || fb % 8 || nslots_max <= 0 || gxc_vdif_slots(cfg, (int)fb, 1) <= 0) {
    *why = "VDIF frame size " + std::to_string(fb) + " does not hold a whole number of samples";
    return GXC_ERR_INVALID_ARG;
  }
  const long long nframes = (long long)(nbytes / (size_t)fb);
  if (nframes >= (long long)UINT32_MAX) return *why = "too many VDIF frames in one call", GXC_ERR_INVALID_ARG;
  if (station_bytes / 8 >= (size_t)UINT32_MAX) {
    *why = "station buffers of 32 GiB or more are not supported by the VDIF ingest";
    return GXC_ERR_UNSUPPORTED;
  }
  call->frame_bytes = fb;
  call->nslots_max = nslots_max;
  call->nframes = nframes;
  return GXC_OK;
}

bool vdif_needs_growth(const VdifState &v, int station, const VdifCall &call) {
  return !v.d_flag_ptrs || v.flag_cap[station] < call.nslots_max || v.slot_cap < call.nslots_max;
}

gxc_status vdif_reserve(VdifState &v, int nant, int station, const VdifCall &call, std::string *why) {
  if (!v.d_flag_ptrs) {
    gxc_status st = alloc(&v.d_flag_ptrs, (size_t)nant, why);
    if (st == GXC_OK) st = alloc(&v.d_spf, (size_t)nant, why);
    if (st == GXC_OK) st = alloc(&v.d_nslots, (size_t)nant, why);
    if (st == GXC_OK) st = alloc(&v.d_vstats, (size_t)kVdifCounters, why);
    if (st != GXC_OK) return st;
    v.d_flags.assign(nant, nullptr);
    v.flag_cap.assign(nant, 0);
    v.done.assign(nant, 0);
  }
  if (v.flag_cap[station] < call.nslots_max) {
    hipFree(v.d_flags[station]);
    v.d_flags[station] = nullptr;
    v.flag_cap[station] = 0;
    gxc_status st = alloc(&v.d_flags[station], (size_t)call.nslots_max, why);
    if (st != GXC_OK) return st;
    v.flag_cap[station] = call.nslots_max;
    hipError_t e = hipMemcpy(v.d_flag_ptrs + station, &v.d_flags[station], sizeof(void *), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy of the frame-flag pointer", why);
  }
  if (v.slot_cap < call.nslots_max) {
    hipFree(v.d_slot_src);
    v.d_slot_src = nullptr;
    v.slot_cap = 0;
    gxc_status st = alloc(&v.d_slot_src, (size_t)call.nslots_max, why);
    if (st != GXC_OK) return st;
    v.slot_cap = call.nslots_max;
  }
  return GXC_OK;
}

gxc_status vdif_ingest(VdifState &v, const gxc_config *cfg, int station, uint8_t *dst, size_t station_bytes,
                       const void *d_frames, const VdifCall &call, const gxc_vdif_spec *spec,
                       gxc_vdif_stats *stats, hipStream_t s, std::string *why) {
  VdifParams vp{};
  vp.frames = static_cast<const uint8_t *>(d_frames);
  vp.nframes = call.nframes;
  vp.frame_bytes = (int)call.frame_bytes;
  vp.thread_id = spec->thread_id;
  vp.fps = spec->frames_per_second;
  vp.ref_epoch = spec->ref_epoch;
  vp.ref_frame = spec->ref_frame;
  vp.ref_second = spec->ref_second;
  vp.nbit = cfg->nbit;
  vp.is_complex = cfg->is_complex ? 1 : 0;
  vp.flip8 = cfg->nbit == 8 && spec->offset_binary8 ? 1 : 0;
  vp.header_bytes = spec->header_bytes;
  vp.slot_src = v.d_slot_src;
  vp.nslots_max = call.nslots_max;
  vp.station = dst;
  vp.station_bytes = station_bytes;
  vp.frame_ok = v.d_flags[station];
  vp.spf_out = v.d_spf + station;
  vp.nslots_out = v.d_nslots + station;
  vp.stats = v.d_vstats;
  hipError_t e = launch_vdif(vp, stats != nullptr, s);
  if (e != hipSuccess) return hip_fail(e, "launch_vdif", why);
  v.done[station] = 1;
  if (stats) {
    unsigned long long c[kVdifCounters];
    e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(c, v.d_vstats, sizeof(c), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "reading the VDIF counters", why);
    stats->frames = call.nframes;
    stats->placed = (long long)c[kVdifPlaced];
    stats->missing = (long long)c[kVdifMissing];
    stats->duplicate = (long long)(c[kVdifCandidates] - c[kVdifPlaced]);
    stats->invalid = (long long)c[kVdifInvalid];
    stats->other_thread = (long long)c[kVdifOtherThread];
    stats->out_of_window = (long long)c[kVdifOutOfWindow];
    stats->malformed = (long long)c[kVdifMalformed];
  }
  return GXC_OK;
}

gxc_status vdif_flags(const VdifState &v, int station, uint8_t *out, size_t n, long long *nslots, hipStream_t s,
                      std::string *why) {
  if (v.done.empty() || !v.done[station]) return *why = "station has no VDIF ingest", GXC_ERR_STATE;
  long long ns = 0;
  hipError_t e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipMemcpy(&ns, v.d_nslots + station, sizeof(ns), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading the slot count", why);
  if (nslots) *nslots = ns;
  if (!out) return GXC_OK;
  if (n < (size_t)ns) return *why = "flag buffer too small", GXC_ERR_INVALID_ARG;
  e = hipMemcpy(out, v.d_flags[station], (size_t)ns, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading the frame flags", why);
  return GXC_OK;
}

}  // namespace host
}  // namespace gxc
