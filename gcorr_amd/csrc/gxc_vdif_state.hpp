// VDIF ingest of one set of stations (a gxc_handle, or one subband of a
// gxc_batch): the device buffers that the ingest kernels of vdif.hip fill
// (frame flags, samples per frame, slot counts, counters) and the host steps
// of gxc_set_input_vdif (ingest/src/ingest.c:267-355) around them.
#pragma once

#include "../../include/gxcorr.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace gxc {
namespace host {

struct VdifState {
  std::vector<char> done;              // stations ingested since entering VDIF mode
  std::vector<uint8_t *> d_flags;      // [nant] frame flags, capacity flag_cap
  std::vector<long long> flag_cap;
  const uint8_t **d_flag_ptrs = nullptr;  // [nant] = d_flags (FParams.frame_ok)
  int *d_spf = nullptr;                   // [nant] time samples per frame (FParams.frame_samples)
  long long *d_nslots = nullptr;          // [nant] slots of the last ingest
  unsigned *d_slot_src = nullptr;         // [slot_cap] ingest scratch
  long long slot_cap = 0;
  unsigned long long *d_vstats = nullptr;  // [kVdifCounters]

  void release();
};

// One ingest call's geometry, from gxc_set_input_vdif's arguments.
struct VdifCall {
  long long frame_bytes = 0, nslots_max = 0, nframes = 0;
};

// Checks the arguments and works out the frame size (read from the first
// header, after synchronising `s`, when spec->frame_bytes is 0) and the slot
// count for a station of `station_bytes`.  On failure returns the status and
// sets *why.
gxc_status vdif_prepare(const gxc_config *cfg, size_t station_bytes, const void *d_frames, size_t nbytes,
                        const gxc_vdif_spec *spec, hipStream_t s, VdifCall *call, std::string *why);

// True when station's buffers must grow (the caller then waits for queued
// work that may still read them).
bool vdif_needs_growth(const VdifState &v, int station, const VdifCall &call);

// Allocates or grows the buffers for `station` of `nant`.
gxc_status vdif_reserve(VdifState &v, int nant, int station, const VdifCall &call, std::string *why);

// Enqueues the ingest of one station's frames into its buffer `dst` on `s`,
// marks the station done, and with `stats` reads the counters back.
gxc_status vdif_ingest(VdifState &v, const gxc_config *cfg, int station, uint8_t *dst, size_t station_bytes,
                       const void *d_frames, const VdifCall &call, const gxc_vdif_spec *spec,
                       gxc_vdif_stats *stats, hipStream_t s, std::string *why);

// gxc_get_frame_flags for one station of `v` (synchronises `s`).
gxc_status vdif_flags(const VdifState &v, int station, uint8_t *out, size_t n, long long *nslots, hipStream_t s,
                      std::string *why);

}  // namespace host
}  // namespace gxc
