// VDIF ingest kernels (vdif.hip): parameters and launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gxc {

// Device counters, in this order.  Duplicates are candidates - placed.
enum {
  kVdifCandidates = 0,
  kVdifPlaced,
  kVdifMissing,
  kVdifInvalid,
  kVdifOtherThread,
  kVdifOutOfWindow,
  kVdifMalformed,
  kVdifCounters
};

// The header size (16 bytes with the legacy bit of the first frame, else 32)
// and so the payload size, slot count and samples per frame are resolved on
// the device from the first frame.
struct VdifParams {
  const uint8_t *frames;  // the stream: nframes frames of frame_bytes
  long long nframes;
  int frame_bytes;        // a multiple of 8, > 32
  int thread_id, fps, ref_epoch, ref_frame;
  long long ref_second;
  int nbit, is_complex, flip8;    // flip8: XOR 0x80 (8-bit offset binary)
  int header_bytes;               // 16 or 32; 0: from the first frame's legacy bit
  unsigned *slot_src;             // [nslots_max] scratch: earliest frame per slot
  long long nslots_max;           // slots for a 32-byte header (the most)
  uint8_t *station;               // [station_bytes], a multiple of 8
  size_t station_bytes;
  uint8_t *frame_ok;              // [nslots_max]
  int *spf_out;                   // time samples per frame, written by the ingest
  long long *nslots_out;          // slots, written by the ingest
  unsigned long long *stats;      // [kVdifCounters]
};

// Clear, index and gather on stream s.  frames and station must be 8-byte
// aligned; every frame lies within the stream; nframes < 2^32 - 1 and
// station_bytes / 8 < 2^32; the frame size holds whole time samples for
// both header sizes.
// tally: also count placed and missing slots into stats.
hipError_t launch_vdif(const VdifParams &p, bool tally, hipStream_t s);

}  // namespace gxc
