// VDIF ingest on the GPU: frames resident in HBM -> one station's packed
// input buffer plus per-frame presence flags (gcorr ingest,
// ingest/src/ingest.c:267-355; header fields as examples/mark/dump_vdif.py
// decodes them).
//   vdif_clear_kernel   slot -> frame map set to "none", counters zeroed;
//   vdif_index_kernel   one lane per frame: header checks, slot from the time
//                       stamp, atomicMin so the earliest frame of a slot wins;
//   vdif_gather_kernel  over the station buffer in 8-byte words, kGatherWords
//                       per workgroup, each lane kGatherWords/256 coalesced
//                       words: the winning payload of the word's slot (or
//                       zeros); the lane holding a slot's first word writes
//                       its flag;
//   vdif_tally_kernel   (only when counts are asked for) one workgroup sums
//                       the flags into placed / missing: one atomic per slot
//                       would serialise thousands of same-address atomics.
// HBM traffic is the payloads read once and the station buffer written once.
#include <hip/hip_runtime.h>

#include "gxc_vdif.hpp"

namespace gxc {
namespace {

constexpr unsigned kNone = 0xffffffffu;
constexpr int kGatherWords = 1024;  // 8-byte words per gather workgroup

__device__ __forceinline__ void count(unsigned long long *stats, int k, bool pred) {
  // one atomic per wave and counter
  const unsigned long long m = __ballot(pred);
  if (m && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)m) - 1))
    atomicAdd(stats + k, (unsigned long long)__popcll(m));
}

struct Geom {
  int legacy, hdr, payload;
  long long nslots;
};

// Frame geometry: the header size given, else the first frame's legacy bit.
__device__ __forceinline__ Geom geom(const VdifParams &p) {
  Geom g;
  g.legacy = p.header_bytes ? (p.header_bytes == 16 ? 1 : 0)
                            : (int)((reinterpret_cast<const unsigned *>(p.frames)[0] >> 30) & 1u);
  g.hdr = g.legacy ? 16 : 32;
  g.payload = p.frame_bytes - g.hdr;
  g.nslots = ((long long)p.station_bytes + g.payload - 1) / g.payload;
  return g;
}

__global__ void vdif_clear_kernel(VdifParams p) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p.nslots_max) p.slot_src[i] = kNone;
  if (i < kVdifCounters) p.stats[i] = 0;
  if (i == 0) {
    const Geom g = geom(p);
    *p.nslots_out = g.nslots;
    *p.spf_out = (int)(8LL * g.payload / (p.nbit * 2 * (p.is_complex ? 2 : 1)));
  }
}

__global__ void vdif_index_kernel(VdifParams p) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Geom g = geom(p);
  bool cand = false, other = false, bad = false, inval = false, outw = false;
  long long rel = -1;
  if (i < p.nframes) {
    const unsigned *w = reinterpret_cast<const unsigned *>(p.frames + i * (long long)p.frame_bytes);
    const unsigned w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
    const int thread = (int)((w3 >> 16) & 0x3ffu);
    const long long frame = w1 & 0xffffffu;
    if (p.thread_id >= 0 && thread != p.thread_id) {
      other = true;
    } else if (8LL * (w2 & 0xffffffu) != p.frame_bytes || (int)((w0 >> 30) & 1u) != g.legacy ||
               ((w2 >> 24) & 0x1fu) != 1u || (int)((w3 >> 26) & 0x1fu) + 1 != p.nbit ||
               (int)(w3 >> 31) != p.is_complex || frame >= p.fps) {
      bad = true;
    } else if (w0 >> 31) {
      inval = true;
    } else {
      const long long seconds = w0 & 0x3fffffffu;
      rel = (seconds - p.ref_second) * p.fps + frame - p.ref_frame;
      if ((int)((w1 >> 24) & 0x3fu) != p.ref_epoch || rel < 0 || rel >= g.nslots)
        outw = true;
      else
        cand = true;
    }
  }
  if (cand) atomicMin(p.slot_src + rel, (unsigned)i);
  count(p.stats, kVdifCandidates, cand);
  count(p.stats, kVdifOtherThread, other);
  count(p.stats, kVdifMalformed, bad);
  count(p.stats, kVdifInvalid, inval);
  count(p.stats, kVdifOutOfWindow, outw);
}

__global__ void __launch_bounds__(256) vdif_gather_kernel(VdifParams p) {
  const Geom g = geom(p);
  const unsigned wps = (unsigned)(g.payload / 8);  // words per slot
  const long long nw = (long long)p.station_bytes / 8;
  const unsigned flip = p.flip8 ? 0x80808080u : 0u;
  const long long base = (long long)blockIdx.x * kGatherWords + threadIdx.x;
  uint2 *dst = reinterpret_cast<uint2 *>(p.station);
#pragma unroll
  for (int r = 0; r < kGatherWords / 256; r++) {
    const long long k = base + r * 256;
    if (k < nw) {
      const unsigned s = (unsigned)k / wps, o = (unsigned)k - s * wps;  // nw < 2^32
      const unsigned src = p.slot_src[s];
      uint2 v = make_uint2(0u, 0u);
      if (src != kNone) {
        v = *reinterpret_cast<const uint2 *>(p.frames + (long long)src * p.frame_bytes + g.hdr + 8LL * o);
        v.x ^= flip;
        v.y ^= flip;
      }
      dst[k] = v;
      if (o == 0) p.frame_ok[s] = src != kNone ? 1 : 0;
    }
  }
}

__global__ void __launch_bounds__(1024) vdif_tally_kernel(VdifParams p) {
  const long long nslots = geom(p).nslots;
  unsigned long long placed = 0;
  for (long long s = threadIdx.x; s < nslots; s += blockDim.x) placed += p.frame_ok[s];
  __shared__ unsigned long long part[16];
  for (int off = 32; off > 0; off >>= 1) placed += __shfl_down(placed, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = placed;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += part[i];
    p.stats[kVdifPlaced] = t;
    p.stats[kVdifMissing] = (unsigned long long)nslots - t;
  }
}

}  // namespace

hipError_t launch_vdif(const VdifParams &p, bool tally, hipStream_t s) {
  const long long nclear = p.nslots_max > kVdifCounters ? p.nslots_max : kVdifCounters;
  hipLaunchKernelGGL(vdif_clear_kernel, dim3((unsigned)((nclear + 255) / 256)), dim3(256), 0, s, p);
  if (p.nframes > 0)
    hipLaunchKernelGGL(vdif_index_kernel, dim3((unsigned)((p.nframes + 255) / 256)), dim3(256), 0, s, p);
  const long long nw = (long long)p.station_bytes / 8;
  hipLaunchKernelGGL(vdif_gather_kernel, dim3((unsigned)((nw + kGatherWords - 1) / kGatherWords)), dim3(256), 0,
                     s, p);
  if (tally) hipLaunchKernelGGL(vdif_tally_kernel, dim3(1), dim3(1024), 0, s, p);
  return hipGetLastError();
}

}  // namespace gxc
