// gxc_xengine_noalloc: the X-stage of gxc_xengine over device spectra into
// device visibilities without any device allocation.  Its working memory
// (row map, row counts, packed spectra for the 16/8-bit forms and the split
// partials) is carved from one fixed array of this library, `scratch`, of
// which each GPU holds its own copy for the life of the process.
//
// Host code reaches the array through the address hipGetSymbolAddress
// returns on each device (the array's name is only a host-side stand-in),
// fetched once per device and kept.  Calls on one GPU share the array and
// keep the last call's contents; everything a call reads is written by that
// call first (fill kernels, no hipMemsetAsync, so the call may sit inside a
// captured graph).  Calls that overlap on one GPU race.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>

#include "gxc_handle.hpp"
#include "gxc_xengine.hpp"

// The working memory of gxc_xengine_noalloc, one copy per GPU.
static __device__ __align__(256) unsigned char scratch[GXC_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kScratchBytes = (size_t)GXC_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "whole 256-byte blocks");

// Device address of `scratch` on the current device, looked up once per
// device.
hipError_t scratch_address(unsigned char **out) {
  static std::mutex mu;
  static std::map<int, unsigned char *> addr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(dev);
  if (it != addr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  addr[dev] = static_cast<unsigned char *>(p);
  *out = addr[dev];
  return hipSuccess;
}

// Rows in fxkernel's order (cross baselines (0,1),(0,2)..(N-2,N-1),
// fxkernel.cpp:328-355), then one auto row per station; every row counts
// all nfft FFTs.
__global__ void rows_kernel(int2 *rowmap, int *counts, int nant, int nrows, int nfft) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int nbl = nant * (nant - 1) / 2;
  int2 jk;
  if (r < nbl) {
    int j = 0, rem = r;
    while (rem >= nant - 1 - j) {
      rem -= nant - 1 - j;
      ++j;
    }
    jk = make_int2(j, j + 1 + rem);
  } else {
    jk = make_int2(r - nbl, r - nbl);
  }
  rowmap[r] = jk;
  counts[r] = nfft;
}

inline size_t up256(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

extern "C" gxc_status gxc_xengine_noalloc(const gxc_cf32 *d_spectra, int nant, int nchan, int nfft, int autos,
                                          int precision, gxc_cf32 *d_vis, void *stream) {
  if (!d_spectra || !d_vis || nant < 1 || nchan < 1 || nfft < 1)
    return gxc::api::set_err(nullptr, GXC_ERR_INVALID_ARG, "bad xengine arguments");
  if (precision < GXC_PREC_FP32 || precision > GXC_PREC_FP8)
    return gxc::api::set_err(nullptr, GXC_ERR_INVALID_ARG, "unknown precision");
  const int M = 2 * nant, mode = gxc::host::out_mode_of(precision);
  if (gxc::xm_blocks(M) == 0 || nchan % gxc::xm_channels_per_group(M, mode) != 0)
    return gxc::api::set_err(nullptr, GXC_ERR_UNSUPPORTED, "X-engine shape not supported");
  int dev = 0;
  std::string why;
  const gxc_status ds = gxc::host::check_device(&dev, &why);
  if (ds != GXC_OK) return gxc::api::set_err(nullptr, ds, why);
  const int nrows = nant * (nant - 1) / 2 + (autos ? nant : 0);
  if (nrows < 1) return gxc::api::set_err(nullptr, GXC_ERR_INVALID_ARG, "no rows to correlate");
  // Layout of the scratch: row map, row counts, packed spectra (16/8-bit
  // forms), then the split partials; the splits are those gxc_xengine
  // uses, fewer if they do not fit.
  const size_t nspec = (size_t)nfft * M * nchan;
  const size_t off_counts = up256((size_t)nrows * sizeof(int2));
  const size_t off_spec = off_counts + up256((size_t)nrows * sizeof(int));
  const size_t spec_bytes = mode == gxc::OUT_F32 ? 0 : up256(nspec * gxc::xm_element_bytes(mode));
  const size_t off_part = off_spec + spec_bytes;
  const size_t per_split = gxc::xm_partial_bytes(M, nchan, 1);
  if (off_part > kScratchBytes || kScratchBytes - off_part < per_split)
    return gxc::api::set_err(nullptr, GXC_ERR_INVALID_ARG,
                               "gxc_xengine_noalloc: working memory of this shape exceeds the " +
                                   std::to_string(kScratchBytes >> 20) + " MiB scratch");
  const int splits = (int)std::min<size_t>((size_t)gxc::xm_choose_splits(M, nchan, nfft, mode),
                                           (kScratchBytes - off_part) / per_split);
  unsigned char *base = nullptr;
  hipError_t e = scratch_address(&base);
  if (e != hipSuccess)
    return gxc::api::set_err(nullptr, GXC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  hipStream_t s = static_cast<hipStream_t>(stream);
  int2 *rowmap = reinterpret_cast<int2 *>(base);
  int *counts = reinterpret_cast<int *>(base + off_counts);
  float *partial = reinterpret_cast<float *>(base + off_part);
  hipLaunchKernelGGL(rows_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, s, rowmap, counts, nant,
                     nrows, nfft);
  if ((e = hipGetLastError()) != hipSuccess) return gxc::api::set_err(nullptr, GXC_ERR_HIP, hipGetErrorString(e));
  const void *spec = d_spectra;
  if (mode != gxc::OUT_F32) {
    void *packed = base + off_spec;
    if ((e = gxc::launch_quantise(packed, reinterpret_cast<const float2 *>(d_spectra), nspec, mode, s)) !=
        hipSuccess)
      return gxc::api::set_err(nullptr, GXC_ERR_HIP, hipGetErrorString(e));
    spec = packed;
  }
  gxc::XMParams x{};
  x.spectra = spec;
  x.partial = partial;
  x.comp = nullptr;
  x.nt = nfft;
  x.M = M;
  x.nchan = nchan;
  x.splits = splits;
  x.accumulate = 0;
  if ((e = gxc::launch_xmfma(x, mode, 0, s)) != hipSuccess) return gxc::api::set_err(nullptr, GXC_ERR_HIP, hipGetErrorString(e));
  gxc::FinMParams f{};
  f.partial = partial;
  f.comp = nullptr;
  f.scales = nullptr;  // spectra packed with unit scale
  f.rowmap = rowmap;
  f.counts = counts;
  f.vis = reinterpret_cast<float2 *>(d_vis);
  f.M = M;
  f.nchan = nchan;
  f.nrows = nrows;
  f.splits = splits;
  if ((e = gxc::launch_finalise_mixed(f, s)) != hipSuccess)
    return gxc::api::set_err(nullptr, GXC_ERR_HIP, hipGetErrorString(e));
  return GXC_OK;
}
