// Pre-correlation conversions through the C-ABI: the requantisers
// (requant.hip) and polarisation conversion (polconv.hip, applied to the
// visibilities in enqueue_device_work).
#include "gxc_handle.hpp"

#include <cstdint>

#include "gxc_requant.hpp"

using namespace gxc::api;

extern "C" {

namespace {
gxc_status requant_args(const float *a, const float *b, const void *out, size_t outalign) {
  if (!a || !b || !out) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null buffer");
  if (reinterpret_cast<uintptr_t>(a) % 16 || reinterpret_cast<uintptr_t>(b) % 16 ||
      reinterpret_cast<uintptr_t>(out) % outalign)
    return set_err(nullptr, GXC_ERR_INVALID_ARG, "misaligned requantiser buffer");
  return GXC_OK;
}
}  // namespace

gxc_status gxc_requantise_2bit(const float *d_pol0, const float *d_pol1, size_t nsamples, float threshold,
                               uint8_t *d_out, void *stream) {
  GXC_TRY(requant_args(d_pol0, d_pol1, d_out, 8));
  if (nsamples % 2) return set_err(nullptr, GXC_ERR_INVALID_ARG, "2-bit requantiser needs an even sample count");
  if (!(threshold > 0.f)) return set_err(nullptr, GXC_ERR_INVALID_ARG, "threshold must be positive");
  GXC_HIP(nullptr, gxc::launch_requant2(d_pol0, d_pol1, (long long)nsamples, threshold, d_out,
                                        static_cast<hipStream_t>(stream)));
  return GXC_OK;
}

gxc_status gxc_requantise_8bit(const float *d_pol0, const float *d_pol1, size_t nsamples, float scale,
                               uint8_t *d_out, void *stream) {
  GXC_TRY(requant_args(d_pol0, d_pol1, d_out, 16));
  GXC_HIP(nullptr, gxc::launch_requant8(d_pol0, d_pol1, (long long)nsamples, scale, d_out,
                                        static_cast<hipStream_t>(stream)));
  return GXC_OK;
}

gxc_status gxc_set_pol_conversion(gxc_handle *h, const gxc_cf32 *gains) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  const bool on = gains != nullptr;
  if (on) {
    const size_t n = (size_t)h->cfg.nant * 2 * h->cfg.nchan;
    if (!h->d_polgains) GXC_TRY(dmalloc(h, &h->d_polgains, n));
    GXC_HIP(h, hipMemcpy(h->d_polgains, gains, n * sizeof(float2), hipMemcpyHostToDevice));
  }
  if (on != h->polconv) destroy_graph(h);  // the captured work changes
  h->polconv = on;
  return GXC_OK;
}

}  // extern "C"
