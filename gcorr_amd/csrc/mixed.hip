// Mixed-precision X-engine path (FP16 / BF16 / FP8 spectra on MFMA).
#include "gxc_mixed.hpp"

namespace gxc {

struct MixedState {
  int dummy;
};

gxc_status mixed_create(MixedState **out, const gxc_config &, int, int, int, int, uint8_t *, int2 *, int *,
                        float2 *, std::string *err) {
  *out = nullptr;
  *err = "mixed precision is not available in this build";
  return GXC_ERR_UNSUPPORTED;
}

void mixed_destroy(MixedState *m) { delete m; }

gxc_status mixed_xengine(const gxc_cf32 *, int, int, int, int, int, gxc_cf32 *, std::string *err) {
  *err = "mixed precision is not available in this build";
  return GXC_ERR_UNSUPPORTED;
}

}  // namespace gxc
