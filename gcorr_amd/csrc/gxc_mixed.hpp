// Mixed-precision (FP16 / BF16 / FP8) X-engine path.
#pragma once

#include <string>

#include "../../include/gxcorr.h"
#include "gxc_internal.hpp"

namespace gxc {

struct MixedState;

gxc_status mixed_create(MixedState **out, const gxc_config &cfg, int M, int nrows, int nbl, int chunk,
                        uint8_t *d_valid, int2 *d_rowmap, int *d_counts, float2 *d_vis, std::string *err);
void mixed_destroy(MixedState *m);
gxc_status mixed_xengine(const gxc_cf32 *spectra, int nant, int nchan, int nfft, int autos, int precision,
                         gxc_cf32 *out, std::string *err);

}  // namespace gxc
