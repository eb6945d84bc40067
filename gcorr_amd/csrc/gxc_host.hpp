// Host-side pieces shared by the single-subband handle (gxc_api.cpp) and the
// multi-subband batch (gxc_batch.cpp): configuration checks and derived
// constants, the visibility row order, per-FFT-length constant tables.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/gxcorr.h"
#include "gxc_internal.hpp"

namespace gxc {
namespace host {

// Constants derived from a configuration (fxkernel.cpp:30-180).
struct Derived {
  Geometry g;
  int L = 0, M = 0;
  int maxoffset = 0, nbl = 0, nrows = 0;
  double sampletime = 0;
  size_t station_bytes = 0;
};

// Validate cfg and derive its constants; on failure *why says what is wrong.
gxc_status derive(const gxc_config *c, Derived *d, std::string *why);

// OUT_* spectra format of a GXC_PREC_* precision.
int out_mode_of(int precision);

// Rows: cross baselines in fxkernel order (fxkernel.cpp:328-355), then autos.
std::vector<int2> make_rowmap(int nant, bool autos);

// The device is present and is a gfx950.
gxc_status check_device(int *dev, std::string *why);

// Allocate and fill the twiddle table [L] and the fractional-sample channel
// frequency tables [s], [fracsteps] of fxkernel.cpp:171-175.
gxc_status upload_tables(const Derived &d, double bw, int nchan, float2 **twiddle, float **subfreq,
                         float **stepfreq, std::string *why);

}  // namespace host
}  // namespace gxc
