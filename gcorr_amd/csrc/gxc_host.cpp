// Shared host helpers (gxc_host.hpp).
#include "gxc_host.hpp"
#include "gxc_xengine.hpp"

#include <cmath>
#include <cstring>

namespace gxc {
namespace host {

gxc_status derive(const gxc_config *c, Derived *d, std::string *why) {
  if (!c) {
    *why = "null config";
    return GXC_ERR_INVALID_ARG;
  }
  if (c->nant < 1) {
    *why = "nant must be >= 1";
    return GXC_ERR_INVALID_ARG;
  }
  if (c->numffts < 1) {
    *why = "numffts must be >= 1";
    return GXC_ERR_INVALID_ARG;
  }
  if (c->nchan < 2 || (c->nchan & (c->nchan - 1))) {
    *why = "nchan must be a power of two";
    return GXC_ERR_INVALID_ARG;
  }
  if (!(c->bw > 0)) {
    *why = "bandwidth must be positive";
    return GXC_ERR_INVALID_ARG;
  }
  if (c->npol != 2) {
    *why = "only dual-polarisation data is supported (npol = 2)";
    return GXC_ERR_UNSUPPORTED;
  }
  if (c->is_complex ? c->nbit != 8 : c->nbit != 2) {
    // fxkernel.cpp:46-55,489-496: 2-bit real is the only supported input;
    // complex 8-bit is this build's extension.
    *why = c->is_complex ? "complex data must be 8-bit" : "real data must be 2-bit";
    return GXC_ERR_UNSUPPORTED;
  }
  if (c->precision < GXC_PREC_FP32 || c->precision > GXC_PREC_FP8) {
    *why = "unknown precision";
    return GXC_ERR_INVALID_ARG;
  }
  d->g = Geometry::make(c->nchan, c->is_complex != 0);
  d->L = d->g.L;
  d->M = 2 * c->nant;
  if (!gxc::fstage_supported(d->L, c->is_complex != 0)) {
    *why = "FFT length " + std::to_string(d->L) + " not supported (powers of two 64..32768)";
    return GXC_ERR_UNSUPPORTED;
  }
  if (gxc::xm_blocks(d->M) == 0) {
    *why = "at most 32 stations (64 inputs) per handle";
    return GXC_ERR_UNSUPPORTED;
  }
  if ((long long)c->numffts * d->L + d->L > 0x7fffffffLL) {
    *why = "subintegration too long for 32-bit sample offsets";
    return GXC_ERR_UNSUPPORTED;
  }
  d->maxoffset = (c->numffts - 1) * d->L;  // fxkernel.cpp:272
  d->nbl = c->nant * (c->nant - 1) / 2;
  d->nrows = d->nbl + (c->autos ? c->nant : 0);
  // fxkernel.cpp:31,37-38
  d->sampletime = 1.0 / (2.0 * c->bw);
  if (c->is_complex) d->sampletime *= 2;
  const size_t samples = (size_t)(c->numffts + 1) * (size_t)d->L;  // common.cpp:16
  d->station_bytes = c->is_complex ? 4 * samples : samples / 2;
  return GXC_OK;
}

int out_mode_of(int precision) {
  static const int mode_of[4] = {gxc::OUT_F32, gxc::OUT_F16, gxc::OUT_BF16, gxc::OUT_FP8};
  return mode_of[precision];
}

std::vector<int2> make_rowmap(int nant, bool autos) {
  std::vector<int2> rm;
  for (int j = 0; j < nant - 1; j++)
    for (int k = j + 1; k < nant; k++) rm.push_back(make_int2(j, k));
  if (autos)
    for (int a = 0; a < nant; a++) rm.push_back(make_int2(a, a));
  return rm;
}

gxc_status check_device(int *dev, std::string *why) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    *why = "no HIP device available";
    return GXC_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDevice(dev) != hipSuccess || hipGetDeviceProperties(&prop, *dev) != hipSuccess) {
    *why = "cannot query the current HIP device";
    return GXC_ERR_HIP;
  }
  if (!strstr(prop.gcnArchName, "gfx950")) {
    *why = std::string("library is built for gfx950, device is ") + prop.gcnArchName;
    return GXC_ERR_NO_DEVICE;
  }
  return GXC_OK;
}

gxc_status upload_tables(const Derived &d, double bw, int nchan, float2 **twiddle, float **subfreq,
                         float **stepfreq, std::string *why) {
  const int L = d.L, s = d.g.s, nsf = d.g.fracsteps;
  std::vector<float2> tw(L);
  for (int m = 0; m < L; m++) {
    const double ang = -2.0 * kPi * (double)m / (double)L;
    tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  // fxkernel.cpp:171-175 channel-frequency tables, identical expressions.
  std::vector<float> subf(s), stepf(nsf);
  for (int i = 0; i < s; i++) subf[i] = (float)((2 * kPi * i * bw) / (double)nchan);
  for (int i = 0; i < nsf; i++) stepf[i] = (float)((2 * kPi * i * s * bw) / (double)nchan);
  *twiddle = nullptr;
  *subfreq = *stepfreq = nullptr;
  if (hipMalloc((void **)twiddle, L * sizeof(float2)) != hipSuccess ||
      hipMalloc((void **)subfreq, s * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)stepfreq, nsf * sizeof(float)) != hipSuccess) {
    *why = "hipMalloc of the FFT tables failed";
    return GXC_ERR_OUT_OF_MEMORY;
  }
  if (hipMemcpy(*twiddle, tw.data(), L * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(*subfreq, subf.data(), s * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(*stepfreq, stepf.data(), nsf * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    *why = "upload of the FFT tables failed";
    return GXC_ERR_HIP;
  }
  return GXC_OK;
}

}  // namespace host
}  // namespace gxc
