// FxKernel-compatible C++ facade over the gxcorr C-ABI.
//
// Same class name, constructor and method signatures as
// fxkernel/src/fxkernel.h:18-64, so code written against fxkernel
// (testfxkernel.cpp:81-108, bench_fxkernel.cpp:233-253) compiles unchanged
// apart from the include.  Errors throw std::runtime_error; define
// GXCORR_EXIT_ON_ERROR to print and exit(1) as the reference does
// (fxkernel.cpp:47-63,575-578).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../gxcorr.h"

typedef uint8_t u8;
typedef gxc_cf32 cf32;
typedef float f32;
typedef double f64;

class FxKernel {
 public:
  FxKernel(int nant, int nchan, int nfft, int nbit, double localosc, double bw, bool is_complex,
           int precision = GXC_PREC_FP32, bool autos = false) {
    gxc_config c;
    gxc_config_init(&c);
    c.nant = nant;
    c.nchan = nchan;
    c.numffts = nfft;
    c.nbit = nbit;
    c.lo = localosc;
    c.bw = bw;
    c.is_complex = is_complex ? 1 : 0;
    c.precision = precision;
    c.autos = autos ? 1 : 0;
    check(gxc_create(&c, &h_), nullptr);
  }
  ~FxKernel() { gxc_destroy(h_); }
  FxKernel(const FxKernel &) = delete;
  FxKernel &operator=(const FxKernel &) = delete;
  FxKernel(FxKernel &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }

  void setInputData(u8 **idata) { check(gxc_set_input(h_, idata), h_); }
  void setDelays(double **d, double *f) { check(gxc_set_delays(h_, d, f), h_); }
  void process() { check(gxc_process(h_), h_); }
  void accumulate(cf32 ***odata) { check(gxc_accumulate(h_, odata), h_); }
  void saveVisibilities(const char *outfile) { check(gxc_save_visibilities(h_, outfile), h_); }

  // Extensions beyond fxkernel.h, thin wrappers of the C-ABI calls of the
  // same names (see gxcorr.h for their contracts).
  void setInputDevice(const void *const *dev) { check(gxc_set_input_device(h_, dev), h_); }
  gxc_vdif_stats setInputVDIF(int station, const void *d_frames, size_t nbytes, const gxc_vdif_spec &spec) {
    gxc_vdif_stats st;
    check(gxc_set_input_vdif(h_, station, d_frames, nbytes, &spec, &st), h_);
    return st;
  }
  void setPolConversion(const cf32 *gains) { check(gxc_set_pol_conversion(h_, gains), h_); }
  void integrateBegin() { check(gxc_integrate_begin(h_), h_); }
  void submit(u8 **idata, double **d, double *f) { check(gxc_submit(h_, idata, d, f), h_); }
  int integrateEnd(cf32 *out, size_t nelem) {
    int n = 0;
    check(gxc_integrate_end(h_, out, nelem, &n), h_);
    return n;
  }
  void saveVisibilitiesBinary(const char *outfile) { check(gxc_save_visibilities_binary(h_, outfile), h_); }
  void saveDiFX(const char *outfile, const gxc_difx_meta &meta) { check(gxc_save_difx(h_, outfile, &meta), h_); }
  void lagSNR(float *snr, int *peak, size_t n) { check(gxc_lag_snr(h_, snr, peak, n), h_); }

  gxc_handle *handle() { return h_; }

 private:
  static void check(gxc_status s, const gxc_handle *h) {
    if (s == GXC_OK) return;
    std::string msg = std::string("gxcorr: ") + gxc_status_string(s) + ": " + gxc_last_error(h);
#ifdef GXCORR_EXIT_ON_ERROR
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(1);
#else
    throw std::runtime_error(msg);
#endif
  }
  gxc_handle *h_ = nullptr;
};
