// Internal to the C-ABI: the handle behind include/gxcorr.h and the helpers
// that its translation units share (gxc_api.cpp: lifetime, input, delays,
// process, results; gxc_vdif_api.cpp: VDIF ingest; gxc_convert.cpp:
// polarisation conversion and requantisers; gxc_stream.cpp: streamed
// subintegrations; gxc_output.cpp: binary/DiFX output and lag S/N;
// gxc_tune.cpp: the autotuner).
#pragma once

#include "../../include/gxcorr.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "gxc_host.hpp"
#include "gxc_internal.hpp"
#include "gxc_vdif_state.hpp"
#include "gxc_xengine.hpp"

struct gxc_handle {
  gxc_config cfg{};
  gxc::host::Derived d;
  int device = 0;
  hipStream_t stream = nullptr;   // F-stage, input copies, final join
  hipStream_t xstream = nullptr;  // X-stage when overlapped
  std::string err;

  // input
  bool input_set = false, input_device = false;
  std::vector<const uint8_t *> host_in;
  std::vector<const void *> dev_in;
  uint8_t *d_input_owned = nullptr;
  const uint8_t **d_station_ptrs = nullptr;

  // VDIF ingest (gxc_set_input_vdif): payloads go into d_input_owned, and
  // the per-station frame flags make FFTs over missing frames invalid
  bool vdif = false;  // input comes from gxc_set_input_vdif
  gxc::host::VdifState vs;

  // polarisation conversion of the visibilities (gxc_set_pol_conversion)
  bool polconv = false;
  float2 *d_polgains = nullptr;  // [nant][2][nchan]

  // delays
  bool delays_set = false;
  std::vector<double> h_delays;  // [nant][4]
  double *d_delays = nullptr;

  // constant tables
  float2 *d_twiddle = nullptr;
  float *d_subfreq = nullptr, *d_stepfreq = nullptr;

  // spectra, X-stage and output
  int out_mode = gxc::OUT_F32;
  int chunk = 0, nbuf = 1, splits = 1;
  int fblock = 0, xtile = 0;  // launch shapes (0 = default)
  int partial_cap = 0;        // splits the partial buffer holds
  void *d_spec[2] = {nullptr, nullptr};  // [chunk][M][nchan] packed complex
  float *d_partial = nullptr;             // [splits][nchan][2][M][M]
  float *d_comp = nullptr;                // its compensation across chunks
                                          // (subintegrations of several chunks)
  float *d_scales = nullptr;              // [M][nchan] (reduced precision)
  float *d_power = nullptr;               // [kPowerRows][M][nchan]
  uint8_t *d_valid = nullptr;             // [nant][numffts]
  int *d_counts = nullptr;
  int2 *d_rowmap = nullptr;
  float2 *d_vis = nullptr;

  // ordering between the two streams
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_f[2] = {nullptr, nullptr}, ev_x[2] = {nullptr, nullptr};

  // graph replay
  hipGraphExec_t graph_exec = nullptr;

  // streaming integration (gxc_integrate_begin / gxc_submit / gxc_integrate_end):
  // two input slots, so the copy of subintegration k+1 (copy stream) runs
  // while subintegration k is correlated
  hipStream_t cstream = nullptr;
  struct Slot {
    uint8_t *d_in = nullptr;           // [nant][station_bytes]
    const uint8_t **d_ptrs = nullptr;  // [nant] into d_in
    double *d_delays = nullptr;        // [nant][4]
    double *h_delays = nullptr;        // pinned staging of the same
    hipEvent_t copied = nullptr, done = nullptr;
    bool used = false;
  } slot[2];
  float2 *d_acc = nullptr;  // sum of the subintegrations' visibilities
  long long nsubmitted = 0;
  bool integrating = false;

  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> tev;
  struct Mark {
    int kind;  // 0 F start, 1 F end, 2 X start, 3 X end, 4 finalise start, 5 finalise end, 6 H2D end
    size_t ev;
  };
  std::vector<Mark> marks;
  gxc_timings timings{};
};

namespace gxc {
namespace api {

// Record msg as the handle's (and the calling thread's) last error.
gxc_status set_err(gxc_handle *h, gxc_status s, const std::string &msg);
// The calling thread's last error (gxc_last_error(NULL)).
const std::string &thread_error();

#define GXC_HIP(h, call)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_err((h), e_ == hipErrorOutOfMemory ? GXC_ERR_OUT_OF_MEMORY : GXC_ERR_HIP,   \
                     std::string(#call) + ": " + hipGetErrorString(e_));                     \
  } while (0)

#define GXC_TRY(expr)            \
  do {                           \
    gxc_status s_ = (expr);      \
    if (s_ != GXC_OK) return s_; \
  } while (0)

template <typename T>
gxc_status dmalloc(gxc_handle *h, T **p, size_t nelem) {
  *p = nullptr;
  if (nelem == 0) nelem = 1;
  hipError_t e = hipMalloc((void **)p, nelem * sizeof(T));
  if (e != hipSuccess)
    return set_err(h, GXC_ERR_OUT_OF_MEMORY,
                   "hipMalloc of " + std::to_string(nelem * sizeof(T)) + " bytes failed: " +
                       hipGetErrorString(e));
  return GXC_OK;
}

gxc_status check_device(gxc_handle *h, int *dev);
gxc_status upload_tables(gxc_handle *h);
// Record a profiling event of `kind` on stream s (profiling on only).
void mark(gxc_handle *h, int kind, hipStream_t s);
gxc_status run_fstage(gxc_handle *h, int t0, int nt, int mode, void *out, int tstride, hipStream_t s);
gxc_status run_xstage(gxc_handle *h, const void *spec, int nt, int accumulate, hipStream_t s);
gxc_status run_finalise(gxc_handle *h, hipStream_t s);
// Device work of one subintegration once the input is resident.
gxc_status enqueue_device_work(gxc_handle *h);
gxc_status ensure_owned_input(gxc_handle *h);
gxc_status enqueue_input_copy(gxc_handle *h);
// Input copy, then the device work (replayed as a graph when enabled).
gxc_status enqueue_process(gxc_handle *h);
float elapsed(gxc_handle *h, size_t a, size_t b);
void collect_timings(gxc_handle *h);
void destroy_graph(gxc_handle *h);
// Leave VDIF mode: the frame flags stop applying (a captured graph holds them).
void leave_vdif(gxc_handle *h);

}  // namespace api
}  // namespace gxc
