// Streamed subintegrations through the C-ABI: pinned host buffers, a copy
// stream filling two device slots while the handle's stream correlates, and
// a device sum of the visibilities.
#include "gxc_handle.hpp"

using namespace gxc::api;

extern "C" {

void *gxc_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void gxc_host_free(void *p) {
  if (p) hipHostFree(p);
}

gxc_status gxc_integrate_begin(gxc_handle *h) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  if (h->vdif) {  // submitted input has no frame flags
    leave_vdif(h);
    h->input_set = false;
  }
  if (!h->cstream) GXC_HIP(h, hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking));
  const size_t nvis = (size_t)h->d.nrows * 4 * h->cfg.nchan;
  if (!h->d_acc) GXC_TRY(dmalloc(h, &h->d_acc, nvis));
  for (auto &sl : h->slot) {
    if (sl.d_in) continue;
    const int nant = h->cfg.nant;
    GXC_TRY(dmalloc(h, &sl.d_in, h->d.station_bytes * (size_t)nant));
    GXC_TRY(dmalloc(h, &sl.d_ptrs, (size_t)nant));
    GXC_TRY(dmalloc(h, &sl.d_delays, (size_t)nant * 4));
    GXC_HIP(h, hipHostMalloc((void **)&sl.h_delays, (size_t)nant * 4 * sizeof(double), hipHostMallocDefault));
    GXC_HIP(h, hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    GXC_HIP(h, hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    std::vector<const uint8_t *> ptrs(nant);
    for (int a = 0; a < nant; a++) ptrs[a] = sl.d_in + (size_t)a * h->d.station_bytes;
    GXC_HIP(h, hipMemcpy(sl.d_ptrs, ptrs.data(), nant * sizeof(void *), hipMemcpyHostToDevice));
  }
  GXC_HIP(h, gxc::launch_fill(reinterpret_cast<float *>(h->d_acc), 2 * nvis, 0.f, h->stream));
  for (auto &sl : h->slot) sl.used = false;
  h->nsubmitted = 0;
  h->integrating = true;
  return GXC_OK;
}

gxc_status gxc_submit(gxc_handle *h, const uint8_t *const *per_station, const double *const *poly3,
                      const double *fileoffsets) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (!h->integrating) return set_err(h, GXC_ERR_STATE, "gxc_integrate_begin first");
  if (!per_station || !poly3 || !fileoffsets) return set_err(h, GXC_ERR_INVALID_ARG, "null input or delays");
  const int nant = h->cfg.nant;
  for (int a = 0; a < nant; a++)
    if (!per_station[a] || !poly3[a]) return set_err(h, GXC_ERR_INVALID_ARG, "null station buffer or polynomial");
  auto &sl = h->slot[h->nsubmitted & 1];
  h->marks.clear();
  // the slot's previous copy (subintegration k-2, already waited on by the
  // previous call) has finished reading the delay staging before it is rewritten
  if (sl.used) GXC_HIP(h, hipEventSynchronize(sl.copied));
  for (int a = 0; a < nant; a++) {
    for (int i = 0; i < 3; i++) sl.h_delays[4 * a + i] = poly3[a][i];
    sl.h_delays[4 * a + 3] = fileoffsets[a];
  }
  // copy stream: after subintegration k-2 released the slot
  if (sl.used) GXC_HIP(h, hipStreamWaitEvent(h->cstream, sl.done, 0));
  for (int a = 0; a < nant; a++)
    GXC_HIP(h, hipMemcpyAsync(sl.d_in + (size_t)a * h->d.station_bytes, per_station[a], h->d.station_bytes,
                              hipMemcpyHostToDevice, h->cstream));
  GXC_HIP(h, hipMemcpyAsync(sl.d_delays, sl.h_delays, (size_t)nant * 4 * sizeof(double), hipMemcpyHostToDevice,
                            h->cstream));
  GXC_HIP(h, hipEventRecord(sl.copied, h->cstream));
  // compute stream: correlate from the slot, then add into the sum
  GXC_HIP(h, hipStreamWaitEvent(h->stream, sl.copied, 0));
  const uint8_t **ptrs = h->d_station_ptrs;
  double *dl = h->d_delays;
  h->d_station_ptrs = sl.d_ptrs;
  h->d_delays = sl.d_delays;
  gxc_status st = enqueue_device_work(h);
  h->d_station_ptrs = ptrs;
  h->d_delays = dl;
  GXC_TRY(st);
  const size_t nvis = (size_t)h->d.nrows * 4 * h->cfg.nchan;
  GXC_HIP(h, gxc::launch_add(reinterpret_cast<float *>(h->d_acc), reinterpret_cast<const float *>(h->d_vis),
                             2 * nvis, h->stream));
  GXC_HIP(h, hipEventRecord(sl.done, h->stream));
  sl.used = true;
  // the previous subintegration's host buffers have been read once this
  // returns, so a caller may refill them while this one copies
  if (h->nsubmitted > 0) GXC_HIP(h, hipEventSynchronize(h->slot[(h->nsubmitted - 1) & 1].copied));
  h->nsubmitted++;
  return GXC_OK;
}

gxc_status gxc_integrate_end(gxc_handle *h, gxc_cf32 *out, size_t nelem, int *nsubints) {
  if (!h) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null handle");
  if (!h->integrating) return set_err(h, GXC_ERR_STATE, "no integration in progress");
  const size_t nvis = (size_t)h->d.nrows * 4 * h->cfg.nchan;
  if (out && nelem < nvis) return set_err(h, GXC_ERR_INVALID_ARG, "output buffer too small");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  if (out) GXC_HIP(h, hipMemcpy(out, h->d_acc, nvis * sizeof(gxc_cf32), hipMemcpyDeviceToHost));
  if (nsubints) *nsubints = (int)h->nsubmitted;
  h->integrating = false;
  return GXC_OK;
}

}  // extern "C"
