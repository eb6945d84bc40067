// Visibility output through the C-ABI: binary and DiFX-record writers and
// the lag S/N fringe check (lags.hip).
#include "gxc_handle.hpp"

#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "gxc_lags.hpp"

using namespace gxc::api;

extern "C" {

gxc_status gxc_save_visibilities_binary(gxc_handle *h, const char *path) {
  if (!h || !path) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  const int nchan = h->cfg.nchan, nrows = h->d.nrows;
  std::vector<gxc_cf32> v((size_t)nrows * 4 * nchan);
  std::vector<int> counts(nrows);
  GXC_TRY(gxc_get_visibilities(h, v.data(), v.size()));
  GXC_TRY(gxc_get_row_counts(h, counts.data(), counts.size()));
  const std::vector<int2> rows = gxc::host::make_rowmap(h->cfg.nant, h->cfg.autos != 0);
  std::vector<int32_t> st(2 * (size_t)nrows);
  for (int r = 0; r < nrows; r++) {
    st[2 * r] = rows[r].x;
    st[2 * r + 1] = rows[r].y;
  }
  std::ofstream f(path, std::ios::binary);
  if (!f) return set_err(h, GXC_ERR_IO, std::string("cannot open ") + path);
  const int32_t hdr[8] = {1, h->cfg.nant, nchan, nrows, 4, h->cfg.autos ? 1 : 0, h->cfg.numffts, h->cfg.precision};
  const double freq[2] = {h->cfg.bw, h->cfg.lo};
  f.write("GXCVIS01", 8);
  f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
  f.write(reinterpret_cast<const char *>(freq), sizeof(freq));
  f.write(reinterpret_cast<const char *>(st.data()), st.size() * sizeof(int32_t));
  f.write(reinterpret_cast<const char *>(counts.data()), counts.size() * sizeof(int32_t));
  f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(gxc_cf32));
  if (!f) return set_err(h, GXC_ERR_IO, std::string("write failed: ") + path);
  return GXC_OK;
}

gxc_status gxc_save_difx(gxc_handle *h, const char *path, const gxc_difx_meta *meta) {
  if (!h || !path || !meta) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  const int nchan = h->cfg.nchan, nrows = h->d.nrows;
  std::vector<gxc_cf32> v((size_t)nrows * 4 * nchan);
  std::vector<int> counts(nrows);
  GXC_TRY(gxc_get_visibilities(h, v.data(), v.size()));
  GXC_TRY(gxc_get_row_counts(h, counts.data(), counts.size()));
  const std::vector<int2> rows = gxc::host::make_rowmap(h->cfg.nant, h->cfg.autos != 0);
  std::ofstream f(path, std::ios::binary);
  if (!f) return set_err(h, GXC_ERR_IO, std::string("cannot open ") + path);
  for (int r = 0; r < nrows; r++)
    for (int p = 0; p < 4; p++) {
      // DiFX binary record header, 74 bytes, packed
      char rec[74];
      char *q = rec;
      auto put = [&q](const void *x, size_t n) {
        std::memcpy(q, x, n);
        q += n;
      };
      const int32_t sync = (int32_t)0xFF00FF00u, version = 1;
      const int32_t baseline = 256 * (rows[r].x + 1) + (rows[r].y + 1);
      const int32_t mjd = meta->mjd, cfgi = meta->config_index, src = meta->source_index, fq = meta->freq_index;
      const int32_t bin = 0;
      const double sec = meta->seconds;
      const double weight = h->cfg.numffts ? (double)counts[r] / h->cfg.numffts : 0.0;
      const double uvw[3] = {meta->uvw ? meta->uvw[3 * r] : 0.0, meta->uvw ? meta->uvw[3 * r + 1] : 0.0,
                             meta->uvw ? meta->uvw[3 * r + 2] : 0.0};
      const char pol[2] = {meta->pols[p >> 1], meta->pols[p & 1]};  // p = 2l + m
      put(&sync, 4);
      put(&version, 4);
      put(&baseline, 4);
      put(&mjd, 4);
      put(&sec, 8);
      put(&cfgi, 4);
      put(&src, 4);
      put(&fq, 4);
      put(pol, 2);
      put(&bin, 4);
      put(&weight, 8);
      put(uvw, 24);
      f.write(rec, sizeof(rec));
      f.write(reinterpret_cast<const char *>(&v[((size_t)r * 4 + p) * nchan]), (size_t)nchan * sizeof(gxc_cf32));
    }
  if (!f) return set_err(h, GXC_ERR_IO, std::string("write failed: ") + path);
  return GXC_OK;
}

namespace {
gxc_status lag_snr_on(gxc_handle *h, const void *src, hipMemcpyKind kind, int nseries, int nchan, float *snr,
                      int *peak, hipStream_t s) {
  if (nchan < 2 || (nchan & (nchan - 1))) return set_err(h, GXC_ERR_INVALID_ARG, "nchan must be a power of two >= 2");
  if (nseries < 1) return set_err(h, GXC_ERR_INVALID_ARG, "no series");
  const size_t n = (size_t)nseries * nchan;
  float2 *a = nullptr, *b = nullptr;
  float *dsnr = nullptr;
  int *dpeak = nullptr;
  gxc_status st = dmalloc(h, &a, n);
  if (st == GXC_OK) st = dmalloc(h, &b, n);
  if (st == GXC_OK) st = dmalloc(h, &dsnr, (size_t)nseries);
  if (st == GXC_OK) st = dmalloc(h, &dpeak, (size_t)nseries);
  hipError_t e = hipSuccess;
  if (st == GXC_OK) e = hipMemcpyAsync(a, src, n * sizeof(float2), kind, s);
  if (st == GXC_OK && e == hipSuccess) e = gxc::launch_lag_snr(a, b, nchan, nseries, dsnr, dpeak, s);
  if (st == GXC_OK && e == hipSuccess) e = hipMemcpyAsync(snr, dsnr, nseries * sizeof(float), hipMemcpyDeviceToHost, s);
  if (st == GXC_OK && e == hipSuccess) e = hipMemcpyAsync(peak, dpeak, nseries * sizeof(int), hipMemcpyDeviceToHost, s);
  if (st == GXC_OK && e == hipSuccess) e = hipStreamSynchronize(s);
  if (st == GXC_OK && e != hipSuccess) st = set_err(h, GXC_ERR_HIP, std::string("lag S/N: ") + hipGetErrorString(e));
  hipFree(a);
  hipFree(b);
  hipFree(dsnr);
  hipFree(dpeak);
  return st;
}
}  // namespace

gxc_status gxc_lag_snr(gxc_handle *h, float *snr, int *peak, size_t n) {
  if (!h || !snr || !peak) return set_err(h, GXC_ERR_INVALID_ARG, "null argument");
  const int nseries = h->d.nrows * 4;
  if (n < (size_t)nseries) return set_err(h, GXC_ERR_INVALID_ARG, "output arrays too small");
  GXC_HIP(h, hipStreamSynchronize(h->stream));
  return lag_snr_on(h, h->d_vis, hipMemcpyDeviceToDevice, nseries, h->cfg.nchan, snr, peak, h->stream);
}

gxc_status gxc_lag_snr_series(const gxc_cf32 *vis, int nseries, int nchan, float *snr, int *peak) {
  if (!vis || !snr || !peak) return set_err(nullptr, GXC_ERR_INVALID_ARG, "null argument");
  int dev = 0;
  GXC_TRY(check_device(nullptr, &dev));
  return lag_snr_on(nullptr, vis, hipMemcpyHostToDevice, nseries, nchan, snr, peak, nullptr);
}

}  // extern "C"
