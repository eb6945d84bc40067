// Row counts for the normalisation of fxkernel.cpp:352,358-367.
//
// The F-stage records per (station, FFT) whether the station's delayed
// window lay inside its buffer (fxkernel.cpp:304-309).  A row (station pair,
// or a station with itself for autocorrelations) counts the FFTs in which
// both stations were valid: fxkernel's baselineCount.
#include <map>
#include <mutex>
#include <set>
#include <tuple>

#include "gxc_internal.hpp"

namespace gxc {
namespace {

// Row `row` on a workgroup of 1024 threads.  Validity flags are 0/1 bytes;
// four FFTs are counted per 32-bit word: the AND of the two stations' words
// holds a 1 byte where both were valid, and the byte sum comes from one
// multiply.  numffts need not be a multiple of 4: the tail is counted bytewise.
__device__ __forceinline__ void count_row(const CountParams &p, int row) {
  const int2 jk = p.rowmap[row];
  const uint8_t *vj = p.valid + (size_t)jk.x * p.numffts;
  const uint8_t *vk = p.valid + (size_t)jk.y * p.numffts;
  int n = 0;
  // rows start at station * numffts: word loads need that offset 4-aligned
  if ((((uintptr_t)vj | (uintptr_t)vk) & 3) == 0) {
    const int nw = p.numffts >> 2;
    const uint32_t *wj = reinterpret_cast<const uint32_t *>(vj), *wk = reinterpret_cast<const uint32_t *>(vk);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) {
      const uint32_t both = wj[i] & wk[i] & 0x01010101u;
      n += (int)((both * 0x01010101u) >> 24);
    }
    for (int t = 4 * nw + threadIdx.x; t < p.numffts; t += blockDim.x) n += (vj[t] & vk[t]) ? 1 : 0;
  } else {
    for (int t = threadIdx.x; t < p.numffts; t += blockDim.x) n += (vj[t] & vk[t]) ? 1 : 0;
  }
  __shared__ int red[1024];
  red[threadIdx.x] = n;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.counts[row] = red[0];
  __syncthreads();  // red is reused by the next row of a persistent launch
}

__global__ void __launch_bounds__(1024) counts_kernel(CountParams p) { count_row(p, blockIdx.x); }

__global__ void __launch_bounds__(1024) counts_multi_kernel(const CountSeg *segs, int nseg, long long total) {
  for (long long u = blockIdx.x; u < total; u += gridDim.x) {
    int k = 0;
    while (k + 1 < nseg && u >= segs[k + 1].unit0) ++k;
    count_row(segs[k].p, (int)(u - segs[k].unit0));
  }
}

}  // namespace

long long persistent_grid(const void *kernel, int block, size_t dyn_lds, long long total) {
  // (device, kernel, block, LDS) -> workgroups resident on the whole device
  static std::mutex mu;
  static std::map<std::tuple<int, const void *, int, size_t>, long long> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  const auto key = std::make_tuple(dev, kernel, block, dyn_lds);
  long long g = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) g = it->second;
  }
  if (g == 0) {
    int ncu = 0, per_cu = 0;
    if (dev < 0 || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 1)
      ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, dyn_lds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    g = (long long)ncu * per_cu;
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = g;
  }
  return total < g ? total : g;
}

hipError_t allow_dynamic_lds(const void *kernel, int bytes) {
  static std::mutex mu;
  static std::set<std::tuple<int, const void *, int>> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const auto key = std::make_tuple(dev, kernel, bytes);
  std::lock_guard<std::mutex> lock(mu);
  if (done.count(key)) return hipSuccess;
  e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess) done.insert(key);
  return e;
}

hipError_t launch_counts_multi(const CountSeg *segs, int nseg, long long total, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  const long long grid = persistent_grid(reinterpret_cast<const void *>(&counts_multi_kernel), 1024, 0, total);
  hipLaunchKernelGGL(counts_multi_kernel, dim3((unsigned)grid), dim3(1024), 0, s, segs, nseg, total);
  return hipGetLastError();
}

hipError_t launch_counts(const CountParams &p, hipStream_t s) {
  if (p.nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(counts_kernel, dim3(p.nrows), dim3(1024), 0, s, p);
  return hipGetLastError();
}

}  // namespace gxc
