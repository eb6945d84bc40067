// Per-subband F-stage launch: one workgroup per group of items (station,
// FFT).  FFT lengths 64..8192 run fstage_item (fstage_item.hpp) at the
// default workgroup size here, at half or double that size in
// fstage_half.hip / fstage_double.hip; longer FFTs go to fstage_big.hip.
#include "fstage_launch.hpp"

namespace gxc {

hipError_t launch_fstage_half(const FParams &p, int L, bool cplx, hipStream_t s);
hipError_t launch_fstage_double(const FParams &p, int L, bool cplx, hipStream_t s);

bool fstage_supported(int L, bool cplx) {
  (void)cplx;
  return L >= 64 && L <= 32768 && (L & (L - 1)) == 0;
}

int fstage_blocks(int L, bool cplx, int *out) {
  if (L > 8192) {
    out[0] = L / 32;  // one workgroup per (station, FFT, pol) in fstage_big.hip
    return 1;
  }
  int n = 0;
  for (int mult : {2, 1, 4}) {  // default first
    const int b = fs::fblock_for(L, cplx, mult);
    if (b) out[n++] = b;
  }
  // the register kernel of the long FFTs (fstage_big), L/32 threads
  if (L >= 2048) out[n++] = L / 32;
  return n;
}

hipError_t launch_fstage(const FParams &p, int L, bool cplx, int block, hipStream_t s) {
  if (block == kFBlockMfma)
    return fstage_mfma_supported(L, cplx, p.out_mode == OUT_F32) ? launch_fstage_mfma(p, L, s) : hipErrorInvalidValue;
  if (L > 8192 || (L >= 2048 && block == L / 32)) return launch_fstage_big(p, L, cplx, s);
  if (block == 0 || block == fs::fblock_for(L, cplx, 2)) return fs::launch_fstage_mult<2>(p, L, cplx, s);
  if (block == fs::fblock_for(L, cplx, 1)) return launch_fstage_half(p, L, cplx, s);
  if (block == fs::fblock_for(L, cplx, 4)) return launch_fstage_double(p, L, cplx, s);
  return hipErrorInvalidValue;
}

}  // namespace gxc
