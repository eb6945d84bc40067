// F-stage on the matrix cores for the reduced-precision forms (FP16, BF16
// and FP8 spectra), real 2-bit input, FFT lengths 2048..8192.
//
// The FP32 F-stage issues about one VALU wave-instruction per sample and is
// bound by that.  Here the radix passes run as MFMA products: a radix-R
// pass over columns b is Y[r'][b] = sum_r W_R^(r r') X[r][b], a real 16x16x32
// f16 MFMA per 16 columns with the complex data interleaved along K (re, im
// of 16 points) and the DFT matrix split into two f16 terms (hi + lo, so its
// own rounding does not count), accumulated in FP32:
//   Y_re = [cos | sin] . X,  Y_im = [-sin | cos] . X.
// L = 16 * 16 * RL, three Stockham passes: DFT-16 (one butterfly per MFMA
// column), DFT-16 again, then DFT-RL pruned to the RL/2 outputs that hold
// bins below nchan (fxkernel keeps bins [0, nchan) of the 2*nchan-point FFT),
// 16/(RL/2) butterflies per MFMA column through a block-diagonal matrix.
//
// The data enter every MFMA as f16 and leave it as FP32.  Between passes
// the FP32 outputs are multiplied by the next pass's twiddle in FP32 and
// rounded to f16 once; the twiddles of a lane's points do not depend on the
// unit, so each lane keeps them in registers (as f16 pairs, widened exactly
// in the multiply) for the whole launch.  The fringe rotation is FP32 with
// one rounding of the rotated sample.  The last pass's outputs keep FP32 for
// the fractional-sample rotation and the scaled store.
//
// A workgroup of 256 threads (512 at L = 8192) loops over units
// (persistent grid).  Below L = 8192 a unit is one (station, FFT): the
// rotator and fractional-sample tables are built once and serve both
// polarisations, which run one after the other.  At L = 8192 a unit is one
// polarisation: both in one unit took 25.1 ms for the F-stage of 16
// stations, 4096 channels, 1 s, one per unit 15.4 ms.  The passes go
// through two LDS buffers one 16-column tile at a time.
//
// Everything else follows the FP32 F-stage: the FP64 delay split and the
// rotator and fractional-sample tables (fill_tables), the 2-bit levels of
// fxkernel.cpp:15-28, and store_bin with the per-(input, channel) scales.
#include "fstage_common.hpp"

namespace gxc {
namespace fs {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mma(const uint4 &a, const uint4 &b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8, a), __builtin_bit_cast(half8, b), c, 0, 0,
                                                0);
}

// (re, im) rounded to f16 (round to nearest even, one instruction)
__device__ __forceinline__ h2 toh2(float re, float im) { return __builtin_convertvector((float2v){re, im}, h2); }
__device__ __forceinline__ uint32_t hbits(h2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ h2 hval(uint32_t v) { return __builtin_bit_cast(h2, v); }
// y * w in FP32 with w stored as an f16 pair, rounded to f16 once
__device__ __forceinline__ uint32_t twmul(float yr, float yi, uint32_t w) {
  const float2v t = __builtin_convertvector(hval(w), float2v);
  return hbits(toh2(yr * t.x - yi * t.y, yr * t.y + yi * t.x));
}

template <int L>
struct MG {
  static constexpr int RL = L / 256;                // radix of the last pass
  static constexpr int O = RL / 2;                  // outputs kept per last-pass butterfly
  static constexpr int S = 16 / O;                  // last-pass butterflies per MFMA column
  static constexpr int LQ = L / 16;                 // butterflies of passes 0 and 1
  static constexpr int NTHR = L >= 8192 ? 512 : 256;
  static constexpr int NW = NTHR / 64;              // waves per workgroup
  static constexpr int K01 = L / 256 / NW;          // 16-column tiles per wave, passes 0 and 1
  static constexpr int K2 = O / NW;                 // 16-column tiles per wave, last pass
  static constexpr int PS = L + L / 8;              // padded words per LDS buffer
  static constexpr int PPU = L >= 8192 ? 1 : 2;     // polarisations per unit
  using TG = TabGeom<L, false>;
  static_assert(L >= 2048 && L <= 8192, "MFMA F-stage covers L = 2048..8192");
  static_assert(K01 >= 1 && K2 >= 1 && K01 * NW * 256 == L && K2 * NW == O, "whole tiles per wave");
};

// LDS word of complex point i: 4 words of padding per 32, keeping groups of
// 4 (one 16-byte store) contiguous and aligned.
__device__ __forceinline__ int pad(int i) { return i + ((i >> 5) << 2); }

// W_L^m = exp(-2 pi i m / L) as packed f16
template <int L>
__device__ __forceinline__ uint32_t twh(int m) {
  float s, c;
  sincospif(2.0f * (float)(m % L) / (float)L, &s, &c);
  return hbits(toh2(c, -s));
}

// One lane's operand of a DFT matrix: row `row` of the 16-row tile, complex
// K index kc = 16 ks + 4 g + j (j = 0..3) as (re, im) half pairs.  Output
// row -> butterfly so = row / O, output rp = row % O; input kc -> butterfly
// kc / R, element kc % R; zero off the diagonal blocks.  hi = f16(w),
// lo = f16(w - hi).
template <int R, int O>
__device__ void dft_operand(int row, int g, int ks, uint4 &re_hi, uint4 &re_lo, uint4 &im_hi, uint4 &im_lo) {
  const int so = row / O, rp = row % O;
  uint32_t w[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int kc = 16 * ks + 4 * g + j;
    float c = 0.f, s = 0.f;
    if (kc / R == so) sincospif(2.0f * (float)(((kc % R) * rp) % R) / (float)R, &s, &c);
    // W = cos - i sin:  Y_re = cos Xr + sin Xi,  Y_im = -sin Xr + cos Xi
    const h2 hr = toh2(c, s), hi = toh2(-s, c);
    const float2v fr = __builtin_convertvector(hr, float2v), fi = __builtin_convertvector(hi, float2v);
    w[0][j] = hbits(hr);
    w[1][j] = hbits(toh2(c - fr.x, s - fr.y));
    w[2][j] = hbits(hi);
    w[3][j] = hbits(toh2(-s - fi.x, c - fi.y));
  }
  re_hi = make_uint4(w[0][0], w[0][1], w[0][2], w[0][3]);
  re_lo = make_uint4(w[1][0], w[1][1], w[1][2], w[1][3]);
  im_hi = make_uint4(w[2][0], w[2][1], w[2][2], w[2][3]);
  im_lo = make_uint4(w[3][0], w[3][1], w[3][2], w[3][3]);
}

// Y = A . B for one 16-column tile: both output components, A as hi + lo.
__device__ __forceinline__ void dft_tile(const uint4 &rh, const uint4 &rl, const uint4 &ih, const uint4 &il,
                                         const uint4 &b, floatx4 &yr, floatx4 &yi) {
  yr = mma(rh, b, yr);
  yr = mma(rl, b, yr);
  yi = mma(ih, b, yi);
  yi = mma(il, b, yi);
}

template <int L, int OUT>
__global__ void __launch_bounds__(MG<L>::NTHR) fstage_mfma_kernel(FParams p, long long nunits) {
  using G = MG<L>;
  using TG = typename G::TG;
  constexpr int NW = G::NW, K01 = G::K01, K2 = G::K2, RL = G::RL, O = G::O, S = G::S, LQ = G::LQ;
  constexpr int NCHAN = TG::NCHAN, S2 = TG::S2, NROT = TG::S2 + TG::NSTEPS;
  __shared__ __align__(16) uint32_t sd[2 * G::PS];  // buffers A and B: complex f16
  __shared__ float2 stab[TG::NTAB];                 // rotator tables (fill_tables)
  __shared__ float slev[4];                         // 2-bit levels (fxkernel.cpp:15-28)

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c16 = lane & 15, g = lane >> 4;

  // Per-launch constants of this lane: DFT matrices and twiddles.
  uint4 d_rh, d_rl, d_ih, d_il;  // DFT-16 of passes 0 and 1
  dft_operand<16, 16>(c16, g, 0, d_rh, d_rl, d_ih, d_il);
  uint4 l_rh[2], l_rl[2], l_ih[2], l_il[2];  // pruned DFT-RL of the last pass, K-steps 0 and 1
#pragma unroll
  for (int ks = 0; ks < 2; ks++) dft_operand<RL, O>(c16, g, ks, l_rh[ks], l_rl[ks], l_ih[ks], l_il[ks]);
  uint32_t tw0[K01][4], tw1[K01][4];
#pragma unroll
  for (int k = 0; k < K01; k++) {
    const int bc = 16 * (wave + NW * k) + c16;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      // pass-0 output q0 = 16 b + r' feeds pass 1 as element q0 / LQ of
      // butterfly q0 % LQ: twiddle W_L^((q0 % 16) (L/256) (q0 / LQ))
      const int q0 = 16 * bc + 4 * g + i;
      tw0[k][i] = twh<L>(((q0 & 15) * (L / 256)) * (q0 / LQ));
      // pass-1 output q1 feeds the last pass as element q1 / 256 of
      // butterfly q1 % 256: twiddle W_L^((q1 % 256) (q1 / 256))
      const int q1 = (bc >> 4) * 256 + (bc & 15) + 16 * (4 * g + i);
      tw1[k][i] = twh<L>((q1 & 255) * (q1 >> 8));
    }
  }
  if (tid < 4) slev[tid] = tid == 0 ? -kHiMag : tid == 1 ? -1.0f : tid == 2 ? 1.0f : kHiMag;

  const float2 *fsub = stab + NROT;
  const float2 *fstep = fsub + TG::S;

  for (long long u = blockIdx.x; u < nunits; u += gridDim.x) {
    const int item = (int)(G::PPU == 2 ? u : u >> 1);
    const int pol0 = G::PPU == 2 ? 0 : (int)(u & 1);
    const int tl = item / p.nant;
    const int st = item % p.nant;
    const int t = p.t0 + tl * p.tstride;
    int offset = 0;
    double frac = 0.0, a = 0.0, b = 0.0;
    // the same for every thread: the branches below are uniform
    const bool valid = split_offset<L>(p, st, t, offset, frac, a, b);
    if (tid == 0 && pol0 == 0 && OUT != OUT_POWER) p.valid[(size_t)st * p.numffts + t] = valid ? 1 : 0;
    __syncthreads();  // the previous unit is done with LDS
    if (!valid) {
      for (int pol = pol0; pol < pol0 + G::PPU; pol++) {
        const size_t row = ((size_t)tl * p.nant + st) * 2 + pol;
        for (int q = tid; q < NCHAN; q += G::NTHR)
          store_bin<OUT, NCHAN>(p, row, 2 * st + pol, q, make_float2(0.f, 0.f));
      }
      continue;
    }
    fill_tables<L, false>(p, frac, a, b, stab, tid, G::NTHR);
    __syncthreads();

    // The polarisations in turn.  Pass 0 of the second writes buffer A,
    // which every wave finished reading before the barrier after pass 1 of
    // the first; its pass 1 writes buffer B after the barrier that follows
    // its pass 0, by which every wave is done with the first's last pass.
#pragma unroll 1
    for (int pol = pol0; pol < pol0 + G::PPU; pol++) {
    const size_t row = ((size_t)tl * p.nant + st) * 2 + pol;

    // Pass 0: DFT-16 over window samples n = b + LQ r, rotated on the way in
    // (rotator sub[n % S2] * step[n / S2], fxkernel.cpp:553-558; S2 divides
    // LQ, so sub is one entry per column).
    gbyte *src = (gbyte *)p.station[st];
    uint32_t *sa = sd, *sb = sd + G::PS;
#pragma unroll
    for (int k = 0; k < K01; k++) {
      const int bc = 16 * (wave + NW * k) + c16;
      const int s0 = offset + bc;
      const int sh = (s0 & 1) * 4 + 2 * pol;
      gbyte *bp = src + (s0 >> 1) + (LQ / 2) * 4 * g;
      const float2 rsub = stab[bc % S2];
      const float2 *stp = stab + S2 + bc / S2 + (LQ / S2) * 4 * g;
      uint32_t w[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const float lev = slev[(bp[(LQ / 2) * j] >> sh) & 3];
        const float2 r = cmul(rsub, stp[(LQ / S2) * j]);
        w[j] = hbits(toh2(lev * r.x, lev * r.y));
      }
      floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
      dft_tile(d_rh, d_rl, d_ih, d_il, make_uint4(w[0], w[1], w[2], w[3]), yr, yi);
      uint32_t o[4];
#pragma unroll
      for (int i = 0; i < 4; i++) o[i] = twmul(yr[i], yi[i], tw0[k][i]);
      *reinterpret_cast<uint4 *>(&sa[pad(16 * bc + 4 * g)]) = make_uint4(o[0], o[1], o[2], o[3]);
    }
    __syncthreads();

    // Pass 1: DFT-16 over in[b + LQ r]; output q1 = (b/16) 256 + b%16 + 16 r'.
#pragma unroll
    for (int k = 0; k < K01; k++) {
      const int bc = 16 * (wave + NW * k) + c16;
      uint32_t w[4];
#pragma unroll
      for (int j = 0; j < 4; j++) w[j] = sa[pad(bc + LQ * (4 * g + j))];
      floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
      dft_tile(d_rh, d_rl, d_ih, d_il, make_uint4(w[0], w[1], w[2], w[3]), yr, yi);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int q1 = (bc >> 4) * 256 + (bc & 15) + 16 * (4 * g + i);
        sb[pad(q1)] = twmul(yr[i], yi[i], tw1[k][i]);
      }
    }
    __syncthreads();

    // Last pass: pruned DFT-RL over in[b + 256 r], b = S C + s for MFMA
    // column C; bin q = b + 256 r' for r' < RL/2, fractional-sample rotation
    // fsub[q % S] * fstep[q / S] (fxkernel.cpp:614-625) in FP32, store.
#pragma unroll
    for (int k = 0; k < K2; k++) {
      const int C = 16 * (wave + NW * k) + c16;
      floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int kc = 16 * ks + 4 * g + j;
          w[j] = sb[pad(S * C + kc / RL + 256 * (kc % RL))];
        }
        dft_tile(l_rh[ks], l_rl[ks], l_ih[ks], l_il[ks], make_uint4(w[0], w[1], w[2], w[3]), yr, yi);
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int rw = 4 * g + i;
        const int q = S * C + rw / O + 256 * (rw % O);
        const float2 fr = cmul(fsub[q % TG::S], fstep[q / TG::S]);
        store_bin<OUT, NCHAN>(p, row, 2 * st + pol, q, cmul(fr, make_float2(yr[i], yi[i])));
      }
    }
    }  // pol
  }
}

template <int L, int OUT>
hipError_t launch_l(const FParams &p, hipStream_t s) {
  const long long nunits = (long long)p.nant * p.nt * (MG<L>::PPU == 2 ? 1 : 2);
  if (nunits <= 0) return hipSuccess;
  const void *k = reinterpret_cast<const void *>(&fstage_mfma_kernel<L, OUT>);
  const long long grid = persistent_grid(k, MG<L>::NTHR, 0, nunits);
  hipLaunchKernelGGL((fstage_mfma_kernel<L, OUT>), dim3((unsigned)grid), dim3(MG<L>::NTHR), 0, s, p, nunits);
  return hipGetLastError();
}

template <int L>
hipError_t launch_out(const FParams &p, hipStream_t s) {
  switch (p.out_mode) {
    case OUT_F16: return launch_l<L, OUT_F16>(p, s);
    case OUT_BF16: return launch_l<L, OUT_BF16>(p, s);
    case OUT_FP8: return launch_l<L, OUT_FP8>(p, s);
    case OUT_POWER: return launch_l<L, OUT_POWER>(p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace
}  // namespace fs

bool fstage_mfma_supported(int L, bool cplx) { return !cplx && L >= 2048 && L <= 8192; }

hipError_t launch_fstage_mfma(const FParams &p, int L, hipStream_t s) {
  switch (L) {
    case 2048: return fs::launch_out<2048>(p, s);
    case 4096: return fs::launch_out<4096>(p, s);
    case 8192: return fs::launch_out<8192>(p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gxc
