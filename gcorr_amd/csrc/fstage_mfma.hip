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
// The twiddle of the next pass is applied to each pass's FP32 outputs before
// they are rounded to f16 for LDS, so the data enter every MFMA as f16 and
// leave it as FP32: three f16 roundings of the data in all (rotated input,
// after pass 0, after pass 1) besides the final requantisation.
//
// Everything else follows the FP32 F-stage: the FP64 delay split and the
// rotator and fractional-sample tables (fill_tables), the LDS LUT for the
// 2-bit levels, and store_bin with the per-(input, channel) scales.  A
// workgroup holds one (station, FFT) with both polarisations (L <= 4096) or
// one polarisation (L = 8192), and loops over units (persistent grid) so the
// DFT matrices are built once per workgroup.  The passes go through two LDS
// buffers (pass 0 -> A, pass 1 A -> B, last pass from B), one 16-column tile
// at a time, so no pass holds more than one tile's operands in registers.
#include "fstage_common.hpp"

namespace gxc {
namespace fs {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mma(const uint4 &a, const uint4 &b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8, a), __builtin_bit_cast(half8, b), c, 0, 0,
                                                0);
}

template <int L>
struct MG {
  static constexpr int RL = L / 256;             // radix of the last pass
  static constexpr int O = RL / 2;               // outputs kept per last-pass butterfly
  static constexpr int S = 16 / O;               // last-pass butterflies per MFMA column
  static constexpr int NPW = L <= 4096 ? 2 : 1;  // polarisations per workgroup
  static constexpr int LQ = L / 16;              // butterflies of passes 0 and 1
  static constexpr int NT01 = L / 256;           // 16-column tiles of passes 0 and 1, per polarisation
  static constexpr int NT2 = O;                  // 16-column tiles of the last pass, per polarisation
  static constexpr int PS = L + L / 8;           // padded words per polarisation in LDS
  static constexpr int BUF = NPW * PS;           // words per LDS buffer
  using TG = TabGeom<L, false>;
  static_assert(L >= 2048 && L <= 8192, "MFMA F-stage covers L = 2048..8192");
  static_assert(NT01 % 4 == 0 && NT2 % 4 == 0, "whole tiles per wave");
};

// LDS word of complex point i: 4 words of padding per 32, keeping groups of
// 4 (one 16-byte store) contiguous and aligned.
__device__ __forceinline__ int pad(int i) { return i + ((i >> 5) << 2); }

__device__ __forceinline__ uint32_t f16pair(float lo, float hi) { return pack_f16(lo, hi); }

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
    const float v[4] = {c, s, -s, c};
    _Float16 h[4];
    float l[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      h[e] = (_Float16)v[e];
      l[e] = v[e] - (float)h[e];
    }
    w[0][j] = (uint32_t)__builtin_bit_cast(uint16_t, h[0]) | ((uint32_t)__builtin_bit_cast(uint16_t, h[1]) << 16);
    w[1][j] = f16pair(l[0], l[1]);
    w[2][j] = (uint32_t)__builtin_bit_cast(uint16_t, h[2]) | ((uint32_t)__builtin_bit_cast(uint16_t, h[3]) << 16);
    w[3][j] = f16pair(l[2], l[3]);
  }
  re_hi = make_uint4(w[0][0], w[0][1], w[0][2], w[0][3]);
  re_lo = make_uint4(w[1][0], w[1][1], w[1][2], w[1][3]);
  im_hi = make_uint4(w[2][0], w[2][1], w[2][2], w[2][3]);
  im_lo = make_uint4(w[3][0], w[3][1], w[3][2], w[3][3]);
}

template <int L, int OUT>
__global__ void __launch_bounds__(256) fstage_mfma_kernel(FParams p, long long nunits) {
  using G = MG<L>;
  using TG = typename G::TG;
  constexpr int NPW = G::NPW, RL = G::RL, O = G::O, S = G::S, LQ = G::LQ, PS = G::PS;
  constexpr int NCHAN = TG::NCHAN;
  __shared__ __align__(16) uint32_t sd[2 * G::BUF];  // buffers A and B
  __shared__ float2 stab[TG::NTAB];
  __shared__ float slut[4];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c16 = lane & 15, g = lane >> 4;

  uint4 d_rh, d_rl, d_ih, d_il;  // DFT-16 of passes 0 and 1
  dft_operand<16, 16>(c16, g, 0, d_rh, d_rl, d_ih, d_il);
  uint4 l_rh[2], l_rl[2], l_ih[2], l_il[2];  // pruned DFT-RL of the last pass, K-steps 0 and 1
#pragma unroll
  for (int ks = 0; ks < 2; ks++) dft_operand<RL, O>(c16, g, ks, l_rh[ks], l_rl[ks], l_ih[ks], l_il[ks]);
  // 2-bit levels by code (fxkernel.cpp:15-28)
  if (tid < 4) slut[tid] = tid == 0 ? -kHiMag : tid == 1 ? -1.0f : tid == 2 ? 1.0f : kHiMag;

  const float2 *sub = stab;
  const float2 *step = sub + TG::S2;
  const float2 *fsub = step + TG::NSTEPS;
  const float2 *fstep = fsub + TG::S;
  const float2 *__restrict__ tw = p.twiddle;

  for (long long u = blockIdx.x; u < nunits; u += gridDim.x) {
    const int item = NPW == 2 ? (int)u : (int)(u >> 1);
    const int pol0 = NPW == 2 ? 0 : (int)(u & 1);
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
#pragma unroll
      for (int pp = 0; pp < NPW; pp++) {
        const int pol = pol0 + pp;
        const size_t row = ((size_t)tl * p.nant + st) * 2 + pol;
        for (int q = tid; q < NCHAN; q += 256) store_bin<OUT, NCHAN>(p, row, 2 * st + pol, q, make_float2(0.f, 0.f));
      }
      continue;
    }
    fill_tables<L, false>(p, frac, a, b, stab, tid, 256);
    __syncthreads();

    // Pass 0: DFT-16 over window samples n = b + LQ r, rotated on the way in
    // (rotator sub[n % S2] * step[n / S2], fxkernel.cpp:553-558); output
    // q0 = 16 b + r', times the pass-1 twiddle W_L^((q0 % 16) (L/256) (q0 / LQ)).
    gbyte *src = (gbyte *)p.station[st];
    uint32_t *sa = sd, *sb = sd + G::BUF;
#pragma unroll 1
    for (int k = 0; k < G::NT01 / 4; k++) {
      const int bc = 16 * (wave + 4 * k) + c16;
      uint32_t w[NPW][4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int n = bc + LQ * (4 * g + j);
        const int sidx = offset + n;
        const int byte = src[sidx >> 1];
        const int sh = (sidx & 1) * 4;
        const float2 r = cmul(sub[n % TG::S2], step[n / TG::S2]);
#pragma unroll
        for (int pp = 0; pp < NPW; pp++) {
          const float x = slut[(byte >> (sh + 2 * (pol0 + pp))) & 3];
          w[pp][j] = pack_f16(x * r.x, x * r.y);
        }
      }
#pragma unroll
      for (int pp = 0; pp < NPW; pp++) {
        const uint4 bop = make_uint4(w[pp][0], w[pp][1], w[pp][2], w[pp][3]);
        floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
        yr = mma(d_rh, bop, yr);
        yr = mma(d_rl, bop, yr);
        yi = mma(d_ih, bop, yi);
        yi = mma(d_il, bop, yi);
        uint32_t o[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int q0 = 16 * bc + 4 * g + i;
          const int m = ((q0 & 15) * (L / 256)) * (q0 / LQ);
          const float2 v = cmul(make_float2(yr[i], yi[i]), tw[m]);
          o[i] = pack_f16(v.x, v.y);
        }
        *reinterpret_cast<uint4 *>(&sa[pp * PS + pad(16 * bc + 4 * g)]) = make_uint4(o[0], o[1], o[2], o[3]);
      }
    }
    __syncthreads();

    // Pass 1: DFT-16 over in[b + LQ r]; output q1 = (b/16) 256 + b%16 + 16 r',
    // times the last-pass twiddle W_L^((q1 % 256) (q1 / 256)).
#pragma unroll 1
    for (int k = 0; k < G::NT01 / 4; k++) {
      const int bc = 16 * (wave + 4 * k) + c16;
#pragma unroll
      for (int pp = 0; pp < NPW; pp++) {
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) w[j] = sa[pp * PS + pad(bc + LQ * (4 * g + j))];
        const uint4 bop = make_uint4(w[0], w[1], w[2], w[3]);
        floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
        yr = mma(d_rh, bop, yr);
        yr = mma(d_rl, bop, yr);
        yi = mma(d_ih, bop, yi);
        yi = mma(d_il, bop, yi);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int q1 = (bc >> 4) * 256 + (bc & 15) + 16 * (4 * g + i);
          const float2 v = cmul(make_float2(yr[i], yi[i]), tw[(q1 & 255) * (q1 >> 8)]);
          sb[pp * PS + pad(q1)] = pack_f16(v.x, v.y);
        }
      }
    }
    __syncthreads();

    // Last pass: pruned DFT-RL over in[b + 256 r], b = S C + s for MFMA
    // column C; bin q = b + 256 r' for r' < RL/2, fractional-sample rotation
    // fsub[q % S] * fstep[q / S] (fxkernel.cpp:614-625), store.
#pragma unroll 1
    for (int k = 0; k < G::NT2 / 4; k++) {
      const int C = 16 * (wave + 4 * k) + c16;
#pragma unroll
      for (int pp = 0; pp < NPW; pp++) {
        const int pol = pol0 + pp;
        const size_t row = ((size_t)tl * p.nant + st) * 2 + pol;
        floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
          uint32_t w[4];
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int kc = 16 * ks + 4 * g + j;
            w[j] = sb[pp * PS + pad(S * C + kc / RL + 256 * (kc % RL))];
          }
          const uint4 bop = make_uint4(w[0], w[1], w[2], w[3]);
          yr = mma(l_rh[ks], bop, yr);
          yr = mma(l_rl[ks], bop, yr);
          yi = mma(l_ih[ks], bop, yi);
          yi = mma(l_il[ks], bop, yi);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int rw = 4 * g + i;
          const int q = S * C + rw / O + 256 * (rw % O);
          const float2 fr = cmul(fsub[q % TG::S], fstep[q / TG::S]);
          store_bin<OUT, NCHAN>(p, row, 2 * st + pol, q, cmul(fr, make_float2(yr[i], yi[i])));
        }
      }
    }
  }
}

template <int L, int OUT>
hipError_t launch_l(const FParams &p, hipStream_t s) {
  constexpr int NPW = MG<L>::NPW;
  const long long nunits = (long long)p.nant * p.nt * (NPW == 2 ? 1 : 2);
  if (nunits <= 0) return hipSuccess;
  const void *k = reinterpret_cast<const void *>(&fstage_mfma_kernel<L, OUT>);
  const long long grid = persistent_grid(k, 256, 0, nunits);
  hipLaunchKernelGGL((fstage_mfma_kernel<L, OUT>), dim3((unsigned)grid), dim3(256), 0, s, p, nunits);
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
