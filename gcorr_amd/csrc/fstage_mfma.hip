// F-stage on the matrix cores: real 2-bit input; FP16, BF16 and FP8 spectra
// (and the power pass) at FFT lengths 2048..8192, FP32 spectra at
// 2048..4096.
//
// The FP32 F-stage issues about one VALU wave-instruction per sample and is
// bound by that.  Here the radix passes run as MFMA products: a radix-R
// pass over columns b is Y[r'][b] = sum_r W_R^(r r') X[r][b], a real 16x16x32
// f16 MFMA per 16 columns with the complex data interleaved along K (re, im
// of 16 points), accumulated in FP32:
//   Y_re = [cos | sin] . X,  Y_im = [-sin | cos] . X.
// The reduced-precision forms take one f16 DFT matrix per pass, scaled so
// that its entries are nearly exact (dft_scale); the FP32 form splits the
// matrix into two f16 terms (hi + lo) so that its rounding does not count.
// L = 16 * 16 * RL, three Stockham passes: DFT-16 (one butterfly per MFMA
// column), DFT-16 again, then DFT-RL pruned to the RL/2 outputs that hold
// bins below nchan (fxkernel keeps bins [0, nchan) of the 2*nchan-point FFT),
// 16/(RL/2) butterflies per MFMA column through a block-diagonal matrix.
//
// For the reduced-precision forms the data enter every MFMA as f16 and
// leave it as FP32.  Between passes the FP32 outputs are multiplied by the
// next pass's twiddle (an f16 pair from an LDS table, widened inside the
// FMAs) and rounded to f16 once; at 8192 points pass 0's twiddles are
// folded into its DFT matrices instead.  The FP32 form carries each value as an f16
// pair plus the f16 pair of its remainder (22 significant bits), multiplies
// both into the accumulators (A_hi B_hi + A_lo B_hi + A_hi B_lo) and keeps
// its twiddles in FP32.  The fringe rotation is FP32, rounded to f16 once
// and multiplied by the 2-bit level in packed f16 for the reduced-precision
// forms (exact for the levels +-1), rounded once after the level product in
// FP32 for the FP32 form; the last pass's outputs keep FP32 for the
// fractional-sample rotation and the store.
//
// A workgroup of 256 threads (512 at L = 8192) loops over units (persistent
// grid); a unit is one (station, FFT), whose rotator and fractional-sample
// tables serve both polarisations, run one after the other through one LDS
// buffer.  Registers are held to 128 per lane so that 4 waves share each
// SIMD: the twiddles live in LDS tables, every LDS access is a per-lane base
// plus an immediate offset, and the input bytes and output scales are
// requested ahead of their use.
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
// y * w in FP32 with w stored as an f16 pair, rounded to f16 once.  Each
// product is an FMA with the f16 operand widened inside it (v_fma_mix), so
// w is never converted on its own: re = fma(yr, wr, -(yi wi)),
// im = fma(yr, wi, yi wr), each inner product an fma with +0.  (The same
// product in packed f16, three instructions instead of five, rounds three
// times: at 2048 points it raised the FP16 error against FP64 from 6.8e-4
// to 8.8e-4 for 1 % of the F-stage's time.)
__device__ __forceinline__ uint32_t twmul(float yr, float yi, uint32_t w) {
  const h2 v = hval(w);
  const float wr = (float)v.x, wi = (float)v.y;
  const float a = __builtin_fmaf(yi, wi, 0.0f), b = __builtin_fmaf(yi, wr, 0.0f);
  return hbits(toh2(__builtin_fmaf(yr, wr, -a), __builtin_fmaf(yr, wi, b)));
}

// (re, im) as f16 pair hi and, for the FP32 form, the f16 pair of the
// remainder lo = (re, im) - hi: hi + lo carries 22 significant bits.
template <bool SPLIT>
__device__ __forceinline__ void pack(float re, float im, uint32_t &hi, uint32_t &lo) {
  const h2 h = toh2(re, im);
  hi = hbits(h);
  if constexpr (SPLIT) {
    const float2v b = __builtin_convertvector(h, float2v);
    lo = hbits(toh2(re - b.x, im - b.y));
  }
}

// Twiddle of the passes: an f16 pair (reduced-precision forms) or FP32.
template <bool SPLIT>
using Tw = typename std::conditional<SPLIT, float2, uint32_t>::type;

// y * w, packed as above
template <bool SPLIT>
__device__ __forceinline__ void twist(float yr, float yi, Tw<SPLIT> w, uint32_t &hi, uint32_t &lo) {
  if constexpr (SPLIT) {
    pack<true>(yr * w.x - yi * w.y, yr * w.y + yi * w.x, hi, lo);
  } else {
    hi = twmul(yr, yi, w);
  }
}

// B operand of one tile: four f16 pairs and, for the FP32 form, their remainders.
template <bool SPLIT>
struct Opnd {
  uint4 h;
};
template <>
struct Opnd<true> {
  uint4 h, l;
};

// Delay split of one unit (split_offset's results).
struct Split {
  double frac, a, b;
  int offset, valid;
};
constexpr int kSplitBatch = 16;  // units whose splits one wave computes at once

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
  // LDS word of point i in the buffer pass 0 writes (layout A) and in the
  // one pass 1 writes (layout B): i + a1 (i >> s1) + a2 (i >> s2), chosen so
  // that each 32-lane half of every 4-byte access of the passes falls in 32
  // distinct banks (LDS serves a wave's 4-byte accesses as two halves over
  // 32 banks of 4 bytes) and each 16-lane quarter of the 16-byte pass-0
  // stores covers 64 distinct banks, except for two lanes to a bank in a
  // quarter of the last-pass loads at L = 2048 (no layout of this form does
  // better) and in the pass-1 stores at L = 4096 (the conflict-free layout
  // B (16, 6, 1, 10) there takes 1.8 KB more per workgroup and measured 1 %
  // slower in an 8-subband batch on MI355X).  Terms of layout A are
  // multiples of 4, keeping each lane's 16-byte store aligned.  Every
  // access below is a per-lane base plus an offset that is the same for
  // all lanes and waves, so the offsets are immediates.  (At L = 2048,
  // layout A's second term at bit 9 instead of 10 took the pass-1 loads
  // from two lanes to a bank to none: F-stage 1.92 -> 1.90 ms at 6
  // stations x 1024 channels, FP16.)
  static constexpr int A1 = 4, SA1 = 6, A2 = 16, SA2 = L >= 8192 ? 11 : L == 4096 ? 10 : 9;
  static constexpr int B1 = L == 4096 ? 1 : 16, SB1 = L == 4096 ? 4 : 6;
  static constexpr int B2 = L == 2048 ? 2 : L == 4096 ? 10 : 16, SB2 = L == 4096 ? 7 : 10;
  static constexpr int lay_a(int i) { return i + A1 * (i >> SA1) + A2 * (i >> SA2); }
  static constexpr int lay_b(int i) { return i + B1 * (i >> SB1) + B2 * (i >> SB2); }
  static constexpr int PS = (lay_a(L - 1) > lay_b(L - 1) ? lay_a(L - 1) : lay_b(L - 1)) + 1;  // buffer words
  using TG = TabGeom<L, false>;
  static_assert(L >= 2048 && L <= 8192, "MFMA F-stage covers L = 2048..8192");
  static_assert(K01 >= 1 && K2 >= 1 && K01 * NW * 256 == L && K2 * NW == O, "whole tiles per wave");
};

// Points of the LDS accesses of lane (column c, group g) of wave w:
// pass-0 store of tile k (4 points from here), pass-1 load j of tile k,
// pass-1 store i of tile k, last-pass load j of K-step ks of tile k.
template <int L>
constexpr int pt0w(int w, int c, int g, int k) { return 16 * (16 * (w + MG<L>::NW * k) + c) + 4 * g; }
// the same for tile t of 16 columns
constexpr int pt0t(int t, int c, int g) { return 16 * (16 * t + c) + 4 * g; }
template <int L>
constexpr int pt1r(int w, int c, int g, int k, int j) { return 16 * (w + MG<L>::NW * k) + c + MG<L>::LQ * (4 * g + j); }
template <int L>
constexpr int pt1w(int w, int c, int g, int k, int i) { return 256 * (w + MG<L>::NW * k) + c + 64 * g + 16 * i; }
template <int L>
constexpr int pt2r(int w, int c, int g, int k, int ks, int j) {
  constexpr int RL = MG<L>::RL;
  return MG<L>::S * (16 * (w + MG<L>::NW * k) + c) + (16 * ks + 4 * g + j) / RL + 256 * ((16 * ks + 4 * g + j) % RL);
}

// W_L^m = exp(-2 pi i m / L) times `scale` as packed f16
template <int L>
__device__ __forceinline__ uint32_t twh(int m, float scale) {
  float s, c;
  sincospif(2.0f * (float)(m % L) / (float)L, &s, &c);
  return hbits(toh2(c * scale, -s * scale));
}

// The reduced-precision forms multiply by each DFT-R matrix scaled by a
// constant chosen so that every entry s cos(2 pi k / R), s sin(2 pi k / R)
// lies within a small fraction of an f16 ulp of an f16 value: one f16
// matrix then stands in for the hi + lo pair, and its rounding, the same
// matrix for every FFT, does not add up coherently over the FFTs of a
// subintegration on narrowband signals.  Largest relative entry error: R = 8
// at s = 1.3603517, 1.3e-7; R = 16 at 1.8603300, 1.7e-5; R = 32 at
// 1.6268249, 8.6e-5 (2.9e-4 .. 3.0e-4 unscaled for R = 16, 32).  The data
// carry the scales through the passes (at most s16^2 = 3.5 times the f16
// values of the unscaled transform, far inside the f16 range) and the
// last pass's FP32 fractional-sample factor divides them out, so that the
// twiddle tables keep their exact entries (1, -i, ...).
template <int R>
constexpr float dft_scale() {
  return R == 8 ? 1.3603517f : R == 16 ? 1.8603300f : R == 32 ? 1.6268249f : 1.0f;
}

// One lane's operand of a DFT matrix times `scale`: row `row` of the 16-row
// tile, complex K index kc = 16 ks + 4 g + j (j = 0..3) as (re, im) half
// pairs.  Output row -> butterfly so = row / O, output rp = row % O; input
// kc -> butterfly kc / R, element kc % R; zero off the diagonal blocks.
// hi = f16(w), lo = f16(w - hi).
// The row may carry a complex factor fr + i fi (a twiddle of the output).
template <int R, int O>
__device__ void dft_operand(int row, int g, int ks, float scale, uint4 &re_hi, uint4 &re_lo, uint4 &im_hi,
                            uint4 &im_lo, float fr = 1.f, float fi = 0.f) {
  const int so = row / O, rp = row % O;
  uint32_t w[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int kc = 16 * ks + 4 * g + j;
    float c = 0.f, s = 0.f;
    if (kc / R == so) sincospif(2.0f * (float)(((kc % R) * rp) % R) / (float)R, &s, &c);
    // (c - i s)(fr + i fi), then the scale
    const float cr = c * fr + s * fi, sr = s * fr - c * fi;
    c = cr * scale;
    s = sr * scale;
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

// Y = A . B for one 16-column tile, both output components.  The FP32 form
// takes A and B as hi + lo (A_lo B_lo, below 2^-22 of the product, left
// out); the reduced-precision forms one scaled f16 matrix (dft_scale).
template <bool SPLIT>
__device__ __forceinline__ void dft_op(const uint4 &rh, const uint4 &rl, const uint4 &ih, const uint4 &il,
                                       const Opnd<SPLIT> &b, floatx4 &yr, floatx4 &yi) {
  yr = mma(rh, b.h, yr);
  yi = mma(ih, b.h, yi);
  if constexpr (SPLIT) {
    yr = mma(rl, b.h, yr);
    yi = mma(il, b.h, yi);
    yr = mma(rh, b.l, yr);
    yi = mma(ih, b.l, yi);
  }
}

// W_N^m = exp(-2 pi i m / N) in FP32
template <int N>
__device__ __forceinline__ float2 twf(int m) {
  float s, c;
  sincospif(2.0f * (float)(m % N) / (float)N, &s, &c);
  return make_float2(c, -s);
}

template <int L, int OUT>
__global__ void __launch_bounds__(MG<L>::NTHR, 4) fstage_mfma_kernel(FParams p, long long nunits) {
  using G = MG<L>;
  using TG = typename G::TG;
  constexpr int NW = G::NW, K01 = G::K01, K2 = G::K2, RL = G::RL, O = G::O, S = G::S, LQ = G::LQ;
  constexpr int NCHAN = TG::NCHAN, S2 = TG::S2, NROT = TG::S2 + TG::NSTEPS, TS = TG::S;
  // The FP32 form keeps every operand as hi + lo f16 pairs (planes 0 and 1
  // of the buffer) and its twiddles in FP32.
  constexpr bool SPLIT = OUT == OUT_F32;
  constexpr int TWW = SPLIT ? 2 : 1;                   // words per twiddle
  __shared__ __align__(16) uint32_t sd[TWW * G::PS];  // the transform, layouts A and B
  __shared__ __align__(16) uint32_t st0[TWW * 256];   // pass-0 twiddles, [s][r]
  __shared__ __align__(16) uint32_t st1[TWW * L];     // pass-1 twiddles, [t][c][g][i] (FP32: [t][g][c][i])
  uint32_t *const sl = sd + (SPLIT ? G::PS : 0);      // plane 1 (FP32 form)
  __shared__ float2 stab[TG::NTAB];             // rotator tables (fill_tables)
  // 2-bit levels (fxkernel.cpp:15-28): by 4-bit time slot (the nibble holding
  // both polarisations, A in its low 2 bits) as the f16 pair (level A,
  // level B), and by 2-bit code
  __shared__ uint32_t slevh[16];
  __shared__ float slev4[4];
  __shared__ Split ssplit[kSplitBatch];         // this workgroup's next units' delay splits

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c16 = lane & 15, g = lane >> 4;

  // Per-launch constants of this lane: the DFT matrices.
  // (scaled for the reduced-precision forms, whose lo parts go unused)
  constexpr float S16 = SPLIT ? 1.0f : dft_scale<16>(), SL = SPLIT ? 1.0f : dft_scale<RL>();
  uint4 d_rh, d_rl, d_ih, d_il;  // DFT-16 of passes 0 and 1
  dft_operand<16, 16>(c16, g, 0, S16, d_rh, d_rl, d_ih, d_il);
  // Pruned DFT-RL of the last pass.  Its two K-steps need only one operand
  // per lane: for RL <= 16 a row's butterfly lies in K-step row / 8 and its
  // rows are zero in the other, so the lane keeps the K-step of its own row
  // (rows >= 8 take the product with the second K-step's inputs); for
  // RL = 32 the second K-step's matrix is the first with row r scaled by
  // W_32^(16 r) = (-1)^r.
  uint4 l_rh, l_rl, l_ih, l_il;
  dft_operand<RL, O>(c16, g, RL == 32 ? 0 : c16 >> 3, SL, l_rh, l_rl, l_ih, l_il);
  // At 8192 points the reduced-precision forms fold pass 0's twiddles into
  // its matrices.  Output row r of the tile holding column b takes
  // W_256^(r s), s = b / 32: the same for the 16 columns of a tile, so the
  // twiddle scales row r of the DFT-16 (the lane's row c16 of the MFMA A
  // operand).  Wave w takes tiles 4w .. 4w+3, i.e. s = 2w and 2w + 1, one
  // folded matrix each; pass 0 then rounds its products to f16 directly.
  constexpr bool FOLD0 = !SPLIT && L == 8192;
  auto tile_of = [&](int w, int k) { return FOLD0 ? K01 * w + k : w + NW * k; };
  uint4 f_rh[2], f_ih[2];
  if constexpr (FOLD0) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float sn, cs;
      sincospif(2.0f * (float)((c16 * (2 * wave + h)) % 256) / 256.0f, &sn, &cs);
      uint4 unused0, unused1;
      dft_operand<16, 16>(c16, g, 0, S16, f_rh[h], unused0, f_ih[h], unused1, cs, -sn);
    }
  }
  // Twiddle tables, the same for every unit (f16 pairs, widened exactly in
  // twmul; FP32 for the FP32 form).  Pass-0 output row r of the column tile holding column b takes
  // W_L^(r RL (b / RL)) = W_256^(r s), s = b / RL < 16: st0[16 s + r].
  // Pass-1 output q1 = 256 t + c + 16 r (tile t, column c, row r = 4 g + i)
  // takes W_L^((c + 16 r) t): st1[256 t + 16 c + r] for the f16 tables.  The
  // FP32 form's 8-byte entries sit at st1[256 t + 64 g + 4 c + i] instead, so
  // that 16 lanes of one g read consecutive words rather than 8 of them
  // sharing banks (measured on MI355X at 2048 points: F-stage 3.88 -> 3.79
  // ms; the f16 tables lose about 1 % with the same change at 8192 points).
  for (int e = tid; e < 256; e += G::NTHR) {
    if constexpr (SPLIT) reinterpret_cast<float2 *>(st0)[e] = twf<256>((e & 15) * (e >> 4));
    else st0[e] = twh<256>((e & 15) * (e >> 4), 1.0f);
  }
  for (int e = tid; e < L; e += G::NTHR) {
    const int m = SPLIT ? (((e >> 2) & 15) + 16 * (4 * ((e >> 6) & 3) + (e & 3))) * (e >> 8)
                        : (((e >> 4) & 15) + 16 * (e & 15)) * (e >> 8);
    if constexpr (SPLIT) reinterpret_cast<float2 *>(st1)[e] = twf<L>(m);
    else st1[e] = twh<L>(m, 1.0f);
  }
  if (tid < 16) {
    const float lv[4] = {-kHiMag, -1.0f, 1.0f, kHiMag};
    slevh[tid] = hbits(toh2(lv[tid & 3], lv[tid >> 2]));
    if (tid < 4) slev4[tid] = lv[tid];
  }

  // LDS words of this lane's first access of each kind; the others sit at
  // constant offsets (last-pass input kc = 4 g + j + 16 ks of the butterfly
  // of column C is point S C + kc / RL + 256 (kc % RL)).
  const int p0w = G::lay_a(pt0t(tile_of(wave, 0), c16, g));
  const int p1r = G::lay_a(pt1r<L>(wave, c16, g, 0, 0));
  const int p1w = G::lay_b(pt1w<L>(wave, c16, g, 0, 0));
  const int p2r = G::lay_b(pt2r<L>(wave, c16, g, 0, 0, 0));
  // last-pass bins q = qb + 256 i (+ 16 S NW per tile): row 4 g + i of
  // column C is output r' = (4 g + i) % O of butterfly (4 g + i) / O
  const int qb0 = S * (16 * wave + c16) + (4 * g) / O + 256 * ((4 * g) % O);

  const float2 *fsub = stab + NROT;
  const float2 *fstep = fsub + TS;

  int it = 0;
  for (long long u = blockIdx.x; u < nunits; u += gridDim.x, it++) {
    // The FP64 delay splits of this workgroup's next kSplitBatch units, one
    // per lane of wave 0 (instead of every thread redoing its unit's).
    // Entries are read before the barrier at the top of each unit, so the
    // batch is only rewritten after every thread has read the last one.
    if (it % kSplitBatch == 0) {
      if (tid < kSplitBatch) {
        const long long v = u + (long long)gridDim.x * tid;
        Split sp{0.0, 0.0, 0.0, 0, 0};
        if (v < nunits) {
          const int vt = p.t0 + (int)(v / p.nant) * p.tstride;
          sp.valid = split_offset<L>(p, (int)(v % p.nant), vt, sp.offset, sp.frac, sp.a, sp.b) ? 1 : 0;
        }
        ssplit[tid] = sp;
      }
      __syncthreads();
    }
    const int tl = (int)(u / p.nant);
    const int st = (int)(u % p.nant);
    const int t = p.t0 + tl * p.tstride;
    const Split sp = ssplit[it % kSplitBatch];
    const int offset = sp.offset;
    const double frac = sp.frac, a = sp.a, b = sp.b;
    // the same for every thread: the branches below are uniform
    const bool valid = sp.valid != 0;
    if (tid == 0 && OUT != OUT_POWER) p.valid[(size_t)st * p.numffts + t] = valid ? 1 : 0;
    // This lane's input bytes (both polarisations) for every pass-0 tile:
    // window samples n = b + LQ r, r = 4 g + j, two samples to a byte.  All
    // are requested here so that their latency overlaps the table build.
    // The reduced-precision forms rotate each sample once for both
    // polarisations: polarisation A's pass 0 replaces the bytes here by
    // polarisation B's rotated samples (f16 pairs), its pass-0 operands.
    uint32_t raw[K01][4];
    if (valid) {
      gbyte *src = (gbyte *)p.station[st];
#pragma unroll
      for (int k = 0; k < K01; k++) {
        gbyte *bp = src + ((offset + 16 * tile_of(wave, k) + c16) >> 1) + (LQ / 2) * 4 * g;
#pragma unroll
        for (int j = 0; j < 4; j++) raw[k][j] = bp[(LQ / 2) * j];
      }
    }
    __syncthreads();  // the previous unit is done with LDS (and the twiddles are in)
    if (!valid) {
      for (int pol = 0; pol < 2; pol++) {
        const size_t row = ((size_t)tl * p.nant + st) * 2 + pol;
        for (int q = tid; q < NCHAN; q += G::NTHR)
          store_bin<OUT, NCHAN>(p, row, 2 * st + pol, q, make_float2(0.f, 0.f));
      }
      continue;
    }
    fill_tables<L, false>(p, frac, a, b, stab, tid, G::NTHR);
    __syncthreads();

    // The polarisations in turn, through the one buffer.
#pragma unroll 1
    for (int pol = 0; pol < 2; pol++) {
    const size_t row = ((size_t)tl * p.nant + st) * 2 + pol;

    // four consecutive twiddles of a table
    auto fetch_tw = [&](const uint32_t *tab, int e, Tw<SPLIT> *tw) {
      if constexpr (SPLIT) {
        const float4 a = reinterpret_cast<const float4 *>(tab)[e / 2];
        const float4 c = reinterpret_cast<const float4 *>(tab)[e / 2 + 1];
        tw[0] = make_float2(a.x, a.y);
        tw[1] = make_float2(a.z, a.w);
        tw[2] = make_float2(c.x, c.y);
        tw[3] = make_float2(c.z, c.w);
      } else {
        const uint4 v = *reinterpret_cast<const uint4 *>(&tab[e]);
        tw[0] = v.x;
        tw[1] = v.y;
        tw[2] = v.z;
        tw[3] = v.w;
      }
    };

    // Pass 0: DFT-16 over window samples n = b + LQ r, rotated on the way in
    // (rotator sub[n % S2] * step[n / S2], fxkernel.cpp:553-558; S2 divides
    // LQ, so sub is one entry per column).
    // Each tile's LDS reads are issued one tile ahead of its arithmetic.
    // One tile: the DFT of operand xo, the pass-0 twiddles, the LDS store.
    auto tile0 = [&](int k, const Opnd<SPLIT> &xo, const Tw<SPLIT> *tw) {
      floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
      uint32_t oh[4], ol[4];
      if constexpr (FOLD0) {
        dft_op<SPLIT>(f_rh[k >> 1], f_rh[k >> 1], f_ih[k >> 1], f_ih[k >> 1], xo, yr, yi);
#pragma unroll
        for (int i = 0; i < 4; i++) oh[i] = hbits(toh2(yr[i], yi[i]));
      } else {
        dft_op<SPLIT>(d_rh, d_rl, d_ih, d_il, xo, yr, yi);
#pragma unroll
        for (int i = 0; i < 4; i++) twist<SPLIT>(yr[i], yi[i], tw[i], oh[i], ol[i]);
      }
      const int a0 = p0w + (G::lay_a(pt0t(tile_of(0, k), 0, 0)) - G::lay_a(pt0t(tile_of(0, 0), 0, 0)));
      *reinterpret_cast<uint4 *>(&sd[a0]) = make_uint4(oh[0], oh[1], oh[2], oh[3]);
      if constexpr (SPLIT) *reinterpret_cast<uint4 *>(&sl[a0]) = make_uint4(ol[0], ol[1], ol[2], ol[3]);
    };
    auto tw0_of = [&](int k, Tw<SPLIT> *tw) {
      if constexpr (!FOLD0) {
        const int bc = 16 * tile_of(wave, k) + c16;
        fetch_tw(st0, 16 * (bc / RL) + 4 * g, tw);
      }
    };
    // The FP32 form rotates each polarisation's samples in its own pass 0;
    // the reduced-precision forms rotate a sample once, in polarisation A's
    // pass 0, for both, and keep B's operands in raw.
    if (SPLIT || pol == 0) {
      // levels of polarisation A and B (reduced precision) or of this
      // polarisation alone (FP32 form)
      using Lev = typename std::conditional<SPLIT, float, uint32_t>::type;
      struct In0 {
        float2 rsub, step[4];
        Lev lev[4];
        Tw<SPLIT> tw[4];
      };
      auto fetch0 = [&](int k) {
        In0 f;
        const int bc = 16 * tile_of(wave, k) + c16;
        const int sh = ((offset + bc) & 1) * 4 + (SPLIT ? 2 * pol : 0);
        const float2 *stp = stab + S2 + bc / S2 + (LQ / S2) * 4 * g;
        f.rsub = stab[bc % S2];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if constexpr (SPLIT)
            f.lev[j] = slev4[(raw[k][j] >> sh) & 3];
          else
            f.lev[j] = slevh[(raw[k][j] >> sh) & 15];
          f.step[j] = stp[(LQ / S2) * j];
        }
        tw0_of(k, f.tw);
        return f;
      };
      In0 f0 = fetch0(0);
#pragma unroll
      for (int k = 0; k < K01; k++) {
        In0 f1;
        if (k + 1 < K01) f1 = fetch0(k + 1);
        uint32_t wh[4], wl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const float2 r = cmul(f0.rsub, f0.step[j]);
          if constexpr (SPLIT) {
            pack<SPLIT>(f0.lev[j] * r.x, f0.lev[j] * r.y, wh[j], wl[j]);
          } else {
            // the rotator rounded to f16 once, times each polarisation's
            // f16 level in packed f16 (+-1 exact; 3.3359 is 3.3359375 in
            // f16, 1.1e-5 high)
            const h2 rh = toh2(r.x, r.y), lv = hval(f0.lev[j]);
            wh[j] = hbits(rh * lv.xx);
            raw[k][j] = hbits(rh * lv.yy);
          }
        }
        Opnd<SPLIT> xo;
        xo.h = make_uint4(wh[0], wh[1], wh[2], wh[3]);
        if constexpr (SPLIT) xo.l = make_uint4(wl[0], wl[1], wl[2], wl[3]);
        tile0(k, xo, f0.tw);
        if (k + 1 < K01) f0 = f1;
      }
    } else if constexpr (!SPLIT) {
      Tw<SPLIT> t0[4];
      tw0_of(0, t0);
#pragma unroll
      for (int k = 0; k < K01; k++) {
        Tw<SPLIT> t1[4];
        if (k + 1 < K01) tw0_of(k + 1, t1);
        Opnd<SPLIT> xo;
        xo.h = make_uint4(raw[k][0], raw[k][1], raw[k][2], raw[k][3]);
        tile0(k, xo, t0);
        if (k + 1 < K01)
#pragma unroll
          for (int i = 0; i < 4; i++) t0[i] = t1[i];
      }
    }
    __syncthreads();

    // The output scales of the last pass, requested now to arrive during pass 1.
    float sc[K2][4];
#pragma unroll
    for (int k = 0; k < K2; k++)
#pragma unroll
      for (int i = 0; i < 4; i++) sc[k][i] = bin_scale<OUT, NCHAN>(p, 2 * st + pol, qb0 + 16 * S * NW * k + 256 * i);

    // Pass 1: DFT-16 over in[b + LQ r]; output q1 = (b/16) 256 + b%16 + 16 r'.
    // Every wave reads its inputs before any writes its outputs.
    uint32_t o1h[K01][4], o1l[K01][4];
    auto fetch1 = [&](int k, Opnd<SPLIT> &x, Tw<SPLIT> *tw) {
      uint32_t wh[4], wl[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int a = p1r + (G::lay_a(pt1r<L>(0, 0, 0, k, j)) - G::lay_a(pt1r<L>(0, 0, 0, 0, 0)));
        wh[j] = sd[a];
        if constexpr (SPLIT) wl[j] = sl[a];
      }
      x.h = make_uint4(wh[0], wh[1], wh[2], wh[3]);
      if constexpr (SPLIT) x.l = make_uint4(wl[0], wl[1], wl[2], wl[3]);
      fetch_tw(st1, 256 * (wave + NW * k) + (SPLIT ? 64 * g + 4 * c16 : 16 * c16 + 4 * g), tw);
    };
    Opnd<SPLIT> x0;
    Tw<SPLIT> tw0[4];
    fetch1(0, x0, tw0);
#pragma unroll
    for (int k = 0; k < K01; k++) {
      Opnd<SPLIT> x1;
      Tw<SPLIT> tw1[4];
      if (k + 1 < K01) fetch1(k + 1, x1, tw1);
      floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
      dft_op<SPLIT>(d_rh, d_rl, d_ih, d_il, x0, yr, yi);
#pragma unroll
      for (int i = 0; i < 4; i++) twist<SPLIT>(yr[i], yi[i], tw0[i], o1h[k][i], o1l[k][i]);
      if (k + 1 < K01) {
        x0 = x1;
#pragma unroll
        for (int i = 0; i < 4; i++) tw0[i] = tw1[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K01; k++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int a = p1w + (G::lay_b(pt1w<L>(0, 0, 0, k, i)) - G::lay_b(pt1w<L>(0, 0, 0, 0, 0)));
        sd[a] = o1h[k][i];
        if constexpr (SPLIT) sl[a] = o1l[k][i];
      }
    __syncthreads();

    // Last pass: pruned DFT-RL over in[b + 256 r], b = S C + s for MFMA
    // column C; bin q = b + 256 r' for r' < RL/2, fractional-sample rotation
    // fsub[q % TS] * fstep[q / TS] (fxkernel.cpp:614-625) in FP32, store.
    // TS divides 256 and 16 S NW, so fsub's entry is the lane's for every
    // tile and row.
    // with the 1/s of the three scaled matrices
    constexpr float UNSCALE = 1.0f / (S16 * S16 * SL);
    const float2 fs = make_float2(fsub[qb0 % TS].x * UNSCALE, fsub[qb0 % TS].y * UNSCALE);
#pragma unroll
    for (int k = 0; k < K2; k++) {
      Opnd<SPLIT> x[2];  // B operands of K-steps 0 and 1
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        uint32_t wh[4], wl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int a = p2r + (G::lay_b(pt2r<L>(0, 0, 0, k, ks, j)) - G::lay_b(pt2r<L>(0, 0, 0, 0, 0, 0)));
          wh[j] = sd[a];
          if constexpr (SPLIT) wl[j] = sl[a];
        }
        x[ks].h = make_uint4(wh[0], wh[1], wh[2], wh[3]);
        if constexpr (SPLIT) x[ks].l = make_uint4(wl[0], wl[1], wl[2], wl[3]);
      }
      // output row 4 g + i sits in element i of lane group g
      floatx4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
      dft_op<SPLIT>(l_rh, l_rl, l_ih, l_il, x[1], yr, yi);
      if constexpr (RL == 32) {
        // (-1)^r on the second K-step's product (odd rows: odd i), then add
        // the first K-step's into it
        yr[1] = -yr[1];
        yr[3] = -yr[3];
        yi[1] = -yi[1];
        yi[3] = -yi[3];
        dft_op<SPLIT>(l_rh, l_rl, l_ih, l_il, x[0], yr, yi);
      } else {
        // rows < 8 (g < 2) take the first K-step's product
        floatx4 zr = {0.f, 0.f, 0.f, 0.f}, zi = {0.f, 0.f, 0.f, 0.f};
        dft_op<SPLIT>(l_rh, l_rl, l_ih, l_il, x[0], zr, zi);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          yr[i] = g < 2 ? zr[i] : yr[i];
          yi[i] = g < 2 ? zi[i] : yi[i];
        }
      }
      const int qb = qb0 + 16 * S * NW * k;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float2 fr = cmul(fs, fstep[qb / TS + (256 / TS) * i]);
        const float2 y = cmul(fr, make_float2(yr[i], yi[i]));
        if constexpr (OUT == OUT_F16) {
          // the scaled pair rounded to f16 in one conversion (as store_bin_sc)
          __builtin_nontemporal_store(hbits(toh2(y.x * sc[k][i], y.y * sc[k][i])),
                                      static_cast<uint32_t *>(p.out) + row * NCHAN + qb + 256 * i);
        } else {
          store_bin_sc<OUT, NCHAN>(p, row, qb + 256 * i, y, sc[k][i]);
        }
      }
    }
    if (pol == 0) __syncthreads();  // the next pass 0 rewrites the buffer
    }  // pol
  }
}

template <int L, int OUT>
hipError_t launch_l(const FParams &p, hipStream_t s) {
  const long long nunits = (long long)p.nant * p.nt;
  if (nunits <= 0) return hipSuccess;
  const void *k = reinterpret_cast<const void *>(&fstage_mfma_kernel<L, OUT>);
  const long long grid = persistent_grid(k, MG<L>::NTHR, 0, nunits);
  hipLaunchKernelGGL((fstage_mfma_kernel<L, OUT>), dim3((unsigned)grid), dim3(MG<L>::NTHR), 0, s, p, nunits);
  return hipGetLastError();
}

template <int L>
hipError_t launch_out(const FParams &p, hipStream_t s) {
  switch (p.out_mode) {
    case OUT_F32:
      if constexpr (L <= 4096) return launch_l<L, OUT_F32>(p, s);
      return hipErrorInvalidValue;
    case OUT_F16: return launch_l<L, OUT_F16>(p, s);
    case OUT_BF16: return launch_l<L, OUT_BF16>(p, s);
    case OUT_FP8: return launch_l<L, OUT_FP8>(p, s);
    case OUT_POWER: return launch_l<L, OUT_POWER>(p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace
}  // namespace fs

bool fstage_mfma_supported(int L, bool cplx, bool fp32) { return !cplx && L >= 2048 && L <= (fp32 ? 4096 : 8192); }

hipError_t launch_fstage_mfma(const FParams &p, int L, hipStream_t s) {
  switch (L) {
    case 2048: return fs::launch_out<2048>(p, s);
    case 4096: return fs::launch_out<4096>(p, s);
    case 8192: return fs::launch_out<8192>(p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gxc
