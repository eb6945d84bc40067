// Device helpers shared by the F-stage kernels: fstage_item.hpp (FFT held in
// LDS with both polarisations, L <= 8192) and fstage_big.hpp (one
// polarisation per workgroup, L = 2048..32768).
#pragma once

#include "gxc_internal.hpp"

namespace gxc {
namespace fs {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 mnegi(float2 a) { return make_float2(a.y, -a.x); }  // -i*a

// Compile-time cos/sin for twiddle constants (Taylor series in double,
// |x| <= 2 pi, 40 terms: accurate far below FP32 resolution).
constexpr double cx_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 40; n++) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 40; n++) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}

// Multiply by the compile-time twiddle W_N^K = exp(-2 pi i K / N).  Multiples
// of pi/4 use exact forms; other angles use FP32-rounded constants.
template <int N, int K>
__device__ __forceinline__ float2 twc(float2 a) {
  constexpr int k = K % N;
  constexpr float c = 0.70710678118654752f;
  if constexpr (k == 0) {
    return a;
  } else if constexpr (4 * k == N) {
    return mnegi(a);
  } else if constexpr (2 * k == N) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * k == 3 * N) {
    return make_float2(-a.y, a.x);
  } else if constexpr (8 * k == N) {
    return make_float2((a.x + a.y) * c, (a.y - a.x) * c);
  } else if constexpr (8 * k == 3 * N) {
    return make_float2((a.y - a.x) * c, -(a.x + a.y) * c);
  } else if constexpr (8 * k == 5 * N) {
    return make_float2(-(a.x + a.y) * c, (a.x - a.y) * c);
  } else if constexpr (8 * k == 7 * N) {
    return make_float2((a.x - a.y) * c, (a.x + a.y) * c);
  } else {
    constexpr double ang = 2.0 * kPi * (double)k / (double)N;
    constexpr float wr = (float)cx_cos(ang), wi = (float)-cx_sin(ang);
    return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
  }
}

// plus = a + W_N^K b, minus = a - W_N^K b for a compile-time twiddle.
// Off the multiples of N/4 the twiddle is factored as W = c (1 + i t) with
// |t| <= 1 (|cos| >= |sin|) or W = s (t + i): b' = (1 + i t) b (or (t + i) b)
// takes 2 FMAs and a +- c b' 4 more, 6 instructions where a complex product
// and two complex additions take 8 (the fused multiply-add form of
// Linzer and Feig, 1993).
template <int N, int K>
__device__ __forceinline__ void bfly_tw(float2 a, float2 b, float2 &plus, float2 &minus) {
  constexpr int k = K % N;
  if constexpr (k == 0) {
    plus = cadd(a, b);
    minus = csub(a, b);
  } else if constexpr (4 * k == N || 2 * k == N || 4 * k == 3 * N) {
    const float2 w = twc<N, K>(b);
    plus = cadd(a, w);
    minus = csub(a, w);
  } else {
    constexpr double ang = 2.0 * kPi * (double)k / (double)N;
    constexpr double wr = cx_cos(ang), wi = -cx_sin(ang);
    constexpr bool cosform = (wr < 0 ? -wr : wr) >= (wi < 0 ? -wi : wi);
    constexpr float c = (float)(cosform ? wr : wi);
    constexpr float t = (float)(cosform ? wi / wr : wr / wi);
    const float2 bp = cosform ? make_float2(b.x - t * b.y, b.y + t * b.x)
                              : make_float2(t * b.x - b.y, t * b.y + b.x);
    plus = make_float2(a.x + c * bp.x, a.y + c * bp.y);
    minus = make_float2(a.x - c * bp.x, a.y - c * bp.y);
  }
}

__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
  const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  const float2 t2 = cadd(a1, a3), t3 = mnegi(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <int K>
__device__ __forceinline__ float2 tw32(float2 a) {
  return twc<32, K>(a);
}

// v[q] = e[q] + W32^q o[q], v[q + 16] = e[q] - W32^q o[q] for q = Q .. 15
template <int Q>
__device__ __forceinline__ void bfly32(const float2 *e, const float2 *o, float2 *v) {
  if constexpr (Q < 16) {
    bfly_tw<32, Q>(e[Q], o[Q], v[Q], v[Q + 16]);
    bfly32<Q + 1>(e, o, v);
  }
}

// In-place forward DFT of R points v[0..R), natural order in and out.
template <int R>
__device__ __forceinline__ void dft(float2 *v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    dft4(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    // k = k1 + 2 k2: DFT4 over k2, twiddle W8^(k1 q2), DFT2 over k1.
    dft4(v[0], v[2], v[4], v[6]);
    dft4(v[1], v[3], v[5], v[7]);
    const float2 y00 = v[0], y01 = v[2], y02 = v[4], y03 = v[6];
    const float2 y10 = v[1], y11 = v[3], y12 = v[5], y13 = v[7];
    bfly_tw<8, 0>(y00, y10, v[0], v[4]);
    bfly_tw<8, 1>(y01, y11, v[1], v[5]);
    bfly_tw<8, 2>(y02, y12, v[2], v[6]);
    bfly_tw<8, 3>(y03, y13, v[3], v[7]);
  } else if constexpr (R == 16) {
    // k = k1 + 4 k2, q = 4 q1 + q2: DFT4 over k2, twiddle W16^(k1 q2), DFT4 over k1.
    dft4(v[0], v[4], v[8], v[12]);
    dft4(v[1], v[5], v[9], v[13]);
    dft4(v[2], v[6], v[10], v[14]);
    dft4(v[3], v[7], v[11], v[15]);
    // y[k1][q2] = W16^(k1 q2) v[k1 + 4 q2]; k1 = 2 enters its DFT4 as
    // y0 +- W16^(2 q2) v, one twiddled butterfly (bfly_tw)
    float2 y[4][4];
#pragma unroll
    for (int q2 = 0; q2 < 4; q2++) {
      y[0][q2] = v[4 * q2];
      y[2][q2] = v[2 + 4 * q2];
    }
    y[1][0] = v[1]; y[1][1] = twc<16, 1>(v[5]); y[1][2] = twc<16, 2>(v[9]);  y[1][3] = twc<16, 3>(v[13]);
    y[3][0] = v[3]; y[3][1] = twc<16, 3>(v[7]); y[3][2] = twc<16, 6>(v[11]); y[3][3] = twc<16, 9>(v[15]);
    float2 t0[4], t1[4];
    bfly_tw<16, 0>(y[0][0], y[2][0], t0[0], t1[0]);
    bfly_tw<16, 2>(y[0][1], y[2][1], t0[1], t1[1]);
    bfly_tw<16, 4>(y[0][2], y[2][2], t0[2], t1[2]);
    bfly_tw<16, 6>(y[0][3], y[2][3], t0[3], t1[3]);
#pragma unroll
    for (int q2 = 0; q2 < 4; q2++) {
      const float2 t2 = cadd(y[1][q2], y[3][q2]), t3 = mnegi(csub(y[1][q2], y[3][q2]));
      v[q2] = cadd(t0[q2], t2);
      v[8 + q2] = csub(t0[q2], t2);
      v[4 + q2] = cadd(t1[q2], t3);
      v[12 + q2] = csub(t1[q2], t3);
    }
  } else if constexpr (R == 32) {
    // Even/odd split: X[q] = E[q] + W32^q O[q], X[q+16] = E[q] - W32^q O[q].
    float2 e[16], o[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
      e[m] = v[2 * m];
      o[m] = v[2 * m + 1];
    }
    dft<16>(e);
    dft<16>(o);
    bfly32<0>(e, o, v);
  }
}

// Sizes of the per-item rotator tables for FFT length L (fxkernel.cpp:96-114):
// sub and step fringe-rotator factors, then sub and step fractional-sample
// factors.
template <int L, bool CPLX>
struct TabGeom {
  static constexpr int NCHAN = CPLX ? L : L / 2;
  static constexpr int S = 1 << (ilog2c(NCHAN) / 2);
  static constexpr int S2 = CPLX ? S : 2 * S;
  static constexpr int NSTEPS = L / S2;
  static constexpr int FRACSTEPS = NCHAN / S;
  static constexpr int NTAB = S2 + NSTEPS + S + FRACSTEPS;
};

// FP64 delay model and integer/fractional split for station st, FFT t:
// fxkernel.cpp:425-440 (getStationDelay) and 288-309, one rounding per
// operation (no contraction), so offsets and validity match the CPU exactly.
template <int L>
__device__ __forceinline__ bool split_offset(const FParams &p, int st, int t, int &offset,
                                             double &frac, double &a, double &b) {
#pragma clang fp contract(off)
  const double *dp = p.delays + 4 * st;
  const double c0 = dp[0], c1 = dp[1], c2 = dp[2], fileoffset = dp[3];
  const double t0 = (double)t, th = t + 0.5, t1 = t + 1.0;
  const double d0 = c0 * t0 * t0 + c1 * t0 + c2;
  const double d1 = c0 * th * th + c1 * th + c2;
  const double d2 = c0 * t1 * t1 + c1 * t1 + c2;
  a = d2 - d0;
  b = d0 + (d1 - (a * 0.5 + d0)) / 3.0;
  const double mean = a * 0.5 + b;
  const double ndf = (mean - fileoffset) / p.sampletime;
  if (!(fabs(ndf) < 1.0e9)) {
    offset = 0;
    frac = 0.0;
    return false;
  }
  const int nd = (int)(ndf + 0.5);  // truncation toward zero, fxkernel.cpp:290
  frac = -(ndf - nd) * p.sampletime;
  long long off = (long long)t * L - nd;
  if (off == -1) {
    ++off;
    frac += p.sampletime;
  }
  if (off == (long long)p.maxoffset + 1) {
    --off;
    frac -= p.sampletime;
  }
  offset = (int)off;
  if (off < 0 || off > p.maxoffset) return false;
  if (p.frame_ok) {
    // samples [off, off + L) lie in frames off/spf .. (off+L-1)/spf, all
    // below the station's slot count since off + L <= (numffts+1) L
    const uint8_t *ok = p.frame_ok[st];
    const long long spf = p.frame_samples[st];
    const long long f1 = (off + L - 1) / spf;
    for (long long f = off / spf; f <= f1; f++)
      if (!ok[f]) return false;
  }
  return true;
}

// Entries e = e0, e0+estep, ... of one item's rotator tables: sub, step
// (fringe rotation, fxkernel.cpp:498-567) and fsub, fstep (fractional-sample
// correction, fxkernel.cpp:595-634).  Phases in FP64, reduced, rounded to
// FP32, then the hardware sine/cosine (v_sin_f32/v_cos_f32 via __sincosf).
// Every argument is already reduced: the fringe phases to (-2 pi, 0], the
// fractional-sample phases to |2 pi c bw/nchan * frac| < 2 pi (|frac| is at
// most a sample), where the hardware result is within ~5e-7 of libm
// sincosf (the 2^-24 rounding of x/2 pi times 2 pi, plus the unit's own
// error); IPP's ippsSinCos_32f_A11, which fxkernel uses, is itself only
// ~11-bit accurate.
template <int L, bool CPLX>
__device__ __forceinline__ void fill_tables(const FParams &p, double frac, double a, double b, float2 *tab,
                                            int e0, int estep) {
#pragma clang fp contract(off)
  using TG = TabGeom<L, CPLX>;
  constexpr int S = TG::S, S2 = TG::S2, NSTEPS = TG::NSTEPS;
  // fxkernel.cpp:503-505: remove whole seconds of delay.
  const double intdelay = trunc(b);
  const double bf = b - intdelay;
  const float ffrac = (float)frac;  // ippsMulC_32f takes an Ipp32f (fxkernel.cpp:600)
  for (int e = e0; e < TG::NTAB; e += estep) {
    float sn, cs;
    if (e < S2) {
      // subphase, fxkernel.cpp:107,508,514,519,525-530,535
      double ph = ((double)e / (double)L) * a;
      ph = ph + bf;
      ph = ph * p.lo;
      if (p.fraclo) ph = ph + p.lofrac * intdelay;
      const double arg = (-2.0 * kPi) * (ph - trunc(ph));
      __sincosf((float)arg, &sn, &cs);
    } else if (e < S2 + NSTEPS) {
      // stepphase, fxkernel.cpp:112,511,522,539
      const int i = e - S2;
      double ph = ((double)(i * p.stepmul) / (double)L) * a;
      ph = ph * p.lo;
      const double arg = (-2.0 * kPi) * (ph - trunc(ph));
      __sincosf((float)arg, &sn, &cs);
    } else if (e < S2 + NSTEPS + S) {
      // fxkernel.cpp:600,608
      __sincosf(p.subchanfreq[e - S2 - NSTEPS] * ffrac, &sn, &cs);
    } else {
      // fxkernel.cpp:603,611
      __sincosf(p.stepchanfreq[e - S2 - NSTEPS - S] * ffrac, &sn, &cs);
    }
    tab[e] = make_float2(cs, sn);
  }
}

__device__ __forceinline__ uint32_t pack_f16(float re, float im) {
  const _Float16 a = (_Float16)re, b = (_Float16)im;
  return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

__device__ __forceinline__ uint32_t bf16_bits(float x) {
  const uint32_t u = __builtin_bit_cast(uint32_t, x);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;  // round to nearest even
}

__device__ __forceinline__ uint32_t pack_bf16(float re, float im) {
  return bf16_bits(re) | (bf16_bits(im) << 16);
}

__device__ __forceinline__ uint16_t pack_fp8(float re, float im) {
  // e4m3 with saturation to the largest finite value before conversion.
  re = fminf(fmaxf(re, -448.f), 448.f);
  im = fminf(fmaxf(im, -448.f), 448.f);
  return (uint16_t)(__builtin_amdgcn_cvt_pk_fp8_f32(re, im, 0, false) & 0xffff);
}

// Scale of bin q of input srow (= 2*station+pol) for the reduced-precision
// forms; 1 for the others.
template <int OUT, int NCHAN>
__device__ __forceinline__ float bin_scale(const FParams &p, int srow, int q) {
  if constexpr (OUT == OUT_F32 || OUT == OUT_POWER) {
    return 1.f;
  } else {
    return p.scales[(size_t)srow * NCHAN + q];
  }
}

// Store bin q of output row `row` (= 2*station + pol within FFT row block)
// in the spectra format OUT, with its scale sc = bin_scale(...) already read.
// The spectra are written once and read once, by the X-stage after the
// whole F-stage: streaming (nontemporal) stores of the 4- and 8-byte
// values, so that the X-stage no longer meets the write-back of dirty
// spectra lines (16 stations x 4096 channels on MI355X: FP16 X-stage 3.06
// -> 2.78 ms, FP32 F-stage -0.8 %).  The 2-byte FP8 values keep ordinary
// stores, which the caches combine (as streaming stores the 16-station FP8
// F-stage measured 2-5 % slower).
template <int OUT, int NCHAN>
__device__ __forceinline__ void store_bin_sc(const FParams &p, size_t row, int q, float2 y, float sc) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  if constexpr (OUT == OUT_F32) {
    const f32x2 v = {y.x, y.y};
    __builtin_nontemporal_store(v, reinterpret_cast<f32x2 *>(p.spectra + row * NCHAN + q));
  } else if constexpr (OUT == OUT_POWER) {
    // one row per sampled FFT; the scale kernel sums rows in a fixed order,
    // so the scales do not depend on scheduling
    p.power[row * NCHAN + q] = y.x * y.x + y.y * y.y;
  } else if constexpr (OUT == OUT_FP8) {
    static_cast<uint16_t *>(p.out)[row * NCHAN + q] = pack_fp8(y.x * sc, y.y * sc);
  } else if constexpr (OUT == OUT_F16) {
    __builtin_nontemporal_store(pack_f16(y.x * sc, y.y * sc), static_cast<uint32_t *>(p.out) + row * NCHAN + q);
  } else {
    __builtin_nontemporal_store(pack_bf16(y.x * sc, y.y * sc), static_cast<uint32_t *>(p.out) + row * NCHAN + q);
  }
}

// Store bin q of output row `row`; `srow` is the input index 2*station+pol.
template <int OUT, int NCHAN>
__device__ __forceinline__ void store_bin(const FParams &p, size_t row, int srow, int q, float2 y) {
  store_bin_sc<OUT, NCHAN>(p, row, q, y, bin_scale<OUT, NCHAN>(p, srow, q));
}

// Workgroups are dealt to the 8 XCDs round robin in dispatch order, each XCD
// with its own L2.  Renumbering unit v of `total` so that consecutive units
// run on one XCD one after another lets neighbouring units share lines in
// that L2 (the tail beyond a multiple of 8 keeps its number).
constexpr int kXcdCount = 8;
__device__ __forceinline__ long long xcd_order(long long v, long long total) {
  const long long per = total / kXcdCount;
  if (v >= per * kXcdCount) return v;
  return (v % kXcdCount) * per + v / kXcdCount;
}

// Global-address-space view of the packed station buffers: the loads then
// use the global rather than the flat address path.
typedef const __attribute__((address_space(1))) uint8_t gbyte;
typedef const __attribute__((address_space(1))) int gint;

}  // namespace fs
}  // namespace gxc
