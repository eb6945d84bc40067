// F-stage for gfx950: one workgroup slot ("item") per (station, FFT).
//
// Per item, fused in one kernel so the packed samples are read once and the
// spectra written once:
//   1. FP64 delay model and integer/fractional split (fxkernel.cpp:288-309,
//      425-440), evaluated with one rounding per operation in fxkernel's
//      order so the sample shift and validity decisions are bit-exact.
//   2. Fringe-rotator and fractional-sample tables (fxkernel.cpp:498-567,
//      595-634): FP64 phase, reduced, rounded to FP32, sincosf; kept in LDS.
//   3. First FFT pass straight from global memory: 2-bit (LDS LUT) or
//      complex int8 unpack, multiplied by the rotator sub[k]*step[i].
//   4. Remaining Stockham passes through LDS; each thread keeps 16 points of
//      both polarisations in registers, so twiddles and rotator values are
//      loaded once for two FFTs.
//   5. Last pass writes only bins < nchan, multiplied by the fractional
//      sample rotator, as coalesced stores to spectra[t][2*station+pol][chan].
#include "gxc_internal.hpp"

namespace gxc {
namespace {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 mnegi(float2 a) { return make_float2(a.y, -a.x); }  // -i*a

// Multiply by the compile-time twiddle W_N^K = exp(-2 pi i K / N).
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
    static_assert(N == 16, "general twiddle constants only for N == 16");
    constexpr float c1 = 0.92387953251128676f, s1 = 0.38268343236508977f;
    // W16^k = cos(2 pi k/16) - i sin(2 pi k/16) for odd k
    constexpr float wr = (k == 1 || k == 15) ? c1 : (k == 3 || k == 13) ? s1
                         : (k == 5 || k == 11) ? -s1 : -c1;
    constexpr float wi = (k == 1 || k == 7) ? -s1 : (k == 3 || k == 5) ? -c1
                         : (k == 9 || k == 15) ? s1 : c1;
    return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
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

// In-place forward DFT of R points v[0..R), natural order in and out.
template <int R>
__device__ __forceinline__ void dft(float2 *v) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    dft4(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    // k = k1 + 2 k2: DFT4 over k2, twiddle W8^(k1 q2), DFT2 over k1.
    dft4(v[0], v[2], v[4], v[6]);
    dft4(v[1], v[3], v[5], v[7]);
    const float2 y10 = v[1], y11 = twc<8, 1>(v[3]), y12 = twc<8, 2>(v[5]), y13 = twc<8, 3>(v[7]);
    const float2 y00 = v[0], y01 = v[2], y02 = v[4], y03 = v[6];
    v[0] = cadd(y00, y10); v[4] = csub(y00, y10);
    v[1] = cadd(y01, y11); v[5] = csub(y01, y11);
    v[2] = cadd(y02, y12); v[6] = csub(y02, y12);
    v[3] = cadd(y03, y13); v[7] = csub(y03, y13);
  } else if constexpr (R == 16) {
    // k = k1 + 4 k2, q = 4 q1 + q2: DFT4 over k2, twiddle W16^(k1 q2), DFT4 over k1.
    dft4(v[0], v[4], v[8], v[12]);
    dft4(v[1], v[5], v[9], v[13]);
    dft4(v[2], v[6], v[10], v[14]);
    dft4(v[3], v[7], v[11], v[15]);
    // y[k1][q2] now sits at v[k1 + 4 q2]
    float2 y[4][4];
#pragma unroll
    for (int q2 = 0; q2 < 4; q2++) y[0][q2] = v[4 * q2];
    y[1][0] = v[1]; y[1][1] = twc<16, 1>(v[5]); y[1][2] = twc<16, 2>(v[9]);  y[1][3] = twc<16, 3>(v[13]);
    y[2][0] = v[2]; y[2][1] = twc<16, 2>(v[6]); y[2][2] = twc<16, 4>(v[10]); y[2][3] = twc<16, 6>(v[14]);
    y[3][0] = v[3]; y[3][1] = twc<16, 3>(v[7]); y[3][2] = twc<16, 6>(v[11]); y[3][3] = twc<16, 9>(v[15]);
#pragma unroll
    for (int q2 = 0; q2 < 4; q2++) {
      float2 a0 = y[0][q2], a1 = y[1][q2], a2 = y[2][q2], a3 = y[3][q2];
      dft4(a0, a1, a2, a3);
      v[q2] = a0; v[4 + q2] = a1; v[8 + q2] = a2; v[12 + q2] = a3;
    }
  }
}

// Compile-time geometry for FFT length L.
template <int L, bool CPLX>
struct FG {
  static constexpr int LG = ilog2c(L);
  static constexpr int NCHAN = CPLX ? L : L / 2;
  static constexpr int S = 1 << (ilog2c(NCHAN) / 2);
  static constexpr int S2 = CPLX ? S : 2 * S;
  static constexpr int NSTEPS = L / S2;
  static constexpr int FRACSTEPS = NCHAN / S;
  static constexpr int NTAB = S2 + NSTEPS + S + FRACSTEPS;
  static constexpr int T = L / 16;                   // threads per item
  static constexpr int IPW = T >= 256 ? 1 : 256 / T; // items per workgroup
  static constexpr int BLOCK = T * IPW;
  static constexpr int PADL = L + L / 16;            // LDS slots per pol (1 pad per 16)
  // Radix schedule: radix-16 passes, then one smaller pass if log2 L % 4 != 0.
  static constexpr int N16 = LG / 4;
  static constexpr int REM = 1 << (LG % 4);
  static constexpr int NPASS = N16 + (REM > 1 ? 1 : 0);
  static constexpr int radix(int p) { return p < N16 ? 16 : REM; }
  static constexpr int ns(int p) {
    int n = 1;
    for (int i = 0; i < p; i++) n *= radix(i);
    return n;
  }
  static_assert(L >= 64 && L <= 8192, "LDS-resident F-stage covers 64 <= L <= 8192");
};

__device__ __forceinline__ int padi(int i) { return i + (i >> 4); }

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
  return !(off < 0 || off > p.maxoffset);
}

// LDS index of element i with one pad slot per 16 elements.  When a stride
// is a multiple of 16 the padded address of i + k*stride is the padded
// address of i plus k*stride*17/16, a compile-time offset.
template <int STRIDE>
__device__ __forceinline__ int pad_step(int k) {
  return (STRIDE % 16 == 0) ? k * (STRIDE / 16) * 17 : -1;
}

template <int L, bool CPLX, int P>
__device__ __forceinline__ void pass_from_lds(float2 (&va)[16], float2 (&vb)[16], const float2 *sa,
                                              const float2 *sb, int j, const float2 *__restrict__ tw) {
  using G = FG<L, CPLX>;
  constexpr int R = G::radix(P), NS = G::ns(P), NB = 16 / R, T = G::T;
  constexpr int STR = L / R;
#pragma unroll
  for (int bf = 0; bf < NB; bf++) {
    const int jj = j + bf * T;
    const int b0 = padi(jj);
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int idx = (STR % 16 == 0) ? b0 + pad_step<STR>(k) : padi(jj + k * STR);
      va[bf * R + k] = sa[idx];
      vb[bf * R + k] = sb[idx];
    }
    const unsigned m = (unsigned)(jj % NS) * (unsigned)(L / (NS * R));
#pragma unroll
    for (int k = 1; k < R; k++) {
      const float2 w = tw[m * (unsigned)k];
      va[bf * R + k] = cmul(va[bf * R + k], w);
      vb[bf * R + k] = cmul(vb[bf * R + k], w);
    }
    dft<R>(&va[bf * R]);
    dft<R>(&vb[bf * R]);
  }
}

template <int L, bool CPLX, int P>
__device__ __forceinline__ void pass_to_lds(const float2 (&va)[16], const float2 (&vb)[16], float2 *sa,
                                            float2 *sb, int j) {
  using G = FG<L, CPLX>;
  constexpr int R = G::radix(P), NS = G::ns(P), NB = 16 / R, T = G::T;
#pragma unroll
  for (int bf = 0; bf < NB; bf++) {
    const int jj = j + bf * T;
    const int base = (jj / NS) * NS * R + (jj % NS);
    const int b0 = padi(base);
#pragma unroll
    for (int k = 0; k < R; k++) {
      // NS == 1: base = R*jj, so base + k stays inside one 16-block (R <= 16).
      const int idx = (NS % 16 == 0) ? b0 + pad_step<NS>(k) : (NS == 1 && R == 16) ? b0 + k
                                                                                   : padi(base + k * NS);
      sa[idx] = va[bf * R + k];
      sb[idx] = vb[bf * R + k];
    }
  }
}

// Middle passes P = 1 .. NPASS-2 (compile-time recursion).
template <int L, bool CPLX, int P>
__device__ __forceinline__ void middle_passes(float2 (&va)[16], float2 (&vb)[16], float2 *sa, float2 *sb,
                                              int j, const float2 *__restrict__ tw) {
  using G = FG<L, CPLX>;
  if constexpr (P < G::NPASS - 1) {
    pass_from_lds<L, CPLX, P>(va, vb, sa, sb, j, tw);
    __syncthreads();
    pass_to_lds<L, CPLX, P>(va, vb, sa, sb, j);
    __syncthreads();
    middle_passes<L, CPLX, P + 1>(va, vb, sa, sb, j, tw);
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

template <int L, bool CPLX, int OUT>
__global__ void __launch_bounds__((FG<L, CPLX>::BLOCK)) fstage_kernel(FParams p) {
  using G = FG<L, CPLX>;
  constexpr int T = G::T, S = G::S, S2 = G::S2, NSTEPS = G::NSTEPS;
  __shared__ float2 sdata[G::IPW][2][G::PADL];
  __shared__ float2 stab[G::IPW][G::NTAB];
  __shared__ float2 slut[16];

  const int li = threadIdx.x / T;
  const int j = threadIdx.x % T;
  const int item = blockIdx.x * G::IPW + li;
  const bool active = item < p.nant * p.nt;
  const int tl = active ? item / p.nant : 0;
  const int st = active ? item % p.nant : 0;
  const int t = p.t0 + tl * p.tstride;

  if constexpr (!CPLX) {
    // slut[nibble] = (pol A, pol B) levels of one time slot: the byte layout
    // is [A_t, B_t, A_t+1, B_t+1] (fxkernel.cpp:15-28,443-476), so a time
    // slot is the nibble (byte >> 4*slot) & 15 with A in its low 2 bits.
    // 16 float2 entries occupy 32 consecutive dwords, one per LDS bank:
    // lanes reading different entries never share a bank.
    if (threadIdx.x < 16) {
      const float lv[4] = {-kHiMag, -1.0f, 1.0f, kHiMag};
      slut[threadIdx.x] = make_float2(lv[threadIdx.x & 3], lv[threadIdx.x >> 2]);
    }
  }

  int offset = 0;
  double frac = 0.0, a = 0.0, b = 0.0;
  const bool valid = active && split_offset<L>(p, st, t, offset, frac, a, b);

  float2 *sub = stab[li];
  float2 *step = sub + S2;
  float2 *fsub = step + NSTEPS;
  float2 *fstep = fsub + S;
  if (valid) {
#pragma clang fp contract(off)
    // fxkernel.cpp:503-505: remove whole seconds of delay.
    const double intdelay = trunc(b);
    const double bf = b - intdelay;
    const float ffrac = (float)frac;  // ippsMulC_32f takes an Ipp32f (fxkernel.cpp:600)
    for (int e = j; e < G::NTAB; e += T) {
      float sn, cs;
      if (e < S2) {
        // subphase, fxkernel.cpp:107,508,514,519,525-530,535
        double ph = ((double)e / (double)L) * a;
        ph = ph + bf;
        ph = ph * p.lo;
        if (p.fraclo) ph = ph + p.lofrac * intdelay;
        const double arg = (-2.0 * kPi) * (ph - trunc(ph));
        sincosf((float)arg, &sn, &cs);
      } else if (e < S2 + NSTEPS) {
        // stepphase, fxkernel.cpp:112,511,522,539
        const int i = e - S2;
        double ph = ((double)(i * p.stepmul) / (double)L) * a;
        ph = ph * p.lo;
        const double arg = (-2.0 * kPi) * (ph - trunc(ph));
        sincosf((float)arg, &sn, &cs);
      } else if (e < S2 + NSTEPS + S) {
        // fxkernel.cpp:600,608
        sincosf(p.subchanfreq[e - S2 - NSTEPS] * ffrac, &sn, &cs);
      } else {
        // fxkernel.cpp:603,611
        sincosf(p.stepchanfreq[e - S2 - NSTEPS - S] * ffrac, &sn, &cs);
      }
      stab[li][e] = make_float2(cs, sn);
    }
  }
  __syncthreads();

  float2 va[16], vb[16];
  float2 *sa = sdata[li][0];
  float2 *sb = sdata[li][1];
  const float2 *__restrict__ tw = p.twiddle;

  // ---- pass 0: unpack + rotate from global memory, radix-16, no twiddles ----
  if (valid) {
    // Station buffers are device (global) memory; saying so lets the loads
    // below use the global rather than the flat address path.
    typedef const __attribute__((address_space(1))) uint8_t gbyte;
    gbyte *src = (gbyte *)p.station[st];
    // Sample n = j + k*L/16 of the window is packed sample offset+n.  L/16
    // is even, so every k shares the byte parity of offset+j: one base
    // pointer, compile-time byte offsets k*L/32, one nibble shift.
    const int s0 = offset + j;
    gbyte *bsrc = src + (CPLX ? 4 * (size_t)s0 : (size_t)(s0 >> 1));
    const int shift = (s0 & 1) * 4;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int n = j + k * (L / 16);
      const float2 r = cmul(sub[n % S2], step[n / S2]);  // complexrotator[n], fxkernel.cpp:553-558
      if constexpr (!CPLX) {
        const float2 x = slut[(bsrc[k * (L / 32)] >> shift) & 15];
        va[k] = make_float2(x.x * r.x, x.x * r.y);
        vb[k] = make_float2(x.y * r.x, x.y * r.y);
      } else {
        // [Are, Aim, Bre, Bim] int8, little-endian in one 32-bit word
        const int wv = *reinterpret_cast<const __attribute__((address_space(1))) int *>(
            bsrc + 4 * (size_t)k * (L / 16));
        const float are = (float)(int8_t)(wv & 0xff), aim = (float)(int8_t)((wv >> 8) & 0xff);
        const float bre = (float)(int8_t)((wv >> 16) & 0xff), bim = (float)(int8_t)(wv >> 24);
        va[k] = cmul(r, make_float2(are, aim));
        vb[k] = cmul(r, make_float2(bre, bim));
      }
    }
    dft<16>(va);
    dft<16>(vb);
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++) va[k] = vb[k] = make_float2(0.f, 0.f);
  }
  if constexpr (G::NPASS > 1) {
    pass_to_lds<L, CPLX, 0>(va, vb, sa, sb, j);
    __syncthreads();
    middle_passes<L, CPLX, 1>(va, vb, sa, sb, j, tw);
    pass_from_lds<L, CPLX, G::NPASS - 1>(va, vb, sa, sb, j, tw);
  }

  // ---- last pass outputs: bins q = jj + k*(L/R), keep q < nchan ----
  if (active) {
    constexpr int PL = G::NPASS - 1;
    constexpr int R = G::radix(PL), NB = 16 / R;
    constexpr int KEEP = CPLX ? R : R / 2;
    const size_t rowA = ((size_t)tl * p.nant + st) * 2;  // output row of pol A; pol B follows
#pragma unroll
    for (int bf = 0; bf < NB; bf++) {
      const int jj = j + bf * T;
#pragma unroll
      for (int k = 0; k < KEEP; k++) {
        const int q = jj + k * (L / R);
        float2 ya = make_float2(0.f, 0.f), yb = ya;
        if (valid) {
          const float2 fr = cmul(fsub[q % S], fstep[q / S]);  // fracsamprotator, fxkernel.cpp:614-625
          ya = cmul(fr, va[bf * R + k]);
          yb = cmul(fr, vb[bf * R + k]);
        }
        if constexpr (OUT == OUT_F32) {
          p.spectra[rowA * G::NCHAN + q] = ya;
          p.spectra[(rowA + 1) * G::NCHAN + q] = yb;
        } else if constexpr (OUT == OUT_POWER) {
          // one row per sampled FFT; the scale kernel sums rows in a fixed
          // order, so the scales do not depend on scheduling
          p.power[rowA * G::NCHAN + q] = ya.x * ya.x + ya.y * ya.y;
          p.power[(rowA + 1) * G::NCHAN + q] = yb.x * yb.x + yb.y * yb.y;
        } else {
          const float sa = p.scales[(size_t)(2 * st) * G::NCHAN + q];
          const float sb = p.scales[(size_t)(2 * st + 1) * G::NCHAN + q];
          if constexpr (OUT == OUT_FP8) {
            uint16_t *o = static_cast<uint16_t *>(p.out);
            o[rowA * G::NCHAN + q] = pack_fp8(ya.x * sa, ya.y * sa);
            o[(rowA + 1) * G::NCHAN + q] = pack_fp8(yb.x * sb, yb.y * sb);
          } else {
            uint32_t *o = static_cast<uint32_t *>(p.out);
            if constexpr (OUT == OUT_F16) {
              o[rowA * G::NCHAN + q] = pack_f16(ya.x * sa, ya.y * sa);
              o[(rowA + 1) * G::NCHAN + q] = pack_f16(yb.x * sb, yb.y * sb);
            } else {
              o[rowA * G::NCHAN + q] = pack_bf16(ya.x * sa, ya.y * sa);
              o[(rowA + 1) * G::NCHAN + q] = pack_bf16(yb.x * sb, yb.y * sb);
            }
          }
        }
      }
    }
    if (OUT != OUT_POWER && j == 0) p.valid[(size_t)st * p.numffts + t] = valid ? 1 : 0;
  }
}

template <int L, bool CPLX>
hipError_t launch_one(const FParams &p, hipStream_t s) {
  using G = FG<L, CPLX>;
  const long long items = (long long)p.nant * p.nt;
  const long long blocks = (items + G::IPW - 1) / G::IPW;
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks), block(G::BLOCK);
  switch (p.out_mode) {
    case OUT_F32: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_F32>), grid, block, 0, s, p); break;
    case OUT_F16: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_F16>), grid, block, 0, s, p); break;
    case OUT_BF16: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_BF16>), grid, block, 0, s, p); break;
    case OUT_FP8: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_FP8>), grid, block, 0, s, p); break;
    case OUT_POWER: hipLaunchKernelGGL((fstage_kernel<L, CPLX, OUT_POWER>), grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

bool fstage_supported(int L, bool cplx) {
  (void)cplx;
  return L >= 64 && L <= 8192 && (L & (L - 1)) == 0;
}

hipError_t launch_fstage(const FParams &p, int L, bool cplx, hipStream_t s) {
#define GXC_F_CASE(LL)                                              \
  case LL:                                                          \
    return cplx ? launch_one<LL, true>(p, s) : launch_one<LL, false>(p, s);
  switch (L) {
    GXC_F_CASE(64)
    GXC_F_CASE(128)
    GXC_F_CASE(256)
    GXC_F_CASE(512)
    GXC_F_CASE(1024)
    GXC_F_CASE(2048)
    GXC_F_CASE(4096)
    GXC_F_CASE(8192)
    default:
      return hipErrorInvalidValue;
  }
#undef GXC_F_CASE
}

}  // namespace gxc
