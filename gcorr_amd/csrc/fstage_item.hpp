// The LDS-resident F-stage (FFT lengths 64..8192) as a device function over
// one workgroup's items, shared by the per-subband launch (fstage.hip) and
// the persistent multi-subband launch (fstage_multi.hip).
//
// Per item (one station, one FFT), fused so the packed samples are read
// once and the spectra written once:
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
#pragma once

#include "fstage_common.hpp"

namespace gxc {
namespace fs {

// Compile-time geometry for FFT length L.
template <int L, bool CPLX>
struct FG {
  using TG = TabGeom<L, CPLX>;
  static constexpr int LG = ilog2c(L);
  static constexpr int NCHAN = TG::NCHAN, S = TG::S, S2 = TG::S2, NSTEPS = TG::NSTEPS, NTAB = TG::NTAB;
  static constexpr int T = L / 16;                           // threads per item
  static constexpr int DEFAULT_BLOCK = T >= 256 ? T : 256;  // workgroup size of the plain launch
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

// LDS of one workgroup of BLOCK threads at FFT length L: IPW = BLOCK/(L/16)
// items x 2 pols x (L + L/16) points, IPW rotator tables, the 2-bit LUT.
constexpr int flds_slots(int L, bool cplx, int block) {
  const int nchan = cplx ? L : L / 2;
  const int s = 1 << (ilog2c(nchan) / 2);
  const int s2 = cplx ? s : 2 * s;
  const int ntab = s2 + L / s2 + s + nchan / s;
  const int ipw = block / (L / 16);
  return ipw * 2 * (L + L / 16) + ipw * ntab + 16;
}

template <int L, bool CPLX, int BLOCK>
struct FLds {
  using G = FG<L, CPLX>;
  static_assert(BLOCK % G::T == 0, "whole items per workgroup");
  static constexpr int IPW = BLOCK / G::T;  // items per workgroup
  static constexpr int N = flds_slots(L, CPLX, BLOCK);  // float2 slots
  static_assert(N == IPW * 2 * G::PADL + IPW * G::NTAB + 16, "LDS layout");
};

// Items blk*IPW .. blk*IPW+IPW-1 of one F-stage launch (item = station +
// nant * FFT) on a workgroup of BLOCK threads.  Every thread of the
// workgroup calls it with the same blk (it contains __syncthreads).
template <int L, bool CPLX, int OUT, int BLOCK>
__device__ __forceinline__ void fstage_item(const FParams &p, int blk, float2 *smem) {
  using G = FG<L, CPLX>;
  using LD = FLds<L, CPLX, BLOCK>;
  constexpr int T = G::T, S = G::S, S2 = G::S2, NSTEPS = G::NSTEPS, IPW = LD::IPW;
  float2 (*sdata)[2][G::PADL] = reinterpret_cast<float2 (*)[2][G::PADL]>(smem);
  float2 (*stab)[G::NTAB] = reinterpret_cast<float2 (*)[G::NTAB]>(smem + IPW * 2 * G::PADL);
  float2 *slut = smem + IPW * 2 * G::PADL + IPW * G::NTAB;

  const int li = threadIdx.x / T;
  const int j = threadIdx.x % T;
  const int item = blk * IPW + li;
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
  if (valid) fill_tables<L, CPLX>(p, frac, a, b, stab[li], j, T);
  __syncthreads();

  float2 va[16], vb[16];
  float2 *sa = sdata[li][0];
  float2 *sb = sdata[li][1];
  const float2 *__restrict__ tw = p.twiddle;

  // ---- pass 0: unpack + rotate from global memory, radix-16, no twiddles ----
  if (valid) {
    // Station buffers are device (global) memory; saying so lets the loads
    // below use the global rather than the flat address path.
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
        const int wv = *reinterpret_cast<gint *>(bsrc + 4 * (size_t)k * (L / 16));
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
        store_bin<OUT, G::NCHAN>(p, rowA, 2 * st, q, ya);
        store_bin<OUT, G::NCHAN>(p, rowA + 1, 2 * st + 1, q, yb);
      }
    }
    if (OUT != OUT_POWER && j == 0) p.valid[(size_t)st * p.numffts + t] = valid ? 1 : 0;
  }
}

}  // namespace fs
}  // namespace gxc
