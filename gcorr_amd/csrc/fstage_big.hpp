// The register/plane F-stage as a device function over one (station, FFT,
// polarisation), shared by the per-subband launch (fstage_big.hip) and the
// multi-subband launch (fstage_multi.hip).
//   - L/32 threads, 32 points per thread in registers, radix-32 Stockham
//     passes (then one smaller pass when log2 L is not a multiple of 5);
//   - between passes the points go through LDS one component plane at a
//     time (real parts, then imaginary parts): L floats plus one pad per 32;
//   - the two polarisations of one (station, FFT) are neighbouring units, so
//     the second reads the shared input words from L2.
#pragma once

#include "fstage_common.hpp"

namespace gxc {
namespace fs {

template <int L, bool CPLX>
struct BG {
  using TG = TabGeom<L, CPLX>;
  static constexpr int LG = ilog2c(L);
  static constexpr int PT = 32;     // points per thread
  static constexpr int T = L / PT;  // threads per item = workgroup size
  static constexpr int N32 = LG / 5;
  static constexpr int REM = 1 << (LG % 5);
  static constexpr int NPASS = N32 + (REM > 1 ? 1 : 0);
  static constexpr int radix(int p) { return p < N32 ? 32 : REM; }
  static constexpr int ns(int p) {
    int n = 1;
    for (int i = 0; i < p; i++) n *= radix(i);
    return n;
  }
  static constexpr int PADL = L + L / 32;
  static_assert(L >= 2048 && L <= 32768, "register/plane F-stage covers L = 2048 .. 32768");
  static_assert(T <= 1024, "one workgroup per FFT");
};

__device__ __forceinline__ int padb(int i) { return i + (i >> 5); }

// padb(i + k*STRIDE) - padb(i) for STRIDE a multiple of 32 (any i), or for
// STRIDE == 1 when i is a multiple of 32 and k < 32: a compile-time offset.
template <int STRIDE>
__device__ __forceinline__ constexpr int padb_step(int k) {
  return STRIDE == 1 ? k : k * (STRIDE / 32) * 33;
}

// Store one component of the outputs of pass P (Stockham order).
template <int L, bool CPLX, int P, int COMP>
__device__ __forceinline__ void put_plane(const float2 (&v)[32], float *s, int j) {
  using G = BG<L, CPLX>;
  constexpr int R = G::radix(P), NS = G::ns(P), NB = 32 / R, T = G::T;
#pragma unroll
  for (int bf = 0; bf < NB; bf++) {
    const int jj = j + bf * T;
    const int base = (jj / NS) * NS * R + (jj % NS);
    static_assert(NS % 32 == 0 || (NS == 1 && R == 32), "compile-time padded offsets");
    float *d = s + padb(base);
#pragma unroll
    for (int k = 0; k < R; k++) d[padb_step<NS>(k)] = COMP ? v[bf * R + k].y : v[bf * R + k].x;
  }
}

// Load one component of the inputs of pass P.
template <int L, bool CPLX, int P, int COMP>
__device__ __forceinline__ void get_plane(float2 (&v)[32], const float *s, int j) {
  using G = BG<L, CPLX>;
  constexpr int R = G::radix(P), NB = 32 / R, T = G::T, STR = L / R;
#pragma unroll
  for (int bf = 0; bf < NB; bf++) {
    const int jj = j + bf * T;
    static_assert(STR % 32 == 0, "compile-time padded offsets");
    const float *d = s + padb(jj);
#pragma unroll
    for (int k = 0; k < R; k++) {
      const float x = d[padb_step<STR>(k)];
      if (COMP)
        v[bf * R + k].y = x;
      else
        v[bf * R + k].x = x;
    }
  }
}

// v[k] *= w^k for k = 1 .. R-1, w = W_L^m, table entry m of the byte
// offset mb (entries m*k < L).  Only w^1 .. w^3 and w^(4 * 2^i) come from
// the table (in L2: at 8192 points the whole table is 64 KiB, and loading
// every twiddle cost a third of the F-stage); w^(4a) is the product of the
// loaded powers of the bits of a and w^(4a + b) = w^(4a) w^b, at most three
// FP32 products (a few ulps) from table entries.
template <int R>
__device__ __forceinline__ void twiddle_powers(float2 *v, const unsigned char *twb, unsigned mb) {
  auto ld = [&](unsigned k) { return *reinterpret_cast<const float2 *>(twb + mb * k); };
  float2 p[4];
#pragma unroll
  for (int b = 1; b < 4 && b < R; b++) {
    p[b] = ld((unsigned)b);
    v[b] = cmul(v[b], p[b]);
  }
  if constexpr (R > 4) {
    constexpr int NQ = ilog2c(R) - 2;  // w^4, w^8, .., w^(R/2)
    float2 q[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) q[i] = ld(4u << i);
#pragma unroll
    for (int a = 1; 4 * a < R; a++) {
      float2 w4 = make_float2(1.f, 0.f);
      bool first = true;
#pragma unroll
      for (int i = 0; i < NQ; i++)
        if (a & (1 << i)) {
          w4 = first ? q[i] : cmul(w4, q[i]);
          first = false;
        }
      v[4 * a] = cmul(v[4 * a], w4);
#pragma unroll
      for (int b = 1; b < 4; b++)
        if (4 * a + b < R) v[4 * a + b] = cmul(v[4 * a + b], cmul(w4, p[b]));
    }
  }
}

// Twiddles and butterflies of pass P >= 1.
template <int L, bool CPLX, int P>
__device__ __forceinline__ void butterflies(float2 (&v)[32], int j, const float2 *__restrict__ tw) {
  using G = BG<L, CPLX>;
  constexpr int R = G::radix(P), NS = G::ns(P), NB = 32 / R, T = G::T;
#pragma unroll
  for (int bf = 0; bf < NB; bf++) {
    const int jj = j + bf * T;
    // table entry m as a 32-bit byte offset from the table's uniform base
    const unsigned mb = (unsigned)(jj % NS) * (unsigned)(L / (NS * R)) * (unsigned)sizeof(float2);
    twiddle_powers<R>(&v[bf * R], reinterpret_cast<const unsigned char *>(tw), mb);
    dft<R>(&v[bf * R]);
  }
}

// Passes P .. NPASS-1: exchange through LDS, then butterflies.
template <int L, bool CPLX, int P>
__device__ __forceinline__ void later_passes(float2 (&v)[32], float *s, int j, const float2 *__restrict__ tw) {
  using G = BG<L, CPLX>;
  if constexpr (P < G::NPASS) {
    put_plane<L, CPLX, P - 1, 0>(v, s, j);
    __syncthreads();
    get_plane<L, CPLX, P, 0>(v, s, j);
    __syncthreads();
    put_plane<L, CPLX, P - 1, 1>(v, s, j);
    __syncthreads();
    get_plane<L, CPLX, P, 1>(v, s, j);
    __syncthreads();
    butterflies<L, CPLX, P>(v, j, tw);
    later_passes<L, CPLX, P + 1>(v, s, j, tw);
  }
}

// LDS of one unit: one component plane of L points (+1 pad per 32) and
// the rotator tables, in float2 slots.
template <int L, bool CPLX>
struct BLds {
  static constexpr int PLANE2 = (BG<L, CPLX>::PADL + 1) / 2;  // plane, in float2 slots
  static constexpr int N = PLANE2 + TabGeom<L, CPLX>::NTAB + 2;  // + 4-float 2-bit LUT
};

// Unit u = 2 * (station + nant * FFT) + pol of one launch (BOTH false), or
// u = station + nant * FFT with the two polarisations in turn (BOTH true:
// the delay split, the rotator tables and the input words, which hold both
// polarisations' samples, then serve two FFTs), on a workgroup of exactly
// BG::T threads (every thread calls it with the same u: it contains
// __syncthreads).  smem holds BLds::N float2 slots.
template <int L, bool CPLX, int OUT, bool BOTH = false>
__device__ __forceinline__ void fstage_big_item(const FParams &p, int u, float2 *smem) {
  using G = BG<L, CPLX>;
  using TG = typename G::TG;
  constexpr int T = G::T, S = TG::S, S2 = TG::S2, NSTEPS = TG::NSTEPS;
  static_assert(T % S2 == 0, "pass-0 points of one thread share one sub-rotator entry");
  float *splane = reinterpret_cast<float *>(smem);
  float2 *stab = smem + BLds<L, CPLX>::PLANE2;

  const int j = threadIdx.x;
  const int pol0 = BOTH ? 0 : u & 1;
  const int item = BOTH ? u : u >> 1;
  const int tl = item / p.nant;
  const int st = item % p.nant;
  const int t = p.t0 + tl * p.tstride;
  const size_t row0 = ((size_t)tl * p.nant + st) * 2;  // output row of pol 0

  int offset = 0;
  double frac = 0.0, a = 0.0, b = 0.0;
  // The same for every thread of the workgroup: the branches below that
  // depend on it are uniform, so every __syncthreads is reached by all.
  const bool valid = split_offset<L>(p, st, t, offset, frac, a, b);
  if (j == 0 && pol0 == 0 && OUT != OUT_POWER) p.valid[(size_t)st * p.numffts + t] = valid ? 1 : 0;
  if (!valid) {
    for (int pol = pol0; pol < (BOTH ? 2 : pol0 + 1); pol++)
      for (int q = j; q < TG::NCHAN; q += T)
        store_bin<OUT, TG::NCHAN>(p, row0 + pol, 2 * st + pol, q, make_float2(0.f, 0.f));
    return;
  }
  // This thread's pass-0 input words, requested before the table build so
  // that their latency overlaps it: complex 8-bit [Are, Aim, Bre, Bim] as
  // one 32-bit word per sample; 2-bit as the aligned word that holds the
  // sample's byte (byte layout [A_t, B_t, A_t+1, B_t+1], fxkernel.cpp:443-476).  Pass-0 point k
  // of thread j is window sample n = j + k*T, packed sample offset + n.
  gbyte *src = (gbyte *)p.station[st];
  const int s0 = offset + j;
  int rawin[32];
  // T is even, so every k shares the nibble parity of s0.  Each byte comes
  // in its aligned 32-bit word (never past the page of a valid byte): the
  // field extract below then needs no byte mask, one VALU less per sample.
  gint *wsrc = CPLX ? reinterpret_cast<gint *>(src + 4 * (size_t)s0)
                    : (gint *)((uintptr_t)(src + (s0 >> 1)) & ~(uintptr_t)3);
  constexpr int WSTRIDE = CPLX ? T : T / 8;  // words between points k and k+1
  auto load_words = [&]() {
#pragma unroll
    for (int k = 0; k < 32; k++) rawin[k] = wsrc[(size_t)k * WSTRIDE];
  };
  load_words();
  fill_tables<L, CPLX>(p, frac, a, b, stab, j, T);
  // 2-bit levels by code (fxkernel.cpp:15-28): an LDS read replaces the
  // compare/select chain that would otherwise cost ~6 VALU per sample
  float *lut = reinterpret_cast<float *>(stab + TabGeom<L, CPLX>::NTAB);
  if (!CPLX && j < 4) lut[j] = j == 0 ? -kHiMag : j == 1 ? -1.0f : j == 2 ? 1.0f : kHiMag;
  __syncthreads();

  auto one_pol = [&](const int pol) {
    // An opaque copy of the thread index (and an opaque zero offset) in
    // each polarisation's copy of this code keeps the compiler from sharing
    // the per-thread LDS and global addresses, the rotator, twiddle and
    // table reads and their products between the two copies, where they
    // would hold ~100 registers over the first FFT.
    int z = 0, j = threadIdx.x;
    if constexpr (BOTH) asm volatile("" : "+s"(z), "+v"(j));
    const float2 *sub = stab + z;
    const float2 *step = sub + S2;
    const float2 *fsub = step + NSTEPS;
    const float2 *fstep = fsub + S;
    const float2 *__restrict__ tw = p.twiddle + z;

    // Pass-0 point k of thread j is window sample n = j + k*T: its rotator
    // is sub[n % S2] * step[n / S2] (fxkernel.cpp:553-558) = sub[j % S2] *
    // step[j / S2 + k*T/S2].
    const float2 rsub = sub[j % S2];
    const size_t row = row0 + pol;
    float2 v[32];
    if constexpr (CPLX) {
      const int sh = 16 * pol;
#pragma unroll
      for (int k = 0; k < 32; k++) {
        const int wv = rawin[k];
        const float re = (float)(int8_t)((wv >> sh) & 0xff), im = (float)(int8_t)((wv >> (sh + 8)) & 0xff);
        v[k] = cmul(cmul(rsub, step[j / S2 + k * (T / S2)]), make_float2(re, im));
      }
    } else {
      // sample s is nibble s&1 of its byte, the byte (address & 3) of its
      // word; pol A in the nibble's low 2 bits
      const int sh = 8 * (int)((uintptr_t)(src + (s0 >> 1)) & 3) + (s0 & 1) * 4 + 2 * pol;
#pragma unroll
      for (int k = 0; k < 32; k++) {
        const int code = (int)__builtin_amdgcn_ubfe((unsigned)rawin[k], (unsigned)sh, 2u);
        const float x = lut[code];
        const float2 r = cmul(rsub, step[j / S2 + k * (T / S2)]);
        v[k] = make_float2(x * r.x, x * r.y);
      }
    }
#pragma unroll
    for (int bf = 0; bf < 32 / G::radix(0); bf++) dft<G::radix(0)>(&v[bf * G::radix(0)]);
    later_passes<L, CPLX, 1>(v, splane, j, tw);

    // last pass outputs: bins q = jj + k*(L/R), keep q < nchan
    constexpr int PL = G::NPASS - 1;
    constexpr int R = G::radix(PL), NB = 32 / R;
    constexpr int KEEP = CPLX ? R : R / 2;
#pragma unroll
    for (int bf = 0; bf < NB; bf++) {
      const int jj = j + bf * T;
#pragma unroll
      for (int k = 0; k < KEEP; k++) {
        const int q = jj + k * (L / R);
        // q % S = j % S and q / S = j / S + a compile-time step when S
        // divides T and L/R, which keeps the table offsets out of the VALU
        const float2 fr = (T % S == 0 && (L / R) % S == 0)
                              ? cmul(fsub[j % S], fstep[j / S + (bf * T + k * (L / R)) / S])
                              : cmul(fsub[q % S], fstep[q / S]);  // fracsamprotator, fxkernel.cpp:614-625
        store_bin<OUT, TG::NCHAN>(p, row, 2 * st + pol, q, cmul(fr, v[bf * R + k]));
      }
    }
  };
  one_pol(pol0);
  if constexpr (BOTH) {
    // the second polarisation reads its words again (from L2): keeping
    // them in registers over the first FFT would cost occupancy.  The last
    // pass of the first FFT read the plane before a barrier, so this FFT's
    // first plane store is safe.
    load_words();
    one_pol(1);
  }
}

}  // namespace fs
}  // namespace gxc
