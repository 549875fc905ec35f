// fft_core.hpp -- register-resident Stockham FFT building blocks for gfx950.
//
// One transform of length L (a power of two, 1..16384) is owned by T = L/E
// threads, each holding E = min(L,16) complex values in registers.  Thread t
// always holds elements t + e*T (e = 0..E-1) of the current stage's input, so
// the first stage loads straight from HBM with consecutive lanes on consecutive
// addresses, and the last stage stores final frequency bins
//   bin = j + q*(L/R)    (j = butterfly index, q = butterfly output)
// which is again lane-contiguous.  Between stages the data is exchanged
// through LDS, or across the lanes of one wave for short rows (fft_wave.hpp).
// The stage schedule is one radix-2/4/8 stage (when log2 L is
// not a multiple of 4) followed by radix-16 stages; every stage is a Stockham
// DIT step:
//   butterfly j, k = j mod Ls, inputs x[j + m*L/R] * w_{Ls*R}^{m*k},
//   R-point DFT, output q to (j-k)*R + k + q*Ls.
// This is the same decomposition as glsCV's radix-4 shader
// (glsCV/glsFft.cpp:247-339: inputs at insertZeroBits(n,q,2)+k*2^q, outputs at
// n + k*N/4, twiddle index (n>>q)<<q), generalised to radix 16 and to a
// register/LDS exchange instead of one texture round trip per stage.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace glsx {

// Pointers that live in device memory (not kernel arguments) are not known to
// address global memory, and the compiler falls back to flat instructions
// with 64-bit per-lane addresses.  Casting them to address space 1 keeps the
// loads and stores global (SGPR base + VGPR offset).
// (The host compilation pass of this file never executes device code; there
// the qualifier is left off.)
#if defined(__HIP_DEVICE_COMPILE__)
#define GLSX_GLOBAL __attribute__((address_space(1)))
#else
#define GLSX_GLOBAL
#endif
template <class T>
__device__ __forceinline__ GLSX_GLOBAL T* gptr(T* p) {
  return (GLSX_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ const GLSX_GLOBAL T* gptr(const T* p) {
  return (const GLSX_GLOBAL T*)p;
}

__host__ __device__ constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x >> 1); }

// EP: elements per thread for long transforms (16, or 32 for the column
// kernels whose two transforms per thread-group halve the thread count).
template <int L, int EP = 16>
struct Sched {
  static constexpr int LOG = ilog2c(L);
  static constexpr int E = L < EP ? L : EP;         // elements per thread
  static constexpr int T = L / E;                   // threads per transform
  static constexpr int R0 = (L < 16) ? L : ((LOG % 4) == 0 ? 16 : (1 << (LOG % 4)));
  static constexpr int NSTAGE = (L < 16) ? 1 : (LOG % 4 == 0 ? LOG / 4 : LOG / 4 + 1);
  static constexpr int radix(int s) { return s == 0 ? R0 : 16; }
  static constexpr int ls(int s) { return s == 0 ? 1 : R0 * (1 << (4 * (s - 1))); }
  // LDS words (float2) one transform occupies with one pad per 16 elements.
  static constexpr int LP = L + L / 16 + (L >= 16 ? 1 : 0);
};

// Offset of the twiddle table of length P (P = 2,4,..) in the concatenated
// table: tables for P = 2,4,8,... are stored back to back, so table P starts
// at P - 2.  Entry i of table P is exp(-2*pi*i*I/P) (generated in FP64).
__host__ __device__ constexpr int tw_offset(int P) { return P - 2; }
constexpr int kMaxLen = 16384;
constexpr int kTwiddleEntries = 2 * kMaxLen - 2;

// LDS swizzle: one pad slot per 16 elements keeps the strided butterfly
// writes of every stage free of bank conflicts.
__device__ __forceinline__ int swz(int pos) { return pos + (pos >> 4); }

// ---------------------------------------------------------------------------
// complex helpers (float2 = re, im)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// Fused multiply-adds are spelled out and the library is compiled with
// -ffp-contract=off: every kernel that runs a tile (one launch per pass, the
// half-LDS column kernel, the wave-exchange row tiles, the plan kernels) then
// rounds identically, so a transform's bits do not depend on which launch
// shape computed it.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// a * (-i)
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// multiply by w_16^e = exp(-2*pi*i*e/16) for the constant exponents a 4x4
// split of the 16-point DFT needs.
template <int e>
__device__ __forceinline__ float2 mul_w16(float2 a) {
  constexpr float C1 = 0.92387953251128675613f;  // cos(pi/8)
  constexpr float S1 = 0.38268343236508977173f;  // sin(pi/8)
  constexpr float R2 = 0.70710678118654752440f;  // sqrt(2)/2
  if constexpr (e == 0) return a;
  else if constexpr (e == 1) return cmul(a, make_float2(C1, -S1));
  else if constexpr (e == 2) return make_float2(R2 * (a.x + a.y), R2 * (a.y - a.x));
  else if constexpr (e == 3) return cmul(a, make_float2(S1, -C1));
  else if constexpr (e == 4) return mul_mi(a);
  else if constexpr (e == 6) return make_float2(R2 * (a.y - a.x), -R2 * (a.x + a.y));
  else if constexpr (e == 9) return cmul(a, make_float2(-C1, S1));
  else static_assert(e == 0, "unsupported w16 exponent");
}

// ---------------------------------------------------------------------------
// small forward DFTs, natural order in and out, operating on a[0..R-1]
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dft2(float2& a0, float2& a1) {
  float2 t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_mi(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <int R>
__device__ __forceinline__ void dft(float2* a);

template <>
__device__ __forceinline__ void dft<1>(float2*) {}

template <>
__device__ __forceinline__ void dft<2>(float2* a) { dft2(a[0], a[1]); }

template <>
__device__ __forceinline__ void dft<4>(float2* a) { dft4(a[0], a[1], a[2], a[3]); }

// 8 = 4 (over m = n2 + 2*n1, n1) x 2 (over n2), output k1 + 4*k2
template <>
__device__ __forceinline__ void dft<8>(float2* a) {
  constexpr float R2 = 0.70710678118654752440f;
  float2 b0[4] = {a[0], a[2], a[4], a[6]};
  float2 b1[4] = {a[1], a[3], a[5], a[7]};
  dft4(b0[0], b0[1], b0[2], b0[3]);
  dft4(b1[0], b1[1], b1[2], b1[3]);
  // twiddle w8^{k1} on the n2 = 1 branch
  b1[1] = make_float2(R2 * (b1[1].x + b1[1].y), R2 * (b1[1].y - b1[1].x));
  b1[2] = mul_mi(b1[2]);
  b1[3] = make_float2(R2 * (b1[3].y - b1[3].x), -R2 * (b1[3].x + b1[3].y));
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    a[k1] = cadd(b0[k1], b1[k1]);
    a[k1 + 4] = csub(b0[k1], b1[k1]);
  }
}

// 16 = 4 (over m = n2 + 4*n1, n1) x 4 (over n2), output k1 + 4*k2
template <>
__device__ __forceinline__ void dft<16>(float2* a) {
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4(a[n2], a[n2 + 4], a[n2 + 8], a[n2 + 12]);
  // now a[n2 + 4*k1] = b[n2][k1]; apply w16^{n2*k1}
  a[5] = mul_w16<1>(a[5]);
  a[9] = mul_w16<2>(a[9]);
  a[13] = mul_w16<3>(a[13]);
  a[6] = mul_w16<2>(a[6]);
  a[10] = mul_w16<4>(a[10]);
  a[14] = mul_w16<6>(a[14]);
  a[7] = mul_w16<3>(a[7]);
  a[11] = mul_w16<6>(a[11]);
  a[15] = mul_w16<9>(a[15]);
  float2 o[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 c0 = a[0 + 4 * k1], c1 = a[1 + 4 * k1], c2 = a[2 + 4 * k1], c3 = a[3 + 4 * k1];
    dft4(c0, c1, c2, c3);
    o[k1] = c0;
    o[k1 + 4] = c1;
    o[k1 + 8] = c2;
    o[k1 + 12] = c3;
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = o[i];
}

// The butterfly of a stage, as a policy the stage functions below take: the
// FP32 DFTs above.  (The split-precision passes substitute matrix-core
// products for the 16-point DFT, mfma_core.hpp.)
struct DftFp32 {
  template <int R>
  __device__ __forceinline__ void run(float2 (&a)[R]) const {
    dft<R>(a);
  }
};

// ---------------------------------------------------------------------------
// Twiddles of one stage: w^{m*k}, m = 1..R-1, w = exp(-2*pi*i/P), P = Ls*R.
// Built from a few "base" powers read from a table (radix 16 and 8: w^k and
// w^{4k}; radix 4: w^k, w^{2k}, w^{3k}; radix 2: w^k); the other powers are
// products of at most three factors, so each twiddle stays within a few ulp of
// the FP64 value.  The bases are read from the global tables (256 KiB for all
// lengths, resident in L2).
// ---------------------------------------------------------------------------
template <int R>
__host__ __device__ constexpr int tw_bases() { return R >= 8 ? 2 : (R == 4 ? 3 : (R == 2 ? 1 : 0)); }

template <int R>
__device__ __forceinline__ void twiddles_from_bases(const float2* b, float2* w) {
  if constexpr (R == 2) {
    w[1] = b[0];
  } else if constexpr (R == 4) {
    w[1] = b[0];
    w[2] = b[1];
    w[3] = b[2];
  } else if constexpr (R == 8) {
    float2 w1 = b[0], w4 = b[1];
    float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
    w[1] = w1; w[2] = w2; w[3] = w3; w[4] = w4;
    w[5] = cmul(w4, w1); w[6] = cmul(w4, w2); w[7] = cmul(w4, w3);
  } else if constexpr (R == 16) {
    float2 w1 = b[0], w4 = b[1];
    float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
    float2 w8 = cmul(w4, w4), w12 = cmul(w8, w4);
    w[1] = w1; w[2] = w2; w[3] = w3;
    w[4] = w4; w[8] = w8; w[12] = w12;
#pragma unroll
    for (int m1 = 1; m1 < 4; ++m1) {
      float2 hi = m1 == 1 ? w4 : (m1 == 2 ? w8 : w12);
      w[4 * m1 + 1] = cmul(hi, w1);
      w[4 * m1 + 2] = cmul(hi, w2);
      w[4 * m1 + 3] = cmul(hi, w3);
    }
  }
}

// Exponent multipliers of the base powers: w^{mult*k}.
template <int R>
__host__ __device__ constexpr int tw_base_mult(int i) {
  return R >= 8 ? (i == 0 ? 1 : 4) : (i + 1);
}

// Run butterflies of stage s for thread t, in place in v: after the call,
// output q of butterfly u is in v[u + q*(E/R)].
template <int L, int S, int EP = 16, class D = DftFp32>
__device__ __forceinline__ void stage_butterflies(float2 (&v)[(Sched<L, EP>::E)], int t, const float2* __restrict__ tw,
                                                  const D& d = D{}) {
  constexpr int E = Sched<L, EP>::E, T = Sched<L, EP>::T;
  constexpr int R = Sched<L, EP>::radix(S), LS = Sched<L, EP>::ls(S);
  constexpr int NB = E / R;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    float2 a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = v[u + m * NB];
    if constexpr (LS > 1) {
      const int j = t + u * T;
      const int k = j & (LS - 1);
      constexpr int NBASE = tw_bases<R>();
      float2 b[NBASE > 0 ? NBASE : 1];
      const auto* tab = gptr(tw) + tw_offset(LS * R);
#pragma unroll
      for (int i = 0; i < NBASE; ++i) b[i] = tab[tw_base_mult<R>(i) * k];
      float2 w[R];
      twiddles_from_bases<R>(b, w);
#pragma unroll
      for (int m = 1; m < R; ++m) a[m] = cmul(a[m], w[m]);
    }
    d.template run<R>(a);
#pragma unroll
    for (int q = 0; q < R; ++q) v[u + q * NB] = a[q];
  }
}

// The twiddle bases stage S reads for butterfly j (the table reads of
// stage_butterflies), so that a kernel can issue them before it needs them.
template <int L, int S, int EP = 16>
__device__ __forceinline__ void stage_bases(int j, const float2* __restrict__ tw,
                                            float2 (&b)[tw_bases<Sched<L, EP>::radix(S)>()]) {
  constexpr int R = Sched<L, EP>::radix(S), LS = Sched<L, EP>::ls(S);
  static_assert(LS > 1, "stage 0 has no twiddles");
  const int k = j & (LS - 1);
  const auto* tab = gptr(tw) + tw_offset(LS * R);
#pragma unroll
  for (int i = 0; i < tw_bases<R>(); ++i) b[i] = tab[tw_base_mult<R>(i) * k];
}

// Stages whose twiddle bases a kernel reads up front: every radix-16 stage
// after the first when a thread holds one butterfly (E == R).  Read after
// the LDS barrier that precedes the stage, their L2 latency stalls every
// wave of the workgroup at once (4096-point row pass: 45.0-46.0 us with the
// reads in place, 40.3 us with them issued at the start).
template <int L, int S, int EP = 16>
__host__ __device__ constexpr bool bases_early() {
  return S > 0 && Sched<L, EP>::E == Sched<L, EP>::radix(S);
}

// The bases of every bases_early() stage of butterfly j, read before the
// tile's data (a kernel that conjugates its input on load waits for the data
// there, and reads issued after that wait would add their latency).
template <int L, int EP = 16>
struct EarlyBases {
  static constexpr int NS = Sched<L, EP>::NSTAGE;
  float2 b[NS][2];
  __device__ __forceinline__ void load(int j, const float2* __restrict__ tw) {
    [&]<int... S>(std::integer_sequence<int, S...>) {
      (
          [&] {
            if constexpr (bases_early<L, S, EP>()) stage_bases<L, S, EP>(j, tw, b[S]);
          }(),
          ...);
    }(std::make_integer_sequence<int, NS>{});
  }
};

// Stage S for a thread with one butterfly (E == R), twiddle bases from
// stage_bases: the same operations as stage_butterflies.
template <int L, int S, int EP = 16, class D = DftFp32>
__device__ __forceinline__ void stage_butterflies_b(float2 (&v)[(Sched<L, EP>::E)],
                                                    const float2 (&b)[tw_bases<Sched<L, EP>::radix(S)>()],
                                                    const D& d = D{}) {
  constexpr int R = Sched<L, EP>::radix(S);
  static_assert(Sched<L, EP>::E == R, "one butterfly per thread");
  float2 a[R];
#pragma unroll
  for (int m = 0; m < R; ++m) a[m] = v[m];
  float2 w[R];
  twiddles_from_bases<R>(b, w);
#pragma unroll
  for (int m = 1; m < R; ++m) a[m] = cmul(a[m], w[m]);
  d.template run<R>(a);
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = a[q];
}

// Output position of element v[u + q*(E/R)] after stage S.
template <int L, int S, int EP = 16>
__device__ __forceinline__ int stage_out_pos(int t, int u, int q) {
  constexpr int T = Sched<L, EP>::T;
  constexpr int R = Sched<L, EP>::radix(S), LS = Sched<L, EP>::ls(S);
  const int j = t + u * T;
  const int k = j & (LS - 1);
  return (j - k) * R + k + q * LS;
}

}  // namespace glsx
