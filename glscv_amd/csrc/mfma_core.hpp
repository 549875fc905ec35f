// mfma_core.hpp -- the 16-point DFT on the gfx950 matrix cores with a 3-term
// BF16 split (the split-precision passes' butterfly).
//
// A radix-16 butterfly is y = F a with F[q][m] = exp(-2*pi*i*q*m/16).  Over
// the reals it is a 32x32 matrix M times the 32-vector (Re a; Im a):
//   M = [ C  S ; -S  C ],  C[q][m] = cos(2*pi*q*m/16), S[q][m] = sin(2*pi*q*m/16)
// (rows 0..15 give Re y, rows 16..31 give Im y).  One wave multiplies M by the
// inputs of 32 butterflies at once with v_mfma_f32_32x32x16_bf16, K = 16: one
// instruction for the Re half of M's columns, one for the Im half.  Operands
// are split into BF16 parts, x = hi + lo with hi = bf16(x), lo = bf16(x - hi),
// and the three products M_hi*A_hi + M_hi*A_lo + M_lo*A_hi accumulate in FP32:
// six MFMAs per 32 butterflies.  What the split drops is M_lo*A_lo plus the
// residue below lo, about 2^-17 relative per product.
//
// Operand layout of v_mfma_f32_32x32x16_bf16 (lane l, h = l / 32, n = l % 32):
//   A: lane holds A[n][8h + i], i = 0..7
//   B: lane holds B[8h + i][n]
//   D: lane holds D[8*(i/4) + 4h + i%4][n], i = 0..15
// With A = M and B = inputs (column n = butterfly n), lane (n, h) supplies
// inputs m = 8h..8h+7 of butterfly n and receives its outputs
// q = 4h..4h+3 and 8+4h..8+4h+3, Re in D[0..7] and Im in D[8..15].
#pragma once
#include "fft_wave.hpp"

namespace glsx {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

// x = hi + lo in BF16: hi = bf16(x), lo = bf16(x - hi), both rounded to
// nearest even by the hardware conversion (v_cvt_pk_bf16_f32); bf16 -> float
// is exact (a 16-bit shift).
__device__ __forceinline__ void split_bf16(float x, short& hi, short& lo) {
  const __bf16 h = (__bf16)x;
  const __bf16 l = (__bf16)(x - (float)h);
  hi = __builtin_bit_cast(short, h);
  lo = __builtin_bit_cast(short, l);
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// The same split for two values at once: one conversion instruction per
// part, the parts packed as the MFMA operands want them (element 0 low).
__device__ __forceinline__ void split_bf16x2(float x0, float x1, unsigned& hi, unsigned& lo) {
  const f32x2_t v = {x0, x1};
  const bf16x2_t h = __builtin_convertvector(v, bf16x2_t);
  const f32x2_t hf = __builtin_convertvector(h, f32x2_t);
  const f32x2_t r = {v.x - hf.x, v.y - hf.y};
  const bf16x2_t l = __builtin_convertvector(r, bf16x2_t);
  hi = __builtin_bit_cast(unsigned, h);
  lo = __builtin_bit_cast(unsigned, l);
}

// This lane's fragments of M, hi and lo parts, for the Re and Im column halves.
struct DftFrag {
  bf16x8_t hi[2], lo[2];
};

// cos(2*pi*r/16), r = 0..15
__constant__ const float kCos16[16] = {
    1.0f,  0.92387953251128675613f,  0.70710678118654752440f,  0.38268343236508977173f,
    0.0f,  -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
    -1.0f, -0.92387953251128675613f, -0.70710678118654752440f, -0.38268343236508977173f,
    0.0f,  0.38268343236508977173f,  0.70710678118654752440f,  0.92387953251128675613f};

__device__ __forceinline__ DftFrag dft16_fragments(int lane) {
  const float* C = kCos16;
  const int row = lane & 31, h = lane >> 5;
  const int q = row & 15;
  const bool im_row = row >= 16;
  DftFrag f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int m = 8 * h + i;
    const int r = (q * m) & 15;
    const float c = C[r], s = C[(r + 12) & 15];  // sin(x) = cos(x - pi/2)
    const float vre = im_row ? -s : c;           // multiplies Re(a_m)
    const float vim = im_row ? c : s;            // multiplies Im(a_m)
    short hi, lo;
    split_bf16(vre, hi, lo);
    f.hi[0][i] = hi;
    f.lo[0][i] = lo;
    split_bf16(vim, hi, lo);
    f.hi[1][i] = hi;
    f.lo[1][i] = lo;
  }
  return f;
}

// The fragments depend only on the lane: they are computed once per device
// (mfma_fragments_kernel, 64 lanes x 4 x 16 bytes) into the bytes after the
// twiddle tables, and every MFMA tile loads its lane's four 16-byte vectors.
constexpr size_t kMfmaFragBytes = 64 * 4 * 16;

__device__ __forceinline__ DftFrag load_fragments(const float2* tw, int lane) {
  const auto* p = (const GLSX_GLOBAL uint4*)(tw + kTwiddleEntries) + lane * 4;
  DftFrag f;
  f.hi[0] = __builtin_bit_cast(bf16x8_t, p[0]);
  f.hi[1] = __builtin_bit_cast(bf16x8_t, p[1]);
  f.lo[0] = __builtin_bit_cast(bf16x8_t, p[2]);
  f.lo[1] = __builtin_bit_cast(bf16x8_t, p[3]);
  return f;
}

// y = F a for the 32 butterflies of one wave.  a[i] is input m = 8h + i of
// this lane's butterfly (already twiddled); y[i] is output mfma_out_q(h, i).
__device__ __forceinline__ void dft16_mfma(const DftFrag& f, const float2 (&a)[8], float2 (&y)[8]) {
  uint4 rh, rl, ih, il;
  split_bf16x2(a[0].x, a[1].x, rh.x, rl.x);
  split_bf16x2(a[2].x, a[3].x, rh.y, rl.y);
  split_bf16x2(a[4].x, a[5].x, rh.z, rl.z);
  split_bf16x2(a[6].x, a[7].x, rh.w, rl.w);
  split_bf16x2(a[0].y, a[1].y, ih.x, il.x);
  split_bf16x2(a[2].y, a[3].y, ih.y, il.y);
  split_bf16x2(a[4].y, a[5].y, ih.z, il.z);
  split_bf16x2(a[6].y, a[7].y, ih.w, il.w);
  const bf16x8_t re_hi = __builtin_bit_cast(bf16x8_t, rh), re_lo = __builtin_bit_cast(bf16x8_t, rl);
  const bf16x8_t im_hi = __builtin_bit_cast(bf16x8_t, ih), im_lo = __builtin_bit_cast(bf16x8_t, il);
  f32x16_t d = {};
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.hi[0], re_hi, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.hi[1], im_hi, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.hi[0], re_lo, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.hi[1], im_lo, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.lo[0], re_hi, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.lo[1], im_hi, d, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 8; ++i) y[i] = make_float2(d[i], d[i + 8]);
}

__device__ __forceinline__ int mfma_out_q(int h, int i) { return i < 4 ? 4 * h + i : 4 + 4 * h + i; }

// The 16-point DFT of every lane's own butterfly on the matrix cores, in the
// FP32 kernels' register layout: a[m] is input m of this lane's butterfly
// before the call and output m after it.  dft16_mfma splits a butterfly over
// lanes n and n + 32 (lane n + 32h supplies inputs 8h..8h+7), so one
// v_permlane32_swap per dword first trades inputs 8..15 of lane n's butterfly
// for inputs 0..7 of lane n + 32's: a[0..7] then feed the butterflies of
// lanes 0..31 (group 0) and a[8..15] those of lanes 32..63 (group 1).  The
// same swap on the outputs returns the 16 outputs of each butterfly to its
// lane.  The products, and so the results, are those of dft16_mfma on the same
// inputs.  Every lane of the wave must execute it.
__device__ __forceinline__ void dft16_lanes(const DftFrag& f, float2 (&a)[16]) {
  const int lane = threadIdx.x & 63;
  swap_step<5, 3>(a, lane);
  float2 g0[8], g1[8], y0[8], y1[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    g0[i] = a[i];
    g1[i] = a[i + 8];
  }
  dft16_mfma(f, g0, y0);
  dft16_mfma(f, g1, y1);
  float2 o[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    o[i] = y0[i];
    o[i + 8] = y1[i];
  }
  // now o[i] is output mfma_out_q(0, i) and o[i + 8] output mfma_out_q(1, i)
  // of this lane's butterfly
  swap_step<5, 3>(o, lane);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[mfma_out_q(0, i)] = o[i];
    a[mfma_out_q(1, i)] = o[i + 8];
  }
}

// The butterfly policy of the split-precision passes (see DftFp32): radix-16
// butterflies on the matrix cores, the leading radix-2/4/8 stage in FP32.
struct DftMfma {
  DftFrag f;
  template <int R>
  __device__ __forceinline__ void run(float2 (&a)[R]) const {
    if constexpr (R == 16) dft16_lanes(f, a);
    else dft<R>(a);
  }
};

// The policy a kernel compiled with or without the matrix-core butterflies
// uses; the MFMA form loads this lane's DFT-matrix fragments.
template <bool MF>
__device__ __forceinline__ auto make_dft(const float2* tw) {
  if constexpr (MF) return DftMfma{load_fragments(tw, threadIdx.x & 63)};
  else return DftFp32{};
}

}  // namespace glsx
