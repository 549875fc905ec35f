// mfma_fft.hpp -- split-precision FFT passes: the radix-16 stages run as
// 16x16 complex DFT-matrix products on the gfx950 matrix cores.
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
// residue below lo, about 2^-17 relative per product.  Twiddle multiplies
// (bases generated in FP64 on the device), a leading radix-2/4/8 stage, the
// load/store fusions and all data movement are the FP32 kernels' own
// (fft_pass.hpp).
//
// Operand layout of v_mfma_f32_32x32x16_bf16 (lane l, h = l / 32, n = l % 32):
//   A: lane holds A[n][8h + i], i = 0..7
//   B: lane holds B[8h + i][n]
//   D: lane holds D[8*(i/4) + 4h + i%4][n], i = 0..15
// With A = M and B = inputs (column n = butterfly n), lane (n, h) supplies
// inputs m = 8h..8h+7 of butterfly n and receives its outputs
// q = 4h..4h+3 and 8+4h..8+4h+3, Re in D[0..7] and Im in D[8..15].
#pragma once
#include "fft_pass.hpp"

namespace glsx {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

// round-to-nearest-even float -> bf16 bits (finite inputs), and back
__device__ __forceinline__ short bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }

__device__ __forceinline__ void split_bf16(float x, short& hi, short& lo) {
  hi = bf16_rne(x);
  lo = bf16_rne(x - bf16_to_f32(hi));
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

// y = F a for the 32 butterflies of one wave.  a[i] is input m = 8h + i of
// this lane's butterfly (already twiddled); y[i] is output mfma_out_q(h, i).
__device__ __forceinline__ void dft16_mfma(const DftFrag& f, const float2 (&a)[8], float2 (&y)[8]) {
  bf16x8_t re_hi, re_lo, im_hi, im_lo;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    short hi, lo;
    split_bf16(a[i].x, hi, lo);
    re_hi[i] = hi;
    re_lo[i] = lo;
    split_bf16(a[i].y, hi, lo);
    im_hi[i] = hi;
    im_lo[i] = lo;
  }
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

// One radix-16 stage S on the LDS-resident tile.  Group g of a wave covers the
// 32 butterflies the FP32 layout gives to threads wave_base + 32g + n.
template <int L, bool COL, int S>
__device__ __forceinline__ void mfma_stage(const PassArgs& a, long long tile, const TileCtx (&g2)[2],
                                           const DftFrag& frag, int h, float2* lds) {
  constexpr int R = 16, LS = Sched<L>::ls(S), LP = Sched<L>::LP;
  constexpr bool LAST = S + 1 == Sched<L>::NSTAGE;
  static_assert(Sched<L>::radix(S) == 16, "MFMA stages are radix 16");
  auto at = [&](int b, int pos) -> int {
    if constexpr (COL) return swz(pos) * a.w + b;
    else return b * LP + swz(pos);
  };
  lds_barrier();  // the previous stage's writes are visible
  float2 in[2][8];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int j = g2[g].t;
    const int k = j & (LS - 1);
    float2 w[16];
    if constexpr (LS > 1) {
      const auto* tab = gptr(a.tw) + tw_offset(LS * R);
      float2 bases[2] = {tab[k], tab[4 * k]};
      twiddles_from_bases<16>(bases, w);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int m = 8 * h + i;
      float2 x = g2[g].in_tile ? lds[at(g2[g].b, j + m * (L / 16))] : make_float2(0.0f, 0.0f);
      if constexpr (LS > 1) {
        // w[m] by selection (no dynamic register indexing); w[0] = 1 is never built
        if (i == 0) {
          if (h) x = cmul(x, w[8]);
        } else {
          x = cmul(x, h ? w[8 + i] : w[i]);
        }
      }
      in[g][i] = x;
    }
  }
  if constexpr (!LAST) lds_barrier();  // every read of this stage is done
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    float2 y[8];
    dft16_mfma(frag, in[g], y);
    const int j = g2[g].t;
    const int k = j & (LS - 1);
    if constexpr (!LAST) {
      if (g2[g].in_tile) {
#pragma unroll
        for (int i = 0; i < 8; ++i) lds[at(g2[g].b, (j - k) * 16 + k + mfma_out_q(h, i) * LS)] = y[i];
      }
    } else if (g2[g].active) {
      const int half = a.shift ? (L >> 1) : 0;
      auto* dst = gptr((float2*)a.dst) + g2[g].base;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int pos = (j + mfma_out_q(h, i) * (L / 16) + half) & (L - 1);
        float2 o = y[i];
        o.x *= a.scale;
        o.y *= a.scale;
        if (a.conj_out) o.y = -o.y;
        dst[(long long)pos * a.es] = o;
      }
    }
  }
}

// One tile of a pass whose radix-16 stages run on MFMA.  Same tile geometry
// as run_tile<L, COL> (16 elements per thread); the workgroup is a whole
// number of waves (threads past the tile feed zeros to the products).
// Output: interleaved complex (OUT_C32) only.
template <int L, bool COL>
__device__ __forceinline__ void run_tile_mfma(const PassArgs& a, long long tile, float2* lds, int tid) {
  using Sc = Sched<L>;
  constexpr int E = Sc::E, T = Sc::T, LP = Sc::LP;
  static_assert(L >= 16 && E == 16, "MFMA passes need at least one radix-16 stage");
  const TileCtx c = tile_ctx<L, COL>(a, tile, tid);
  auto at = [&](int b, int pos) -> int {
    if constexpr (COL) return swz(pos) * a.w + b;
    else return b * LP + swz(pos);
  };
  {
    float2 v[E];
    tile_load<L, COL>(a, c, v);
    if constexpr (Sc::R0 != 16) {
      // leading radix-2/4/8 stage in FP32 from registers, out to LDS
      constexpr int R = Sc::R0, NB = E / R;
      stage_butterflies<L, 0>(v, c.t, a.tw);
      if (c.in_tile) {
#pragma unroll
        for (int u = 0; u < NB; ++u)
#pragma unroll
          for (int q = 0; q < R; ++q) lds[at(c.b, stage_out_pos<L, 0>(c.t, u, q))] = v[u + q * NB];
      }
    } else {
      if (c.in_tile) {
#pragma unroll
        for (int e = 0; e < E; ++e) lds[at(c.b, c.t + e * T)] = v[e];
      }
    }
  }
  const int lane = tid & 63;
  const int h = lane >> 5, n = lane & 31;
  const DftFrag frag = dft16_fragments(lane);
  const int wave_base = tid & ~63;
  const TileCtx g2[2] = {tile_ctx<L, COL>(a, tile, wave_base + n), tile_ctx<L, COL>(a, tile, wave_base + 32 + n)};
  constexpr int FIRST = Sc::R0 != 16 ? 1 : 0;
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (mfma_stage<L, COL, S + FIRST>(a, tile, g2, frag, h, lds), ...);
  }(std::make_integer_sequence<int, Sc::NSTAGE - FIRST>{});
}

template <int L, bool COL>
__global__ void __launch_bounds__(1024) fft_mfma_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  long long tile = blockIdx.x;
  if constexpr (COL) {
    // same XCD-aware tile order as the FP32 column pass
    const unsigned G = gridDim.x;
    if ((G & 7u) == 0) tile = (long long)(blockIdx.x & 7u) * (G >> 3) + (blockIdx.x >> 3);
  }
  run_tile_mfma<L, COL>(a, tile, lds, threadIdx.x);
}

}  // namespace glsx
