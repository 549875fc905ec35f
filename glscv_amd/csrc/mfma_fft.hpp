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

__global__ void __launch_bounds__(64) mfma_fragments_kernel(uint4* out) {
  const int lane = threadIdx.x;
  const DftFrag f = dft16_fragments(lane);
  out[lane * 4 + 0] = __builtin_bit_cast(uint4, f.hi[0]);
  out[lane * 4 + 1] = __builtin_bit_cast(uint4, f.hi[1]);
  out[lane * 4 + 2] = __builtin_bit_cast(uint4, f.lo[0]);
  out[lane * 4 + 3] = __builtin_bit_cast(uint4, f.lo[1]);
}

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

// One radix-16 stage S on the LDS-resident tile.  Group g of a wave covers the
// 32 butterflies the FP32 layout gives to threads wave_base + 32g + n.
template <int L, bool COL, int S>
__device__ __forceinline__ void mfma_stage(const PassArgs& a, long long tile, const TileCtx (&g2)[2],
                                           const DftFrag& frag, int h, float2* lds, const float2 (&tb)[2][2]) {
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
    const bool it = g2[g].in_tile;
    // threads past the tile read a valid slot and feed zeros (no branch per read)
    const int j = it ? g2[g].t : 0, bb = it ? g2[g].b : 0;
    // this lane's twiddles w^m, m = 8h + i: hi[i / 4] * p[i % 4] with
    // hi = (1, w^4) or (w^8, w^12) and p = (1, w, w^2, w^3) -- the products
    // twiddles_from_bases<16> forms for these m
    float2 hi[2], p[4];
    if constexpr (LS > 1) {
      const float2 w1 = tb[g][0], w4 = tb[g][1];  // tab[k], tab[4k], read at the tile's start
      const float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
      const float2 w8 = cmul(w4, w4), w12 = cmul(w8, w4);
      hi[0] = h ? w8 : make_float2(1.0f, 0.0f);
      hi[1] = h ? w12 : w4;
      p[1] = w1;
      p[2] = w2;
      p[3] = w3;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int m = 8 * h + i;
      float2 x = lds[at(bb, j + m * (L / 16))];
      if (!it) x = make_float2(0.0f, 0.0f);
      if constexpr (LS > 1) {
        if (i == 0) {
          const float2 c = cmul(x, hi[0]);  // w^0 = 1 is not applied
          x = h ? c : x;
        } else {
          x = cmul(x, (i & 3) == 0 ? hi[i >> 2] : cmul(hi[i >> 2], p[i & 3]));
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
        if (!COL && a.es == 1) dst[pos] = o;
        else dst[(long long)pos * a.es] = o;
      }
    }
  }
}

// The twiddle bases (w^k, w^4k of table LS*16) of MFMA stage S for both
// groups of this lane, read at the tile's start so that their latency
// overlaps the data loads instead of following each stage's barrier.
template <int L, int S>
__device__ __forceinline__ void mfma_stage_bases(const PassArgs& a, const TileCtx (&g2)[2], float2 (&tb)[2][2]) {
  constexpr int LS = Sched<L>::ls(S);
  if constexpr (LS > 1) {
    const auto* tab = gptr(a.tw) + tw_offset(LS * 16);
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int k = (g2[g].in_tile ? g2[g].t : 0) & (LS - 1);
      tb[g][0] = tab[k];
      tb[g][1] = tab[4 * k];
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
  const int lane = tid & 63;
  const int h = lane >> 5, n = lane & 31;
  const int wave_base = tid & ~63;
  const TileCtx g2[2] = {tile_ctx<L, COL>(a, tile, wave_base + n), tile_ctx<L, COL>(a, tile, wave_base + 32 + n)};
  constexpr int FIRST = Sc::R0 != 16 ? 1 : 0;
  constexpr int NM = Sc::NSTAGE - FIRST;  // MFMA stages
  constexpr int NE = NM < 2 ? NM : 2;     // whose twiddle bases are read before the data
  float2 tbs[NE][2][2];
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (mfma_stage_bases<L, S + FIRST>(a, g2, tbs[S]), ...);
  }(std::make_integer_sequence<int, NE>{});
  {
    float2 v[E];
    if (!COL && a.in_kind == IN_C32 && a.es == 1 && c.active) {
      // contiguous complex rows (every 2D FFT's row pass): 32-bit offsets
      // known at compile time, as fft_row_kernel loads them
      const auto* s = gptr((const float2*)a.src) + c.base + c.t;
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = s[e * T];
      if (a.conj_in) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
      }
    } else {
      tile_load<L, COL>(a, c, v);
    }
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
  const DftFrag frag = load_fragments(a.tw, lane);
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (
        [&] {
          if constexpr (S < NE) {
            mfma_stage<L, COL, S + FIRST>(a, tile, g2, frag, h, lds, tbs[S]);
          } else {
            float2 tb[2][2];
            mfma_stage_bases<L, S + FIRST>(a, g2, tb);
            mfma_stage<L, COL, S + FIRST>(a, tile, g2, frag, h, lds, tb);
          }
        }(),
        ...);
  }(std::make_integer_sequence<int, NM>{});
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

// Column pass of the split-precision variant over C-column tiles with the
// leading radix-2/4/8 stage fed by 16-byte loads: thread (pair p, index jt)
// loads rows j + m*L/R0 of two adjacent columns (one 16-byte access each;
// j = jt + v*L/8) and runs that stage's butterfly j on both columns in
// registers -- the butterflies and output positions of stage_butterflies<L, 0>
// -- into a C-column LDS tile; the MFMA stages then run as in
// run_tile_mfma<L, true> with C columns.  Same bits as fft_mfma_kernel<L, true>,
// half its load instructions.  C*L/16 threads.
template <int L, int C>
__global__ void __launch_bounds__(C * L / 16) fft_mfma_cols_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  using Sc = Sched<L>;
  constexpr int R0 = Sc::R0, NS = Sc::NSTAGE;
  static_assert(R0 < 16 && (8 / R0) * R0 == 8, "a leading radix-2/4/8 stage");
  static_assert(C == 4 || C == 8, "4 or 8 columns");
  constexpr int NP = C / 2;       // column pairs per tile
  constexpr int NJ = L / 8;       // butterfly indices per column pair handled at once
  constexpr int NBJ = 8 / R0;     // stage-0 butterflies per thread and column
  const long long tile = pair_tile_index<false>(blockIdx.x, gridDim.x);
  const int p = threadIdx.x % NP, jt = threadIdx.x / NP;
  const long long col0 = tile * C + 2 * p;
  PassArgs m = a;  // the MFMA stages address a C-column LDS tile
  m.w = C;
  m.log_w = C == 8 ? 3 : 2;
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, n = lane & 31;
  const int wave_base = threadIdx.x & ~63;
  const TileCtx g2[2] = {tile_ctx<L, true>(m, tile, wave_base + n), tile_ctx<L, true>(m, tile, wave_base + 32 + n)};
  float2 tbs[NS - 1][2][2];
  static_assert(NS - 1 <= 2, "two MFMA stages (1024 and 2048 points)");
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (mfma_stage_bases<L, S + 1>(m, g2, tbs[S]), ...);
  }(std::make_integer_sequence<int, NS - 1>{});
  const auto* s = (const GLSX_GLOBAL float4*)(gptr((const float2*)a.src) + col0);
#pragma unroll
  for (int vj = 0; vj < NBJ; ++vj) {
    const int j = jt + vj * NJ;
    float2 xa[R0], xb[R0];
#pragma unroll
    for (int m = 0; m < R0; ++m) {
      const float4 x = s[(long long)(j + m * (L / R0)) * a.es / 2];
      xa[m] = make_float2(x.x, a.conj_in ? -x.y : x.y);
      xb[m] = make_float2(x.z, a.conj_in ? -x.w : x.w);
    }
    dft<R0>(xa);
    dft<R0>(xb);
#pragma unroll
    for (int q = 0; q < R0; ++q) {
      lds[swz(j * R0 + q) * C + 2 * p] = xa[q];
      lds[swz(j * R0 + q) * C + 2 * p + 1] = xb[q];
    }
  }
  const DftFrag frag = load_fragments(a.tw, lane);
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (mfma_stage<L, true, S + 1>(m, tile, g2, frag, h, lds, tbs[S]), ...);
  }(std::make_integer_sequence<int, NS - 1>{});
}

}  // namespace glsx
