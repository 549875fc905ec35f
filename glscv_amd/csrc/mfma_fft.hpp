// mfma_fft.hpp -- split-precision FFT passes whose radix-16 stages run as
// 16x16 complex DFT-matrix products on the gfx950 matrix cores
// (mfma_core.hpp: the split and the products).
//
// The plain row kernel (fft_row_kernel, whole rows of 2048 points and more)
// takes the matrix-core butterfly as a policy and keeps its own data layout
// (dft16_lanes).  This file holds the tiles of every other shape, whose data
// passes through LDS in the products' operand layout at every stage (the
// same policy in the wave-exchange rows and the pair column tiles measured
// slower than these tiles: 1024-point rows 6.0 against 5.8 us, 2048 x 2048
// columns 15.5 against 15.2 us).  Twiddle multiplies (bases generated in FP64 on the device), a
// leading radix-2/4/8 stage, the load/store fusions and all data movement are
// the FP32 kernels' own (fft_pass.hpp); both forms compute the same bits.
// The fragments of the DFT matrix are computed once per device into the bytes
// after the twiddle tables (mfma_fragments_kernel, 64 lanes x 4 x 16 bytes).
#pragma once
#include "fft_pass.hpp"

namespace glsx {

__global__ void __launch_bounds__(64) mfma_fragments_kernel(uint4* out) {
  const int lane = threadIdx.x;
  const DftFrag f = dft16_fragments(lane);
  out[lane * 4 + 0] = __builtin_bit_cast(uint4, f.hi[0]);
  out[lane * 4 + 1] = __builtin_bit_cast(uint4, f.hi[1]);
  out[lane * 4 + 2] = __builtin_bit_cast(uint4, f.lo[0]);
  out[lane * 4 + 3] = __builtin_bit_cast(uint4, f.lo[1]);
}

// One radix-16 stage S on the LDS-resident tile.  Group g of a wave covers the
// 32 butterflies the FP32 layout gives to threads wave_base + 32g + n.
template <int L, bool COL, int S>
__device__ __forceinline__ void mfma_stage(const PassArgs& a, long long tile, const TileCtx (&g2)[2],
                                           const DftFrag& frag, int h, float2* lds, const float2 (&tb)[2][2]) {
  constexpr int LS = Sched<L>::ls(S), LP = Sched<L>::LP;
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
