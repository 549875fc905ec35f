// fft_wave.hpp -- stage exchanges inside one wave for row transforms of up to
// 64 threads (L <= 1024 at 16 elements per thread).
//
// Between two Stockham stages every element moves from (thread t, register i)
// to (thread t', register e), and both indices are bit fields of the element's
// position: the exchange is a permutation of the lane bits and register bits.
// It runs without LDS in three parts:
//   * every lane bit whose element must end in a register swaps with a
//     register bit whose element must end in a lane: one cross-lane step
//     (lane bit 4 or 5: v_permlane16_swap / v_permlane32_swap, one instruction
//     per dword; lane bits 0-3: DPP quad_perm or ds_swizzle plus selects);
//   * the lane permutation that remains is not carried out: the next stage
//     runs as thread t' = that permutation of the lane index (its twiddles and
//     output positions are computed from t');
//   * the register permutation is a renaming.
// The data movement is exact, so the results are bit-identical to the LDS
// exchange.  Measured on MI355X (row passes over 4M points, HIP events, two
// machines): 1024 points 15.0 -> 14.4 us on one and equal on the other, 512
// points 14.8 -> 14.0 and 14.4 -> 14.2 us, 32 and 64 points 1-3% faster, 128
// points equal, 256 points slower (12.4 -> 13.1 us: all four exchanged bits
// lie below lane bit 4 and need the select form).  The plain row pass uses it
// for 32, 64, 512 and 1024 points.  Placing a transform's threads at the top
// lane bits instead (lane = transform + W * thread, so that every swap is a
// permlane swap) measured 1.4x slower below 512 points: adjacent lanes then
// load from different rows and the loads stop coalescing.
#pragma once
#include "fft_core.hpp"

namespace glsx {

// Row lengths whose plain row pass exchanges inside the wave.
__host__ __device__ constexpr bool wave_rows(int L) { return L == 32 || L == 64 || L == 512 || L == 1024; }

// The exchange plan of every stage boundary, derived at compile time from the
// stage schedule.
template <int L, int EP = 16>
struct WavePlan {
  using Sc = Sched<L, EP>;
  static constexpr int E = Sc::E, T = Sc::T, NS = Sc::NSTAGE;
  static constexpr int LT = ilog2c(T), LE = ilog2c(E);
  struct Ex {
    int nswap = 0;
    int lane[8] = {};  // lane bit of swap k
    int reg[8] = {};   // the register-index bit it swaps with
    int tau[8] = {};   // next stage: lane bit p carries bit tau[p] of its thread index
    int ren[64] = {};  // next stage: input e is element ren[e] of this stage's output array
  };
  struct All {
    Ex ex[8];
    bool ok = true;
  };
  static constexpr All plan() {
    All a{};
    int tau[8] = {};
    for (int p = 0; p < LT; ++p) tau[p] = p;
    for (int s = 0; s + 1 < NS; ++s) {
      const int R = Sc::radix(s), LS = Sc::ls(s), NB = E / R;
      const int lr = ilog2c(R), lls = ilog2c(LS), lnb = ilog2c(NB);
      // output position of butterfly j, output q: the low log2(Ls) bits of j,
      // then the log2(R) bits of q, then the rest of j
      auto jbit_to_pos = [&](int x) { return x < lls ? x : x + lr; };
      // which position bit each lane bit and register bit holds after stage s
      // (output array index u + q*NB, butterfly j = thread + u*T)
      int lane_m[8] = {}, reg_m[8] = {};
      for (int p = 0; p < LT; ++p) lane_m[p] = jbit_to_pos(tau[p]);
      for (int b = 0; b < LE; ++b) reg_m[b] = b < lnb ? jbit_to_pos(LT + b) : lls + (b - lnb);
      // the next stage wants position bits below log2(T) in lanes, the rest in registers
      Ex& x = a.ex[s];
      int la[8] = {}, rb[8] = {}, nla = 0, nrb = 0;
      for (int p = 0; p < LT; ++p)
        if (lane_m[p] >= LT) la[nla++] = p;
      for (int b = 0; b < LE; ++b)
        if (reg_m[b] < LT) rb[nrb++] = b;
      if (nla != nrb) a.ok = false;
      x.nswap = nla;
      int lane_c[8] = {}, reg_c[8] = {};
      for (int p = 0; p < LT; ++p) lane_c[p] = lane_m[p];
      for (int b = 0; b < LE; ++b) reg_c[b] = reg_m[b];
      for (int k = 0; k < nla && k < nrb; ++k) {
        x.lane[k] = la[k];
        x.reg[k] = rb[k];
        lane_c[la[k]] = reg_m[rb[k]];
        reg_c[rb[k]] = lane_m[la[k]];
      }
      for (int p = 0; p < LT; ++p) {
        if (lane_c[p] >= LT) a.ok = false;
        x.tau[p] = lane_c[p];
        tau[p] = lane_c[p];
      }
      for (int e = 0; e < E; ++e) {
        int i = 0;
        for (int b = 0; b < LE; ++b) {
          if (reg_c[b] < LT) a.ok = false;
          i |= ((e >> (reg_c[b] - LT)) & 1) << b;
        }
        x.ren[e] = i;
      }
    }
    return a;
  }
  static constexpr All P = plan();
  static_assert(P.ok, "stage exchange is not a permutation of lane and register bits");
};

// The value held by lane ^ (1 << P), P < 4 (within a row of 16 lanes).
template <int P>
__device__ __forceinline__ float lane_xor(float x) {
  const int b = __float_as_int(x);
  int r;
  if constexpr (P == 0) r = __builtin_amdgcn_mov_dpp(b, 0xB1, 0xF, 0xF, false);       // quad_perm [1,0,3,2]
  else if constexpr (P == 1) r = __builtin_amdgcn_mov_dpp(b, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  else if constexpr (P == 2) r = __builtin_amdgcn_ds_swizzle(b, 0x1F | (4 << 10));    // bitmask mode, xor 4
  else r = __builtin_amdgcn_ds_swizzle(b, 0x1F | (8 << 10));                          // bitmask mode, xor 8
  return __int_as_float(r);
}

// Swap lane bit P with register-index bit B: the element at (lane bit x,
// register bit y) moves to (lane bit y, register bit x).  All lanes of the
// wave must execute it.
template <int P, int B, int E>
__device__ __forceinline__ void swap_step(float2 (&v)[E], int lane) {
  constexpr int M = 1 << B;
  if constexpr (P >= 4) {
    // v_permlane16_swap: rows 1 and 3 of the first operand trade places with
    // rows 0 and 2 of the second; v_permlane32_swap: the upper half of the
    // first with the lower half of the second
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if (i & M) continue;
      unsigned lo[2] = {__float_as_uint(v[i].x), __float_as_uint(v[i].y)};
      unsigned hi[2] = {__float_as_uint(v[i | M].x), __float_as_uint(v[i | M].y)};
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const auto r = P == 4 ? __builtin_amdgcn_permlane16_swap(lo[c], hi[c], false, false)
                              : __builtin_amdgcn_permlane32_swap(lo[c], hi[c], false, false);
        lo[c] = r[0];
        hi[c] = r[1];
      }
      v[i] = make_float2(__uint_as_float(lo[0]), __uint_as_float(lo[1]));
      v[i | M] = make_float2(__uint_as_float(hi[0]), __uint_as_float(hi[1]));
    }
  } else {
    const bool up = (lane >> P) & 1;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if (i & M) continue;
      // selects on values held in opaque registers: a select between two
      // array elements would become a select of addresses and move the whole
      // array to scratch memory
      float2 a0 = v[i], a1 = v[i | M];
      asm volatile("" : "+v"(a0.x), "+v"(a0.y), "+v"(a1.x), "+v"(a1.y));
      const float2 send = make_float2(up ? a0.x : a1.x, up ? a0.y : a1.y);
      const float2 recv = make_float2(lane_xor<P>(send.x), lane_xor<P>(send.y));
      v[i] = make_float2(up ? recv.x : a0.x, up ? recv.y : a0.y);
      v[i | M] = make_float2(up ? a1.x : recv.x, up ? a1.y : recv.y);
    }
  }
}

template <int L, int EP, int S, int... K>
__device__ __forceinline__ void wave_swaps(float2 (&v)[(Sched<L, EP>::E)], int t, std::integer_sequence<int, K...>) {
  using WP = WavePlan<L, EP>;
  (swap_step<WP::P.ex[S].lane[K], WP::P.ex[S].reg[K]>(v, t), ...);
}

template <int L, int EP, int S, int... I>
__device__ __forceinline__ void wave_rename(float2 (&v)[(Sched<L, EP>::E)], std::integer_sequence<int, I...>) {
  using WP = WavePlan<L, EP>;
  const float2 w[] = {v[std::integral_constant<int, WP::P.ex[S].ren[I]>::value]...};
  ((v[I] = w[I]), ...);
}

template <int L, int EP, int S, int... P>
__device__ __forceinline__ int wave_thread(int t, std::integer_sequence<int, P...>) {
  using WP = WavePlan<L, EP>;
  return (0 | ... | (((t >> P) & 1) << std::integral_constant<int, WP::P.ex[S].tau[P]>::value));
}

// Stage S and the exchange after it; tv is the thread index stage S runs as
// and is updated for stage S + 1.
template <int L, int EP, int S, class D>
__device__ __forceinline__ void wave_stage(float2 (&v)[(Sched<L, EP>::E)], int t, int& tv, const float2* tw, const D& d) {
  using WP = WavePlan<L, EP>;
  stage_butterflies<L, S, EP>(v, tv, tw, d);
  if constexpr (S + 1 < WP::NS) {
    wave_swaps<L, EP, S>(v, t, std::make_integer_sequence<int, WP::P.ex[S].nswap>{});
    wave_rename<L, EP, S>(v, std::make_integer_sequence<int, WP::E>{});
    tv = wave_thread<L, EP, S>(t, std::make_integer_sequence<int, WP::LT>{});
  }
}

template <int L, int EP, class D, int... S>
__device__ __forceinline__ int wave_stages_seq(float2 (&v)[(Sched<L, EP>::E)], int t, const float2* tw, const D& d,
                                               std::integer_sequence<int, S...>) {
  int tv = t;
  (wave_stage<L, EP, S>(v, t, tv, tw, d), ...);
  return tv;
}

// All stages of a transform whose T threads are consecutive lanes of one wave
// (row layout, T <= 64); t = the thread's index in its transform (its low lane
// bits).  Returns the thread index the last stage ran as: its outputs are
// stored as that thread's.
template <int L, int EP = 16, class D = DftFp32>
__device__ __forceinline__ int wave_stages(float2 (&v)[(Sched<L, EP>::E)], int t, const float2* tw, const D& d = D{}) {
  static_assert(Sched<L, EP>::T <= 64, "the transform must fit in one wave");
  return wave_stages_seq<L, EP>(v, t, tw, d, std::make_integer_sequence<int, Sched<L, EP>::NSTAGE>{});
}

}  // namespace glsx
