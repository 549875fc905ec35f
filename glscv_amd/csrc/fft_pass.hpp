// fft_pass.hpp -- one FFT pass over a batch of 1D transforms: the row pass and
// the column pass of a 2D FFT are both instances of this kernel.
//
//   row mode (COL=false): transform b = row, element i at base + b*bs + i
//     (i contiguous).  Threads of one transform are lane-contiguous, so every
//     global load/store of a wave covers consecutive addresses.
//   column mode (COL=true): transform b = column, element i at base + i*es + b.
//     W neighbouring columns share a workgroup and are lane-fastest, so each
//     row segment of W*8 bytes is read and written by consecutive lanes; the
//     full column (all L rows) stays in LDS for the exchange.
//
// Fused load variants (in_kind) and store variants (out_kind) fold the
// reference's separate merge / multiply / mulSpectrumsPhaseOnly / split
// passes into the FFT's own sweep:
//   IN_C32     interleaved complex float        (GlsMat CV_32FC2)
//   IN_R32     real float, imag = 0              (CV_32FC1 -> merge, glsFft.cpp:683-689)
//   IN_R32PAIR two real images as one complex one, z = a*w + i*b*w (w: the
//              optional window, gls::multiply of glsPhaseCorrelate.cpp:178-184)
//   IN_POCZ    row pass over the spectrum Z of such a pair: the two spectra
//              are split off Z, A(k) = (Z(k) + conj Z(-k))/2 and
//              B(k) = (Z(k) - conj Z(-k))/2i, and A*conj(B)/|A*conj(B)| is
//              transformed (glsBasicOperation.cpp:195-206); see run_tile_pocz
//   OUT_C32    interleaved complex float
//   OUT_R32    real part only                    (gls::split, glsPhaseCorrelate.cpp:203-204)
//   OUT_MAG    |z|                               (gls::magSpectrums, glsBasicOperation.cpp:337-343)
//   OUT_LOGMAG log(|z| + offset)                 (gls::logMagSpectrums, glsBasicOperation.cpp:345-351)
// plus per-workgroup argmax partials of the stored real values
// (gls::minMaxLoc, glsMinMaxLoc.cpp:65-118: first maximum in row-major order).
#pragma once
#include "fft_core.hpp"
#include "fft_wave.hpp"
#include "mfma_core.hpp"

namespace glsx {

enum InKind { IN_C32 = 0, IN_R32 = 1, IN_R32PAIR = 2, IN_POCZ = 3 };
enum OutKind { OUT_C32 = 0, OUT_R32 = 1, OUT_MAG = 2, OUT_LOGMAG = 3 };

// The display-chain value of a complex bin: |z|, or log(abs(|z| + offset))
// (the scalar shader's cases 9 and 13, glsBasicOperation.cpp:121-135).
__device__ __forceinline__ float spectrum_value(float2 z, bool log_mag, float offset) {
  const float rr = z.x * z.x, ii = z.y * z.y;
  const float m = sqrtf(rr + ii);
  return log_mag ? logf(fabsf(m + offset)) : m;
}

struct PassArgs {
  const void* src;      // IN_C32/IN_POCZ: float2*, IN_R32/IN_R32PAIR: float*
  const void* src2;     // IN_R32PAIR: second image (float*)
  const float* win;     // IN_R32PAIR: optional window, same layout as src
  void* dst;            // OUT_C32: float2*, OUT_R32/OUT_MAG/OUT_LOGMAG: float*
  const float2* tw;     // concatenated twiddle tables
  int2* argmax;         // optional per-workgroup partials: OUT_R32 (value bits, linear index),
                        // OUT_MAG/OUT_LOGMAG (max bits, min bits)
  long long n_batch;    // number of transforms
  long long es;         // element stride (in elements) of a transform
  long long es_out;     // pair column tiles: element stride of the output, 0 = es (the column
                        // transforms of a decimated 2D FFT read every other row of src and
                        // write dst's rows contiguously)
  long long bs;         // stride between transforms (in elements)
  long long out_cols;   // row pitch of the output, for the argmax linear index
  int in_kind, out_kind;
  int w, log_w;         // transforms per workgroup (power of two)
  int conj_in, conj_out, shift;
  float scale;
  int col_shift;        // argmax: this pass is the column pass (index = pos*out_cols + b)
  float mag_offset;     // OUT_LOGMAG: offset added to |z| before the log
};

// total order used for the argmax: larger value wins, ties go to the lower
// linear index; NaN is never selected.
__device__ __forceinline__ bool am_better(float v1, int i1, float v2, int i2) {
  return v1 > v2 || (v1 == v2 && i1 < i2);
}

template <int L, bool COL>
struct PassLaunch {
  static constexpr int T = Sched<L>::T;
  static constexpr int MAXT = COL ? 1024 : (T > 256 ? T : 256);
};

// Barrier for the LDS exchanges: orders LDS traffic only, so global loads
// issued earlier (the next tile's prefetch) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Where a thread sits in a tile: W transforms of length L (W = a.w), processed
// by the first W*T threads of the workgroup (the rest only join barriers).
struct TileCtx {
  int b, t;
  bool in_tile, active;
  long long bg, base;
};

template <int L, bool COL, int EP = 16>
__device__ __forceinline__ TileCtx tile_ctx(const PassArgs& a, long long tile, int tid) {
  constexpr int T = Sched<L, EP>::T;
  TileCtx c;
  c.in_tile = tid < a.w * T;
  if constexpr (COL) {
    c.b = tid & (a.w - 1);
    c.t = tid >> a.log_w;
  } else {
    c.t = tid & (T - 1);
    c.b = tid / T;
  }
  c.bg = tile * a.w + c.b;
  c.active = c.in_tile && c.bg < a.n_batch;
  c.base = c.bg * a.bs;
  return c;
}

// ---- load: thread t holds elements t + e*T ----------------------------------
template <int L, bool COL, int EP = 16>
__device__ __forceinline__ void tile_load(const PassArgs& a, const TileCtx& c, float2 (&v)[(Sched<L, EP>::E)]) {
  constexpr int E = Sched<L, EP>::E, T = Sched<L, EP>::T;
  const int t = c.t;
  const long long base = c.base;
  if (c.active) {
    if (a.in_kind == IN_C32) {
      // per-lane pointer plus a uniform stride: one 64-bit add per element
      const auto* s = gptr((const float2*)a.src) + base + (long long)t * a.es;
      const long long step = (long long)T * a.es;
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = s[e * step];
    } else if (a.in_kind == IN_R32) {
      const auto* s = gptr((const float*)a.src);
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = make_float2(s[base + (long long)(t + e * T) * a.es], 0.0f);
    } else {  // IN_R32PAIR (IN_POCZ tiles load in run_tile_pocz)
      const auto* sa = gptr((const float*)a.src);
      const auto* sb = gptr((const float*)a.src2);
      if (a.win) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const long long o = base + (long long)(t + e * T) * a.es;
          const float w = gptr(a.win)[o];
          v[e] = make_float2(sa[o] * w, sb[o] * w);
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const long long o = base + (long long)(t + e * T) * a.es;
          v[e] = make_float2(sa[o], sb[o]);
        }
      }
    }
    if (a.conj_in) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = make_float2(0.0f, 0.0f);
  }
}

// ---- stages: all butterflies, LDS exchanges between them ---------------------
// On return v[u + q*(E/R)] holds bin j + q*(L/R) of the last stage (radix R).
// The LDS is free again only after the caller's next lds_barrier().
template <int L, bool COL, int EP = 16, int HALVES = 1>
__device__ __forceinline__ void tile_stages(const PassArgs& a, const TileCtx& c, float2 (&v)[(Sched<L, EP>::E)],
                                            float2* lds, const EarlyBases<L, EP>& eb) {
  constexpr int E = Sched<L, EP>::E, T = Sched<L, EP>::T, NS = Sched<L, EP>::NSTAGE;
  constexpr int LP = Sched<L, EP>::LP;
  const int t = c.t, b = c.b;
  const bool in_tile = c.in_tile;
  // HALVES > 1 (column mode): the LDS holds w/HALVES columns, and the
  // exchange runs in HALVES rounds, each for one group of columns.
  const int wh = a.w / HALVES;
  const int bl = b & (wh - 1), bgrp = b / wh;
  auto lds_index = [&](int pos) -> int {
    if constexpr (COL) return swz(pos) * wh + bl;
    else return b * LP + swz(pos);
  };
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (
        [&] {
          if constexpr (bases_early<L, S, EP>()) stage_butterflies_b<L, S, EP>(v, eb.b[S]);
          else stage_butterflies<L, S, EP>(v, t, a.tw);
          if constexpr (S + 1 < NS) {
            constexpr int R = Sched<L, EP>::radix(S);
            constexpr int NB = E / R;
#pragma unroll
            for (int h = 0; h < HALVES; ++h) {
              if (S > 0 || h > 0) lds_barrier();  // earlier reads of the buffer are done
              const bool mine = in_tile && bgrp == h;
              if (mine) {
#pragma unroll
                for (int u = 0; u < NB; ++u)
#pragma unroll
                  for (int q = 0; q < R; ++q) lds[lds_index(stage_out_pos<L, S, EP>(t, u, q))] = v[u + q * NB];
              }
              lds_barrier();
              if (mine) {
#pragma unroll
                for (int e = 0; e < E; ++e) v[e] = lds[lds_index(t + e * T)];
              }
            }
          }
        }(),
        ...);
  }(std::make_integer_sequence<int, NS>{});
}

// ---- store: v[u + q*NB] is frequency bin j + q*(L/R) --------------------------
template <int L, bool COL, int EP = 16>
__device__ __forceinline__ void tile_store(const PassArgs& a, const TileCtx& c, float2 (&v)[(Sched<L, EP>::E)],
                                           long long am_slot, float* am_v, int* am_i, int tid) {
  constexpr int E = Sched<L, EP>::E, T = Sched<L, EP>::T, NS = Sched<L, EP>::NSTAGE;
  constexpr int RL = Sched<L, EP>::radix(NS - 1);
  constexpr int NBL = E / RL;
  const int t = c.t;
  const long long base = c.base, bg = c.bg;
  const int half = a.shift ? (L >> 1) : 0;
  const bool minmax = a.out_kind == OUT_MAG || a.out_kind == OUT_LOGMAG;
  float best = -__FLT_MAX__;
  int best_i = minmax ? __float_as_int(__FLT_MAX__) : 0x7fffffff;
  if (c.active && a.out_kind == OUT_C32) {
    // complex output (every FFT pass but the last of a phase correlation or a
    // spectrum view): one uniform test instead of one per element
    auto* d = gptr((float2*)a.dst) + base;
    const float sy = a.conj_out ? -a.scale : a.scale;
#pragma unroll
    for (int u = 0; u < NBL; ++u)
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int pos = ((t + u * T + q * (L / RL)) + half) & (L - 1);
        const float2 y = v[u + q * NBL];
        d[(long long)pos * a.es] = make_float2(y.x * a.scale, y.y * sy);
      }
  } else if (c.active) {
#pragma unroll
    for (int u = 0; u < NBL; ++u) {
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int pos = ((t + u * T + q * (L / RL)) + half) & (L - 1);
        float2 y = v[u + q * NBL];
        y.x *= a.scale;
        y.y *= a.scale;
        if (a.conj_out) y.y = -y.y;
        const long long o = base + (long long)pos * a.es;
        if (minmax) {
          const float m = spectrum_value(y, a.out_kind == OUT_LOGMAG, a.mag_offset);
          gptr((float*)a.dst)[o] = m;
          if (a.argmax) {  // min/max partials: best = max, best_i = bits of min
            best = fmaxf(best, m);
            best_i = __float_as_int(fminf(__int_as_float(best_i), m));
          }
        } else {  // OUT_R32
          gptr((float*)a.dst)[o] = y.x;
          if (a.argmax) {
            const int li = a.col_shift ? (int)((long long)pos * a.out_cols + bg) : (int)(bg * a.out_cols + pos);
            if (am_better(y.x, li, best, best_i)) { best = y.x; best_i = li; }
          }
        }
      }
    }
  }
  if (a.argmax) {
    // wave reduction, then across waves through LDS
    // (workgroups are powers of two; offsets >= blockDim would read idle lanes)
    // argmax: (value, first index); OUT_MAG/OUT_LOGMAG: (max, bits of min)
    auto combine = [&](float ov, int oi) {
      if (minmax) {
        best = fmaxf(best, ov);
        best_i = __float_as_int(fminf(__int_as_float(best_i), __int_as_float(oi)));
      } else if (am_better(ov, oi, best, best_i)) {
        best = ov;
        best_i = oi;
      }
    };
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      if (off >= (int)blockDim.x) continue;
      float ov = __shfl_xor(best, off);
      int oi = __shfl_xor(best_i, off);
      combine(ov, oi);
    }
    const int wave = tid >> 6, nw = (blockDim.x + 63) >> 6;
    if ((tid & 63) == 0) { am_v[wave] = best; am_i[wave] = best_i; }
    lds_barrier();
    if (tid == 0) {
      for (int i = 1; i < nw; ++i) combine(am_v[i], am_i[i]);
      gptr(a.argmax)[am_slot] = make_int2(__float_as_int(best), best_i);
    }
  }
}

// One tile of a pass, start to finish.  Used by the one-launch-per-pass
// kernel below and by the task-graph kernels.  `am_slot` is where the tile's
// argmax partial goes.
template <int L, bool COL>
__device__ __forceinline__ void run_tile(const PassArgs& a, long long tile, long long am_slot, float2* lds,
                                         float* am_v, int* am_i, int tid) {
  float2 v[Sched<L>::E];
  const TileCtx c = tile_ctx<L, COL>(a, tile, tid);
  EarlyBases<L> eb;
  eb.load(c.t, a.tw);
  tile_load<L, COL>(a, c, v);
  tile_stages<L, COL>(a, c, v, lds, eb);
  tile_store<L, COL>(a, c, v, am_slot, am_v, am_i, tid);
}

// A row tile whose transforms each fit in one wave, with the stage exchanges
// done across lanes instead of through LDS (fft_wave.hpp; same bits).
template <int L>
__device__ __forceinline__ void run_tile_wave(const PassArgs& a, long long tile, long long am_slot, float* am_v,
                                              int* am_i, int tid) {
  float2 v[Sched<L>::E];
  TileCtx c = tile_ctx<L, false>(a, tile, tid);
  tile_load<L, false>(a, c, v);
  c.t = wave_stages<L>(v, c.t, a.tw);
  tile_store<L, false>(a, c, v, am_slot, am_v, am_i, tid);
}

// Plain wave-exchange rows (wave_rows(L)): complex, real or real-pair input
// (IN_C32, IN_R32, IN_R32PAIR), complex output, contiguous rows, whole
// workgroups (n_batch a multiple of a.w), no partials -- the row pass of
// every 2D FFT and phase correlation with rows of these lengths.  The shared
// tile code's predicates and 64-bit strided addressing drop out; the rows
// are latency-bound (one wave per SIMD at 1024 x 1024), so the instructions
// they save are time (1024^2 row pass 6.0 -> 5.4 us).  Same bits as
// run_tile_wave.
// (Workgroup `wg` of the pass: rows wg*a.w .. wg*a.w + a.w - 1.)
template <int L>
__device__ __forceinline__ void row_wave_tile(const PassArgs& a, long long wg) {
  constexpr int E = Sched<L>::E, T = Sched<L>::T, NS = Sched<L>::NSTAGE;
  constexpr int RL = Sched<L>::radix(NS - 1), NBL = E / RL;
  static_assert(T <= 64, "one row per wave or less");
  const int t = threadIdx.x & (T - 1);
  const long long row = wg * a.w + (int)(threadIdx.x / T);
  const long long base = row * a.bs + t;
  float2 v[E];
  if (a.in_kind == IN_C32) {
    const auto* s = gptr((const float2*)a.src) + base;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = s[e * T];
  } else if (a.in_kind == IN_R32) {
    const auto* s = gptr((const float*)a.src) + base;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = make_float2(s[e * T], 0.0f);
  } else {  // IN_R32PAIR: two real images (times the optional window)
    const auto* sa = gptr((const float*)a.src) + base;
    const auto* sb = gptr((const float*)a.src2) + base;
    if (a.win) {
      const auto* sw = gptr(a.win) + base;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const float w = sw[e * T];
        v[e] = make_float2(sa[e * T] * w, sb[e * T] * w);
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = make_float2(sa[e * T], sb[e * T]);
    }
  }
  if (a.conj_in) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
  }
  const int tq = wave_stages<L>(v, t, a.tw);
  auto* d = gptr((float2*)a.dst) + row * a.bs;
  const int half = a.shift ? (L >> 1) : 0;
  const float sy = a.conj_out ? -a.scale : a.scale;
#pragma unroll
  for (int u = 0; u < NBL; ++u)
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int pos = ((tq + u * T + q * (L / RL)) + half) & (L - 1);
      const float2 y = v[u + q * NBL];
      d[pos] = make_float2(y.x * a.scale, y.y * sy);
    }
}

template <int L>
__global__ void __launch_bounds__(64) fft_row_wave_kernel(PassArgs a) {
  row_wave_tile<L>(a, blockIdx.x);
}

// A whole 2D FFT of a small frame (R x C, both 16..128) in one workgroup of
// R*C/16 threads: the rows in the wave (16 elements per thread), through LDS
// (row pitch C + 1) to the columns, the columns in the wave.  The same stages,
// twiddles, flags and roundings as the row pass (conj_in, scale_r, shift)
// followed by the column pass (scale_c, conj_out, shift), so the same bits,
// in one launch instead of two: a call on a frame this small is bound by the
// host's launch cost (two launches cost about 7 us per call on the host).
// In place is allowed: every load precedes the barrier, every store follows it.
struct SmallArgs {
  const void* src;  // IN_C32: float2*, IN_R32: float*
  void* dst;        // float2*
  const float2* tw;
  int in_kind;
  int conj;         // INVERSE: conjugate the input of the rows and the output of the columns
  int shift;
  float scale_r, scale_c;
};

template <int R, int C>
__global__ void __launch_bounds__(R * C / 16) fft2d_small_kernel(SmallArgs a) {
  extern __shared__ float2 lds[];
  constexpr int TR = C / 16, TC = R / 16, P = C + 1;
  static_assert(Sched<C>::E == 16 && Sched<R>::E == 16, "16 elements per thread in both dimensions");
  const int tid = threadIdx.x;
  {
    const int b = tid / TR, t = tid % TR;  // row b, thread t of the row
    float2 v[16];
    if (a.in_kind == IN_C32) {
      const auto* s = gptr((const float2*)a.src) + b * C + t;
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = s[e * TR];
    } else {
      const auto* s = gptr((const float*)a.src) + b * C + t;
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = make_float2(s[e * TR], 0.0f);
    }
    if (a.conj) {
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e].y = -v[e].y;
    }
    const int tq = wave_stages<C>(v, t, a.tw);
    constexpr int RL = Sched<C>::radix(Sched<C>::NSTAGE - 1), NBL = 16 / RL;
    const int half = a.shift ? (C >> 1) : 0;
#pragma unroll
    for (int u = 0; u < NBL; ++u)
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int pos = ((tq + u * TR + q * (C / RL)) + half) & (C - 1);
        const float2 y = v[u + q * NBL];
        lds[b * P + pos] = make_float2(y.x * a.scale_r, y.y * a.scale_r);
      }
  }
  __syncthreads();
  {
    const int c = tid / TC, t = tid % TC;  // column c, thread t of the column
    float2 v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = lds[(t + e * TC) * P + c];
    const int tq = wave_stages<R>(v, t, a.tw);
    constexpr int RL = Sched<R>::radix(Sched<R>::NSTAGE - 1), NBL = 16 / RL;
    const int half = a.shift ? (R >> 1) : 0;
    const float sy = a.conj ? -a.scale_c : a.scale_c;
    auto* d = gptr((float2*)a.dst) + c;
#pragma unroll
    for (int u = 0; u < NBL; ++u)
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int pos = ((tq + u * TC + q * (R / RL)) + half) & (R - 1);
        const float2 y = v[u + q * NBL];
        d[pos * C] = make_float2(y.x * a.scale_c, y.y * sy);
      }
  }
}

// Phase-only cross power A*conj(B)/|A*conj(B)|, 0 where it vanishes
// (glsBasicOperation.cpp:195-206 divides without a guard).  Computed as
// c * rsqrt(|c|^2): one reciprocal square root and two multiplies per bin
// instead of a square root and two IEEE divisions (a third of this pass's
// vector instructions).  The guard |c| > 1e-30 of c / |c| is |c|^2 > 0 in
// FP32 (a nonzero |c|^2 is at least 1.4e-45, so |c| >= 3.7e-23); NaN still
// gives 0, and the result differs from c / |c| by about one ulp.  A
// denormal |c|^2 is scaled by 2^64 before the hardware reciprocal square
// root (which need not handle denormal inputs) and the result by 2^32.
__device__ __forceinline__ float2 poc_phase(float2 x, float2 y) {
  const float re = x.x * y.x + x.y * y.y;
  const float im = -x.x * y.y + x.y * y.x;
  const float s = re * re + im * im;
  if (!(s > 0.0f)) return make_float2(0.0f, 0.0f);
  const bool tiny = s < 1.17549435e-38f;  // below the smallest normal float
  const float inv = rsqrtf(tiny ? s * 0x1p64f : s) * (tiny ? 0x1p32f : 1.0f);
  return make_float2(re * inv, im * inv);
}

// Row tile of the IN_POCZ pass over an M x L spectrum Z = FFT2(a + i b) of two
// real images.  Tile k (0 <= k <= M/2) holds rows k and (M - k) mod M, the
// rows whose bins are each other's mirrors: bin (r, c) pairs with
// (-r mod M, -c mod L).  Each thread reads its bins and their mirrors (the
// other row of the tile, read by the other half of the workgroup at the same
// time), the two spectra separate as
//   2A = Z + conj(Zm),   2B = -i (Z - conj(Zm)),
// and the phase-only cross power (the factors 2 cancel in the normalisation)
// goes through the row FFT like any loaded row.  Rows 0 and M/2 are their own
// mirrors: their tiles hold the row twice and store it once.  The tile needs
// pocz_threads(L) threads (two rows of 16-element threads); a.w must be 2,
// a.n_batch = M.
template <int L>
__device__ __forceinline__ void run_tile_pocz(const PassArgs& a, long long tile, float2* lds, float* am_v,
                                              int* am_i, int tid) {
  using Sc = Sched<L>;
  constexpr int E = Sc::E, T = Sc::T;
  const long long M = a.n_batch;
  const long long r0 = tile, r1 = (M - tile) & (M - 1);
  TileCtx c;
  c.t = tid & (T - 1);
  c.b = tid / T;
  EarlyBases<L> eb;
  eb.load(c.t, a.tw);
  c.in_tile = tid < 2 * T;
  c.bg = c.b == 0 ? r0 : r1;
  c.base = c.bg * a.bs;
  c.active = c.in_tile && (c.b == 0 || r1 != r0);
  float2 v[E];
  if (c.in_tile) {
    // each bin and its mirror (the other row of the pair, column -c), both
    // from global memory: the workgroup's other half loads the mirror row at
    // the same time, so those reads hit the cache
    const auto* s = gptr((const float2*)a.src) + c.base;
    const auto* sm = gptr((const float2*)a.src) + (c.b == 0 ? r1 : r0) * a.bs;
    float2 zmv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = s[(long long)(c.t + e * T) * a.es];
#pragma unroll
    for (int e = 0; e < E; ++e) zmv[e] = sm[(long long)((L - (c.t + e * T)) & (L - 1)) * a.es];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float2 z = v[e];
      const float2 zm = zmv[e];
      const float2 A = make_float2(z.x + zm.x, z.y - zm.y);
      const float2 B = make_float2(z.y + zm.y, zm.x - z.x);
      v[e] = poc_phase(A, B);
      if (a.conj_in) v[e].y = -v[e].y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = make_float2(0.0f, 0.0f);
  }
  // (in place: a row is stored only after the stages' LDS barriers, which every
  // thread reaches after using its mirror values; lengths below 32 have one
  // stage and one wave, whose loads complete before its stores)
  tile_stages<L, false>(a, c, v, lds, eb);
  tile_store<L, false>(a, c, v, tile, am_v, am_i, tid);
}

// LDS bytes and threads of an IN_POCZ tile (two rows).
__host__ __device__ constexpr size_t pocz_lds_bytes(int L) {
  return 2 * (size_t)(L + L / 16 + (L >= 16 ? 1 : 0)) * sizeof(float2);
}
__host__ __device__ constexpr int pocz_threads(int L) { return L < 16 ? 2 : L / 8; }

// Column tiles that share 128-byte lines (W < 16) must run on the same XCD
// so that its L2 merges their partial-line reads and writes: workgroups are
// dispatched round-robin over the 8 XCDs, so tile = (wg % 8) * (G/8) + wg / 8
// gives every XCD a contiguous run of tiles (G workgroups).
__host__ __device__ __forceinline__ long long xcd_tile_index(long long wg, long long G) {
  return (G & 7) == 0 ? (wg & 7) * (G >> 3) + (wg >> 3) : wg;
}

// One tile of the one-launch-per-pass kernel; `slot` is the tile's argmax
// partial.
template <int L, bool COL>
__device__ __forceinline__ void pass_tile(const PassArgs& a, long long tile, long long slot, float2* lds, float* am_v,
                                          int* am_i) {
  if constexpr (!COL && wave_rows(L)) run_tile_wave<L>(a, tile, slot, am_v, am_i, threadIdx.x);
  else run_tile<L, COL>(a, tile, slot, lds, am_v, am_i, threadIdx.x);
}

template <int L, bool COL>
__global__ void __launch_bounds__((PassLaunch<L, COL>::MAXT)) fft_pass_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  const long long tile = COL ? xcd_tile_index(blockIdx.x, gridDim.x) : (long long)blockIdx.x;
  pass_tile<L, COL>(a, tile, blockIdx.x, lds, am_v, am_i);
}

// The stages and the store of a whole row held as 16 elements per thread
// (thread t: elements t + e*T), output row at dst + out_base.  The LDS is free
// again only after the caller's next lds_barrier().
template <int L, class D>
__device__ __forceinline__ void row_stages_store(const PassArgs& a, float2 (&v)[(Sched<L>::E)], int t,
                                                 long long out_base, float2* lds, const EarlyBases<L>& eb,
                                                 const D& dft) {
  constexpr int E = Sched<L>::E, T = Sched<L>::T, NS = Sched<L>::NSTAGE;
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (
        [&] {
          if constexpr (bases_early<L, S>()) stage_butterflies_b<L, S>(v, eb.b[S], dft);
          else stage_butterflies<L, S>(v, t, a.tw, dft);
          if constexpr (S + 1 < NS) {
            constexpr int R = Sched<L>::radix(S), NB = E / R;
            if constexpr (S > 0) lds_barrier();  // earlier reads of the buffer are done
#pragma unroll
            for (int u = 0; u < NB; ++u)
#pragma unroll
              for (int q = 0; q < R; ++q) lds[swz(stage_out_pos<L, S>(t, u, q))] = v[u + q * NB];
            lds_barrier();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[swz(t + e * T)];
          }
        }(),
        ...);
  }(std::make_integer_sequence<int, NS>{});
  constexpr int RL = Sched<L>::radix(NS - 1), NBL = E / RL;
  const int half = a.shift ? (L >> 1) : 0;
  const float sy = a.conj_out ? -a.scale : a.scale;
  auto* d = gptr((float2*)a.dst) + out_base;
#pragma unroll
  for (int u = 0; u < NBL; ++u)
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int pos = ((t + u * T + q * (L / RL)) + half) & (L - 1);
      const float2 y = v[u + q * NBL];
      d[pos] = make_float2(y.x * a.scale, y.y * sy);
    }
}


// Row pass of whole rows, one row per workgroup: the plain row pass of 2048
// to 16384 points with complex output and no partials, from complex, real
// or real-pair input (IN_C32, IN_R32, IN_R32PAIR).
// The shared tile code above carries predicates for partial tiles, tests of
// the input and output kinds and the argmax tail, which the task-graph
// kernels and the other passes need; without them the 4096-point row pass of
// a 4096 x 4096 transform takes 40.1-40.3 us instead of 41.8-42.9 us (its
// loads and stores alone: 38.7 us).  Rows are contiguous (element stride 1,
// as in every row pass here), addressed with compile-time offsets.  Same
// bits as run_tile.  MF: radix-16 butterflies on the matrix cores (split
// precision), the same bits as the split-precision tile of mfma_fft.hpp.
template <int L, bool MF = false>
__device__ __forceinline__ void row_tile(const PassArgs& a, long long row, float2* lds) {
  constexpr int E = Sched<L>::E, T = Sched<L>::T;
  static_assert(T >= 64 && E == 16, "one row of 16-element threads per workgroup");
  const int t = threadIdx.x;
  const long long base = row * a.bs;
  EarlyBases<L> eb;
  eb.load(t, a.tw);
  const auto dft = make_dft<MF>(a.tw);
  float2 v[E];
  if (a.in_kind == IN_C32) {
    const auto* s = gptr((const float2*)a.src) + base + t;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float2 x = s[e * T];
      v[e] = make_float2(x.x, a.conj_in ? -x.y : x.y);
    }
  } else {
    // real input (imag 0) or a pair of real images (one as the imaginary
    // part, both times the optional window): tile_load's values and order
    if (a.in_kind == IN_R32) {
      const auto* s = gptr((const float*)a.src) + base + t;
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = make_float2(s[e * T], 0.0f);
    } else {
      const auto* sa = gptr((const float*)a.src) + base + t;
      const auto* sb = gptr((const float*)a.src2) + base + t;
      if (a.win) {
        const auto* sw = gptr(a.win) + base + t;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const float w = sw[e * T];
          v[e] = make_float2(sa[e * T] * w, sb[e * T] * w);
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = make_float2(sa[e * T], sb[e * T]);
      }
    }
    if (a.conj_in) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
    }
  }
  row_stages_store<L>(a, v, t, base, lds, eb, dft);
}

template <int L, bool MF = false>
__global__ void __launch_bounds__(Sched<L>::T) fft_row_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  row_tile<L, MF>(a, blockIdx.x, lds);
}

// Row pass of a 2D FFT whose column transforms were decimated in time by R
// (2 or 4; fft2d_dit in fft_kernels.hip): rows q*M/R + j of src (q < R, j <
// M/R) hold, at every column, the M/R-point transform Y_q of the image rows
// R*i + q at column frequency j.  The last stage of the column transform,
//   X[k] = sum_q w_M^(q*k) Y_q[k mod M/R],
// runs on the loaded rows (the twiddles are uniform over the workgroup), then
// row k is transformed and stored; SHIFT moves it to row k + M/2.  The input
// conjugation of an inverse transform belongs to the column pass, which ran
// first.
//
// One workgroup per output row, each reading its R input rows; src and dst
// must not overlap.  Workgroups go round-robin over the 8 XCDs, so the R
// workgroups w + 8s (s < R) within each run of 8R, which share k mod M/R and
// so their input rows, run on one XCD at about the same time and the later
// reads of a row find it in that XCD's L2.  M: a multiple of 8R.
template <int R>
__host__ __device__ constexpr long long dit_row(long long w, long long M) {
  return (w / (8 * R)) * 8 + (w & 7) + ((w >> 3) & (R - 1)) * (M / R);
}

template <int L, int R>
__global__ void __launch_bounds__(Sched<L>::T) fft_row_dit_kernel(PassArgs a, int M) {
  extern __shared__ float2 lds[];
  constexpr int E = Sched<L>::E, T = Sched<L>::T;
  static_assert(T >= 64 && E == 16, "one row of 16-element threads per workgroup");
  static_assert(R == 2 || R == 4, "radix of the decimation");
  const int t = threadIdx.x;
  const long long k = dit_row<R>(blockIdx.x, M);
  const long long k0 = k & (M / R - 1);
  const long long out = ((k + (a.shift ? M / 2 : 0)) & (M - 1)) * a.bs;
  EarlyBases<L> eb;
  eb.load(t, a.tw);
  float2 v[E];
  {
    const auto* s = gptr((const float2*)a.src) + t;
    float2 y[R][E];
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int e = 0; e < E; ++e) y[q][e] = s[((long long)q * (M / R) + k0) * a.bs + e * T];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = y[0][e];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      const float2 wq = gptr(a.tw)[tw_offset(M) + ((q * k) & (M - 1))];
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = cadd(v[e], cmul(y[q][e], wq));
    }
  }
  row_stages_store<L>(a, v, t, out, lds, eb, DftFp32{});
}

// Column pass for long columns at two workgroups per CU.  A tile of W
// columns holds 32 elements per thread (half the threads of the 16-element
// layout) and exchanges through LDS one half of its columns at a time, so the
// tile needs half the LDS; two tiles share a CU and one's exchanges overlap
// the other's HBM traffic, while the row segments stay W*8 bytes wide.
template <int L>
__device__ __forceinline__ void col_half_tile(const PassArgs& a, long long tile, long long slot, float2* lds,
                                              float* am_v, int* am_i) {
  float2 v[Sched<L, 32>::E];
  const TileCtx c = tile_ctx<L, true, 32>(a, tile, threadIdx.x);
  EarlyBases<L, 32> eb;  // (no stage has one butterfly per thread at 32 elements)
  tile_load<L, true, 32>(a, c, v);
  tile_stages<L, true, 32, 2>(a, c, v, lds, eb);
  tile_store<L, true, 32>(a, c, v, slot, am_v, am_i, threadIdx.x);
}

template <int L>
__global__ void __launch_bounds__(512) fft_col_half_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  col_half_tile<L>(a, xcd_tile_index(blockIdx.x, gridDim.x), blockIdx.x, lds, am_v, am_i);
}

// Tile order of a pair-kernel launch.  Workgroups are dispatched round-robin
// over the 8 XCDs (workgroup w runs on XCD w % 8).  The two 8-column tiles 2k
// and 2k + 1 share every 128-byte line, so both run on one XCD, back to back
// (workgroups w and w + 8), and its L2 merges their half-line traffic.  SPREAD
// gives XCD x the line pairs x, x + 8, x + 16, ... (the workgroups in flight
// cover each row from its start); otherwise XCD x takes a contiguous range of
// tiles (G / 8 of them).  Measured at 4096^2 with the in-wave second exchange:
// 49.9 us spread against 51.2 us contiguous on one machine (the loads and
// stores alone 42.1 against 45.2 us), equal on another; at 1024 points spread
// is 3% slower, so only the 4096-point tiles use it.
template <bool SPREAD>
__host__ __device__ __forceinline__ long long pair_tile_index(long long w, long long G) {
  if constexpr (SPREAD) {
    if ((G & 15) == 0) return 2 * ((w >> 4) * 8 + (w & 7)) + ((w >> 3) & 1);
  }
  return xcd_tile_index(w, G);
}

// Column pass with two adjacent columns per lane: a tile of 2P columns is P
// column pairs of T threads, every thread holding 16 elements of each column
// of its pair, loaded and stored as one 16-byte access per row (half the
// memory instructions of one column per lane, each covering twice the rows;
// the two columns share their twiddles).  The LDS exchanges run in two
// rounds, one column of every pair at a time.  Same bits as the other column
// kernels.  Measured at P = 4 (8-column tiles, 64-byte row segments) against
// the one-column-per-lane tiles, same process, in place: 4096 points 52.3-52.6
// us against 64.9 us (58.9 us at P = 2), 2048 points 12.9 us against 15.2 us,
// 1024 points 7.5 us against 8.5 us.
//
// With three stages (1024, 2048 and 4096 points: a radix-4/8/16 stage, then two
// radix-16 stages) the second exchange runs inside the wave.  Thread t of a
// pair sits at lane bits 2-5 = t & 15 and wave t >> 4.  With S1 = L/256, the
// span of stage 1, the LDS exchange after stage 0 hands it butterfly
// tv = (t >> 4) + S1 (t & 15) instead of t; stage 1's output q of that
// butterfly is element T (t & 15) + (t >> 4) + S1 q (T = L/16), so swapping
// lane bits 2..5 with register bits 0..3 (permlane swaps and swizzles, no LDS,
// no barrier) leaves element tv + T e in register e: stage 2's inputs, again
// as butterfly tv.  Its outputs are rows tv + T q.  Measured at 4096 points
// 52.6-53.6 -> 51.0-51.4 us on two machines, same bits; reading the twiddle
// bases of stages 1 and 2 before the data (EarlyBases), 50.0-50.2 -> 49.5 us.
// At 1024 and 2048 points (4 and 8 waves per tile) the same exchange measured
// slower than the LDS one, 8.3 against 7.9 us and 14.3 against 13.0 us, so
// only the 4096-point tiles use it.
//
// Complex input; complex output (MAG = false: conj, scale, shift), or with
// MAG the |z| / log(|z| + offset) plane with optional per-workgroup (max, min)
// partials (OUT_MAG / OUT_LOGMAG), or with REAL the real plane with optional
// per-workgroup argmax partials (OUT_R32: the phase-correlation surface, 1024
// and 2048 points).  The
// host checks n_batch % (2P) == 0, unit
// batch stride, an even element stride and aligned buffers (pair_tile()).
// Whether the pair kernel of L points and P pairs spreads its tiles (see
// pair_tile_index).
__host__ __device__ constexpr bool pair_spread(int L, int P) { return L == 4096 && P == 4; }

template <int L, int P, bool MAG = false, bool REAL = false>
__device__ __forceinline__ void col_pair_tile(const PassArgs& a, long long tile, long long slot, float2* lds,
                                              float* am_v, int* am_i) {
  constexpr int E = Sched<L>::E, T = Sched<L>::T, NS = Sched<L>::NSTAGE;
  // the in-wave second exchange: 4096 points, 4 pairs per tile (a pair's 256
  // threads span lane bits 2-5 and 16 waves)
  constexpr bool WX = P == 4 && NS == 3 && E == 16 && L == 4096;
  const int b = threadIdx.x % P, t = threadIdx.x / P;  // pair b of the tile, thread t of its columns
  const long long col0 = tile * (2 * P) + 2 * b;
  // the butterfly index of stages 1.. (WX: tv, see above; otherwise t)
  constexpr int LS1 = L / 256;  // the span of stage 1 (its leading stage's radix)
  const int tv = WX ? ((t >> 4) | ((t & 15) * LS1)) : t;
  // twiddle bases of the later stages, read before the data; up to three
  // stages (at 8192 points, 1024 threads, a third set would spill)
  constexpr bool EARLY = NS <= 3;
  EarlyBases<L> eb;
  if constexpr (EARLY) eb.load(tv, a.tw);
  float2 va[E], vb[E];
  {
    const auto* s = (const GLSX_GLOBAL float4*)(gptr((const float2*)a.src) + col0 + (long long)t * a.es);
    const long long step = (long long)T * a.es / 2;  // in float4
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float4 x = s[e * step];
      va[e] = make_float2(x.x, a.conj_in ? -x.y : x.y);
      vb[e] = make_float2(x.z, a.conj_in ? -x.w : x.w);
    }
  }
  auto lidx = [&](int pos) { return swz(pos) * P + b; };
  // the butterfly index the last stage runs as (its outputs are rows tq + u*T + q*L/R)
  int tq = t;
  if constexpr (WX) {
    static_assert(T == 16 * LS1, "stages 1 and 2 are radix 16");
    const int lane = threadIdx.x & 63;
    stage_butterflies<L, 0>(va, t, a.tw);
    stage_butterflies<L, 0>(vb, t, a.tw);
    constexpr int R0 = Sched<L>::radix(0), NB0 = E / R0;
    auto exchange = [&](float2(&v)[E], bool first) {
      if (!first) lds_barrier();  // earlier reads of the buffer are done
#pragma unroll
      for (int u = 0; u < NB0; ++u)
#pragma unroll
        for (int q = 0; q < R0; ++q) lds[lidx(stage_out_pos<L, 0>(t, u, q))] = v[u + q * NB0];
      lds_barrier();
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = lds[lidx(tv + e * T)];
    };
    exchange(va, true);
    exchange(vb, false);
    stage_butterflies_b<L, 1>(va, eb.b[1]);
    stage_butterflies_b<L, 1>(vb, eb.b[1]);
    swap_step<2, 0>(va, lane);
    swap_step<3, 1>(va, lane);
    swap_step<4, 2>(va, lane);
    swap_step<5, 3>(va, lane);
    swap_step<2, 0>(vb, lane);
    swap_step<3, 1>(vb, lane);
    swap_step<4, 2>(vb, lane);
    swap_step<5, 3>(vb, lane);
    stage_butterflies_b<L, 2>(va, eb.b[2]);
    stage_butterflies_b<L, 2>(vb, eb.b[2]);
    tq = tv;
  } else {
    [&]<int... S>(std::integer_sequence<int, S...>) {
      (
          [&] {
            if constexpr (EARLY && bases_early<L, S>()) {
              stage_butterflies_b<L, S>(va, eb.b[S]);
              stage_butterflies_b<L, S>(vb, eb.b[S]);
            } else {
              stage_butterflies<L, S>(va, t, a.tw);
              stage_butterflies<L, S>(vb, t, a.tw);
            }
            if constexpr (S + 1 < NS) {
              constexpr int R = Sched<L>::radix(S), NB = E / R;
              auto exchange = [&](float2(&v)[E], bool first) {
                if (!first) lds_barrier();  // earlier reads of the buffer are done
#pragma unroll
                for (int u = 0; u < NB; ++u)
#pragma unroll
                  for (int q = 0; q < R; ++q) lds[lidx(stage_out_pos<L, S>(t, u, q))] = v[u + q * NB];
                lds_barrier();
#pragma unroll
                for (int e = 0; e < E; ++e) v[e] = lds[lidx(t + e * T)];
              };
              exchange(va, S == 0);
              exchange(vb, false);
            }
          }(),
          ...);
    }(std::make_integer_sequence<int, NS>{});
  }
  constexpr int RL = Sched<L>::radix(NS - 1), NBL = E / RL;
  const int half = a.shift ? (L >> 1) : 0;
  const float sy = a.conj_out ? -a.scale : a.scale;
  const long long es2 = (a.es_out ? a.es_out : a.es) / 2;
  if constexpr (REAL) {
    // the real parts of the two columns of a row as one 8-byte store; the
    // tile's first maximum in row-major order (linear index pos*out_cols +
    // column), as tile_store keeps it for OUT_R32
    auto* d = (GLSX_GLOBAL float2*)(gptr((float*)a.dst) + col0);
    float best = -__FLT_MAX__;
    int best_i = 0x7fffffff;
#pragma unroll
    for (int u = 0; u < NBL; ++u)
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int pos = ((tq + u * T + q * (L / RL)) + half) & (L - 1);
        const float r0 = va[u + q * NBL].x * a.scale, r1 = vb[u + q * NBL].x * a.scale;
        d[(long long)pos * es2] = make_float2(r0, r1);
        const int li = (int)((long long)pos * a.out_cols + col0);
        if (am_better(r0, li, best, best_i)) { best = r0; best_i = li; }
        if (am_better(r1, li + 1, best, best_i)) { best = r1; best_i = li + 1; }
      }
    if (a.argmax) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(best, off);
        const int oi = __shfl_xor(best_i, off);
        if (am_better(ov, oi, best, best_i)) { best = ov; best_i = oi; }
      }
      const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
      if ((threadIdx.x & 63) == 0) {
        am_v[wave] = best;
        am_i[wave] = best_i;
      }
      lds_barrier();
      if (threadIdx.x == 0) {
        for (int i = 1; i < nw; ++i)
          if (am_better(am_v[i], am_i[i], best, best_i)) { best = am_v[i]; best_i = am_i[i]; }
        gptr(a.argmax)[slot] = make_int2(__float_as_int(best), best_i);
      }
    }
  } else if constexpr (!MAG) {
    auto* d = (GLSX_GLOBAL float4*)(gptr((float2*)a.dst) + col0);
#pragma unroll
    for (int u = 0; u < NBL; ++u)
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int pos = ((tq + u * T + q * (L / RL)) + half) & (L - 1);
        const float2 x = va[u + q * NBL], y = vb[u + q * NBL];
        d[(long long)pos * es2] = make_float4(x.x * a.scale, x.y * sy, y.x * a.scale, y.y * sy);
      }
  } else {
    // the two columns' values of a row as one 8-byte store; (max, min) of
    // everything stored, as the one-column kernels keep them
    const bool lm = a.out_kind == OUT_LOGMAG;
    auto* d = (GLSX_GLOBAL float2*)(gptr((float*)a.dst) + col0);
    float mx = -__FLT_MAX__, mn = __FLT_MAX__;
#pragma unroll
    for (int u = 0; u < NBL; ++u)
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int pos = ((tq + u * T + q * (L / RL)) + half) & (L - 1);
        const float2 x = va[u + q * NBL], y = vb[u + q * NBL];
        const float m0 = spectrum_value(make_float2(x.x * a.scale, x.y * sy), lm, a.mag_offset);
        const float m1 = spectrum_value(make_float2(y.x * a.scale, y.y * sy), lm, a.mag_offset);
        d[(long long)pos * es2] = make_float2(m0, m1);
        mx = fmaxf(fmaxf(mx, m0), m1);
        mn = fminf(fminf(mn, m0), m1);
      }
    if (a.argmax) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, off));
        mn = fminf(mn, __shfl_xor(mn, off));
      }
      const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
      if ((threadIdx.x & 63) == 0) {
        am_v[wave] = mx;
        am_i[wave] = __float_as_int(mn);
      }
      lds_barrier();
      if (threadIdx.x == 0) {
        for (int i = 1; i < nw; ++i) {
          mx = fmaxf(mx, am_v[i]);
          mn = fminf(mn, __int_as_float(am_i[i]));
        }
        gptr(a.argmax)[slot] = make_int2(__float_as_int(mx), __float_as_int(mn));
      }
    }
  }
}

template <int L, int P, bool MAG = false, bool REAL = false>
__global__ void __launch_bounds__(P * Sched<L>::T) fft_col_pair_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  col_pair_tile<L, P, MAG, REAL>(a, pair_tile_index<pair_spread(L, P)>(blockIdx.x, gridDim.x), blockIdx.x, lds, am_v,
                                 am_i);
}

// The value the neighbouring lane (lane ^ 1) holds.
__device__ __forceinline__ float lane_pair_swap(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
}

// Column pass with one column per lane and the pair tiles' 16-byte accesses:
// lanes 2k and 2k+1 hold columns 2p and 2p+1 of the tile at the same thread
// index t, so both need rows t + e*T (e = 0..15) of the same two columns.
// Lane h loads rows of odd (h = 1) or even (h = 0) e as 16-byte pairs, and one
// DPP swap per value moves each column's elements to its lane; the stores
// mirror it.  A tile of C columns is C*T threads, each with the 16 elements of
// the one-column-per-lane kernels: twice the waves of a pair tile with half
// the work each, so at 1024 points (a pair tile of 4 columns runs two waves per
// CU) every SIMD gets a wave.  Same stages, so the same bits as the other
// column kernels.  Measured against the pair tiles (same process, in place):
// 1024 x 1024 6.7 against 7.9 us, 1024 x 16384 46.1 against 64.4 us; at 2048
// points 13.0 against 13.6 us at 2048 columns but 109.7 against 107.5 us at
// 16384, so 2048-point columns keep the pair tiles.
//
// Complex output (MAG = REAL = false), |z| / log(|z| + offset) with (max, min)
// partials (MAG), or the real plane with argmax partials (REAL), as the pair
// kernel stores them.  The host checks n_batch % C == 0, unit batch stride,
// an even element stride and aligned buffers (split_tile()).
template <int L, int C, bool MAG = false, bool REAL = false>
__device__ __forceinline__ void col_split_tile(const PassArgs& a, long long tile, long long slot, float2* lds,
                                               float* am_v, int* am_i) {
  constexpr int E = Sched<L>::E, T = Sched<L>::T, NS = Sched<L>::NSTAGE;
  static_assert(E == 16 && C % 2 == 0, "16 elements per thread, whole column pairs");
  const int tid = threadIdx.x;
  const int h = tid & 1;               // column of the lane pair
  const int p = (tid >> 1) % (C / 2);  // pair of the tile
  const int t = (tid >> 1) / (C / 2);  // thread index of the column
  const int col = 2 * p + h;
  const long long col0 = tile * C + 2 * p;
  EarlyBases<L> eb;
  eb.load(t, a.tw);
  float2 v[E];
  {
    // element e of this lane's column is row t + e*T; lane h loads rows e = 2j + h
    const auto* s = (const GLSX_GLOBAL float4*)(gptr((const float2*)a.src) + col0 + (long long)(t + h * T) * a.es);
    const long long step = (long long)T * a.es;  // two schedule rows, in float4
    float4 x[E / 2];
#pragma unroll
    for (int j = 0; j < E / 2; ++j) x[j] = s[j * step];
#pragma unroll
    for (int j = 0; j < E / 2; ++j) {
      // lane 0 keeps column 0 of row 2j and takes column 0 of row 2j+1 from lane 1;
      // lane 1 keeps column 1 of row 2j+1 and takes column 1 of row 2j from lane 0
      // (selects on values in opaque registers: a select between array
      // elements would become a select of addresses and move the array to
      // scratch memory)
      float x0 = x[j].x, x1 = x[j].y, x2 = x[j].z, x3 = x[j].w;
      asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3));
      const float rx = lane_pair_swap(h ? x0 : x2), ry = lane_pair_swap(h ? x1 : x3);
      const float ox = h ? x2 : x0, oy = h ? x3 : x1;
      v[2 * j] = make_float2(h ? rx : ox, h ? ry : oy);
      v[2 * j + 1] = make_float2(h ? ox : rx, h ? oy : ry);
    }
    if (a.conj_in) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
    }
  }
  auto lidx = [&](int pos) { return swz(pos) * C + col; };
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (
        [&] {
          if constexpr (bases_early<L, S>()) stage_butterflies_b<L, S>(v, eb.b[S]);
          else stage_butterflies<L, S>(v, t, a.tw);
          if constexpr (S + 1 < NS) {
            constexpr int R = Sched<L>::radix(S), NB = E / R;
            if constexpr (S > 0) lds_barrier();  // earlier reads of the buffer are done
#pragma unroll
            for (int u = 0; u < NB; ++u)
#pragma unroll
              for (int q = 0; q < R; ++q) lds[lidx(stage_out_pos<L, S>(t, u, q))] = v[u + q * NB];
            lds_barrier();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[lidx(t + e * T)];
          }
        }(),
        ...);
  }(std::make_integer_sequence<int, NS>{});
  // v[q] is row t + q*T (bin t + q*L/16 of the last radix-16 stage); lane h
  // stores rows 2j + h of both columns
  static_assert(Sched<L>::radix(NS - 1) == 16, "last stage radix 16");
  const int half = a.shift ? (L >> 1) : 0;
  const float sy = a.conj_out ? -a.scale : a.scale;
  const long long es2 = a.es / 2;
  if constexpr (!MAG && !REAL) {
    auto* d = (GLSX_GLOBAL float4*)(gptr((float2*)a.dst) + col0);
#pragma unroll
    for (int j = 0; j < E / 2; ++j) {
      float y0x = v[2 * j].x, y0y = v[2 * j].y, y1x = v[2 * j + 1].x, y1y = v[2 * j + 1].y;
      asm volatile("" : "+v"(y0x), "+v"(y0y), "+v"(y1x), "+v"(y1y));
      // lane 0 sends column 0 of row 2j+1, lane 1 column 1 of row 2j
      const float gx = lane_pair_swap(h ? y0x : y1x), gy = lane_pair_swap(h ? y0y : y1y);
      const float c0x = h ? gx : y0x, c0y = h ? gy : y0y, c1x = h ? y1x : gx, c1y = h ? y1y : gy;
      const int pos = ((t + (2 * j + h) * T) + half) & (L - 1);
      d[(long long)pos * es2] = make_float4(c0x * a.scale, c0y * sy, c1x * a.scale, c1y * sy);
    }
  } else {
    // one float per value: the lane's own column, then the pair's two values
    // of a row as one 8-byte store
    const bool lm = a.out_kind == OUT_LOGMAG;
    auto* d = (GLSX_GLOBAL float2*)(gptr((float*)a.dst) + col0);
    float best = -__FLT_MAX__;
    int best_i = REAL ? 0x7fffffff : __float_as_int(__FLT_MAX__);  // MAG: (max, bits of min)
#pragma unroll
    for (int j = 0; j < E / 2; ++j) {
      float r0, r1;  // this lane's column at rows 2j and 2j+1
      if constexpr (REAL) {
        r0 = v[2 * j].x * a.scale;
        r1 = v[2 * j + 1].x * a.scale;
      } else {
        r0 = spectrum_value(make_float2(v[2 * j].x * a.scale, v[2 * j].y * sy), lm, a.mag_offset);
        r1 = spectrum_value(make_float2(v[2 * j + 1].x * a.scale, v[2 * j + 1].y * sy), lm, a.mag_offset);
      }
      asm volatile("" : "+v"(r0), "+v"(r1));
      const float g = lane_pair_swap(h ? r0 : r1);
      const float c0 = h ? g : r0, c1 = h ? r1 : g;
      const int pos = ((t + (2 * j + h) * T) + half) & (L - 1);
      d[(long long)pos * es2] = make_float2(c0, c1);
      if constexpr (REAL) {
        const int li = (int)((long long)pos * a.out_cols + col0);
        if (am_better(c0, li, best, best_i)) { best = c0; best_i = li; }
        if (am_better(c1, li + 1, best, best_i)) { best = c1; best_i = li + 1; }
      } else {
        best = fmaxf(fmaxf(best, c0), c1);
        best_i = __float_as_int(fminf(fminf(__int_as_float(best_i), c0), c1));
      }
    }
    if (a.argmax) {
      auto combine = [&](float ov, int oi) {
        if constexpr (REAL) {
          if (am_better(ov, oi, best, best_i)) { best = ov; best_i = oi; }
        } else {
          best = fmaxf(best, ov);
          best_i = __float_as_int(fminf(__int_as_float(best_i), __int_as_float(oi)));
        }
      };
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) combine(__shfl_xor(best, off), __shfl_xor(best_i, off));
      const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
      if ((threadIdx.x & 63) == 0) {
        am_v[wave] = best;
        am_i[wave] = best_i;
      }
      lds_barrier();
      if (threadIdx.x == 0) {
        for (int i = 1; i < nw; ++i) combine(am_v[i], am_i[i]);
        gptr(a.argmax)[slot] = make_int2(__float_as_int(best), best_i);
      }
    }
  }
}

template <int L, int C, bool MAG = false, bool REAL = false>
__global__ void __launch_bounds__(C * Sched<L>::T) fft_col_split_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  col_split_tile<L, C, MAG, REAL>(a, pair_tile_index<false>(blockIdx.x, gridDim.x), blockIdx.x, lds, am_v, am_i);
}

// Columns per split-kernel tile (1024 points: 8 columns, 512 threads; 4- and
// 16-column tiles measured 7.3 and 9.4 us at 1024 x 1024).
constexpr int kSplitCols = 8;

// Column pairs per pair-kernel tile: 4 (8 columns) at 2048 and 4096 points,
// 2 at 8192 points (1024 threads) and at 1024 points.  Measured at 8192^2 (an
// array larger than the last-level cache): 437 us against 921 us for the
// half kernel.  At 1024^2 the 8-column tiles give only 128 workgroups, half
// the CUs: 4-column tiles (256 workgroups of two waves) take the forward +
// inverse step from 37.8 to 36.4 us (same bits); 2-column tiles measured
// slower (42.1 us).
__host__ __device__ constexpr int col_pairs(int L) { return L >= 8192 || L == 1024 ? 2 : 4; }

// LDS bytes of a pair-kernel tile: P columns (one of every pair) at a time.
__host__ __device__ constexpr size_t col_pair_lds_bytes(int L, int P) {
  return (size_t)P * (L + L / 16) * sizeof(float2);
}

}  // namespace glsx
