// fft_r2c.hip -- gls::fft of a real image (CV_32FC1) through half-size spectra.
//
// The reference merges the real image with an (unset) imaginary plane and
// runs the full complex FFT (glsCV/glsFft.cpp:683-689); a real image's
// spectrum is Hermitian, Y[r][c] = conj(Y[-r][-c]), so half of that work and
// of its traffic is redundant.  Here an M x N real image is transformed in
// two passes:
//   row pass    two real rows as one complex row, z_j = x[2j] + i x[2j+1]
//               (the existing row kernels, IN_R32PAIR): Z_j = X_2j + i X_2j+1,
//               M/2 complex rows (reads 4 and writes 4 bytes per pixel);
//   column pass slot k (1 <= k < N/2) rebuilds column k of the row spectra,
//               X_2j[k] = (Z_j[k] + conj Z_j[N-k]) / 2 and
//               X_2j+1[k] = (Z_j[k] - conj Z_j[N-k]) / 2i,
//               transforms its even and odd rows as two M/2-point FFTs
//               (Ea, Eb: the two transforms per thread of the pair tiles),
//               combines Y[s] = Ea[s] + w_M^s Eb[s], Y[s + M/2] = Ea[s] -
//               w_M^s Eb[s], and stores column k and its mirror column N-k,
//               Y[-s][N-k] = conj Y[s][k].  Columns 0 and N/2 (real row
//               spectra) share slot 0 as the real and imaginary parts of one
//               transform, separated after it through LDS.
//               (reads 4 and writes 8 bytes per pixel).
// 20 bytes per pixel in all, against 28 for the complex transform of the real
// image.  Flags as in fft_kernels.hip: INVERSE conjugates the output (the
// input conjugation is void on real data), SCALE multiplies by 1/(M*N), SHIFT
// moves output (r, c) to (r + M/2, c + N/2).  The column pass can also store
// magnitudes with (max, min) partials, for the spectrum view.
#include <algorithm>

#include "fft_pass.hpp"
#include "fft_wave.hpp"
#include "glsx_internal.hpp"

namespace glsx {

namespace {

struct R2cArgs {
  const float2* z;   // M/2 rows of Z, row j at z + j * zpitch
  long long zpitch;  // in float2
  void* dst;         // M x N: float2 (OUT_C32) or float (OUT_MAG / OUT_LOGMAG)
  const float2* tw;
  int2* partials;    // OUT_MAG / OUT_LOGMAG: per-workgroup (max bits, min bits), or null
  int n;             // columns N (power of two, >= 2)
  int out_kind;
  int conj_out, shift;
  float scale, mag_offset;
  float* half;       // magnitude planes: columns 0..N/2 only, unshifted, row r at
  long long hpitch;  // half + r * hpitch (no mirror stores), or null
};

// LDS of a tile: P transforms of length L2 (one of the two per lane at a
// time), at least 2*L2 for the slot-0 separation.
template <int L2>
constexpr size_t r2c_lds_bytes(int P) {
  return (size_t)P * Sched<L2>::LP * sizeof(float2);
}

// 16 bytes at an 8-byte aligned address (the mirror stores start one column
// off the 16-byte grid)
struct __attribute__((aligned(8))) F4a8 {
  float x, y, z, w;
};

// A tile covers P slots (P >= 2, even) with P groups of T lanes.  Lane group
// b holds slots k = tile*P + (b & ~1) and k + 1 -- per row it loads bins k, k+1
// and N-k-1, N-k as two 16-byte pairs -- and transforms, for both slots, the even rows of
// the column signal (b even: A) or the odd rows (b odd: B).  The neighbouring
// lane (lane ^ 1) holds the other half, so after the stages the two lanes
// swap values (DPP) and combine: the A lane then holds rows s, the B lane rows
// s + M/2, of both columns k and k + 1, stored as one 16-byte access, and
// their mirrors N-k-1, N-k as another.
template <int L2, int P, bool MAG, bool HALF = false>
__global__ void __launch_bounds__(P * Sched<L2>::T) fft_col_r2c_kernel(R2cArgs a) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  constexpr int E = Sched<L2>::E, T = Sched<L2>::T, NS = Sched<L2>::NSTAGE, M = 2 * L2;
  static_assert(P >= 2 && P % 2 == 0, "lane pairs; slot 0's separation needs 2*L2 LDS entries");
  static_assert(MAG || !HALF, "the half plane holds magnitudes");
  const int b = threadIdx.x % P, t = threadIdx.x / P;
  const bool odd = b & 1;  // this lane transforms the odd rows (B) of its two slots
  const int n = a.n, half_n = n >> 1;
  // Workgroups go round-robin over the 8 XCDs; XCD x takes a contiguous run of
  // tiles, so the tiles whose row segments share 128-byte lines -- neighbours
  // on both the column-k side and the (one column off) mirror side -- run on
  // one XCD and its L2 merges their accesses.
  const unsigned G = gridDim.x, w = blockIdx.x;
  const int tile = (G & 7u) == 0 ? (int)((w & 7u) * (G >> 3) + (w >> 3)) : (int)w;
  const int k = tile * P + (b & ~1);  // slots k and k + 1 (k even; slot 0: columns 0 and N/2)
  const bool active = k < half_n;     // (N/2 is even: slot k + 1 exists too)
  float2 va[E], vb[E];                // slot k, slot k + 1
  if (active) {
    // Per row: bins k, k+1 (16 bytes, aligned) and N-k-1, N-k (16 bytes one
    // column off the grid; slot 0: bins N/2 and, unused, N/2+1).  These are
    // the columns this lane stores, so a tile reads only what it overwrites
    // and the pass may run in place (the half spectra in dst's even rows).
    const long long row0 = (long long)t * a.zpitch, step = (long long)T * a.zpitch;
    const auto* zl = gptr(a.z) + row0 + k;
    const auto* zr = gptr(a.z) + row0 + (k == 0 ? half_n : n - k - 1);
    // X_2j[k] = (Z[k] + conj Z[N-k]) / 2, X_2j+1[k] = (Z[k] - conj Z[N-k]) / 2i
    auto half = [&](float2 z1, float2 z2) {
      return odd ? make_float2(0.5f * (z1.y + z2.y), 0.5f * (z2.x - z1.x))
                 : make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
    };
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float4 xl = *(const GLSX_GLOBAL float4*)(zl + e * step);
      const F4a8 xr = *(const GLSX_GLOBAL F4a8*)(zr + e * step);
      const float2 zk = make_float2(xl.x, xl.y), zk1 = make_float2(xl.z, xl.w);
      if (k == 0) {
        // slot 0: real row spectra at bins 0 and N/2 as column 0 + i column
        // N/2 (xr = bins N/2, N/2+1; the second is not used); slot 1's mirror
        // is bin N-1
        va[e] = odd ? make_float2(zk.y, xr.y) : make_float2(zk.x, xr.x);
        vb[e] = half(zk1, (gptr(a.z) + row0 + e * step)[n - 1]);
      } else {
        // xr = bins N-k-1, N-k
        va[e] = half(zk, make_float2(xr.z, xr.w));
        vb[e] = half(zk1, make_float2(xr.x, xr.y));
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) va[e] = vb[e] = make_float2(0.0f, 0.0f);
  }
  auto lidx = [&](int pos) { return swz(pos) * P + b; };
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (
        [&] {
          stage_butterflies<L2, S>(va, t, a.tw);
          stage_butterflies<L2, S>(vb, t, a.tw);
          if constexpr (S + 1 < NS) {
            constexpr int R = Sched<L2>::radix(S), NB = E / R;
            auto exchange = [&](float2(&v)[E], bool first) {
              if (!first) lds_barrier();  // earlier reads of the buffer are done
#pragma unroll
              for (int u = 0; u < NB; ++u)
#pragma unroll
                for (int q = 0; q < R; ++q) lds[lidx(stage_out_pos<L2, S>(t, u, q))] = v[u + q * NB];
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
  // bins held: v[q] is bin s = t + q*(L2/RL) of the lane's transforms
  constexpr int RL = Sched<L2>::radix(NS - 1), NBL = E / RL;
  static_assert(NBL == 1, "the last stage leaves one butterfly per thread");
  auto bin = [&](int q) { return t + q * (L2 / RL); };
  {
    // Y[s] = Ea[s] + w_M^s Eb[s] (A lane), Y[s + M/2] = Ea[s] - w_M^s Eb[s]
    // (B lane); w_M^s = w_M^t (one table read per thread) * w_{2RL}^q (the
    // same for every lane)
    const float2 wt = gptr(a.tw)[tw_offset(M) + t];
    const auto* wq = gptr(a.tw) + tw_offset(2 * RL);
    const float sg = odd ? -1.0f : 1.0f;
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const float2 wm = q == 0 ? wt : cmul(wt, wq[q]);
      const float2 oa = make_float2(lane_xor<0>(va[q].x), lane_xor<0>(va[q].y));
      const float2 ob = make_float2(lane_xor<0>(vb[q].x), lane_xor<0>(vb[q].y));
      const float2 ea = odd ? oa : va[q], eba = odd ? va[q] : oa;
      const float2 eb = odd ? ob : vb[q], ebb = odd ? vb[q] : ob;
      const float2 pa = cmul(eba, wm), pb = cmul(ebb, wm);
      va[q] = make_float2(fmaf(sg, pa.x, ea.x), fmaf(sg, pa.y, ea.y));
      vb[q] = make_float2(fmaf(sg, pb.x, eb.x), fmaf(sg, pb.y, eb.y));
    }
  }
  const int rofs = odd ? L2 : 0;  // this lane's rows: s + rofs
  // Every thread has consumed its loads: with the source in the destination
  // (no SHIFT), no store below can overwrite a value still to be loaded.
  lds_barrier();
  if (tile == 0) {
    // slot 0 holds G = Y0 + i Y(N/2) over all M rows: publish G for the mirror reads
    if (k == 0) {
#pragma unroll
      for (int q = 0; q < RL; ++q) lds[bin(q) + rofs] = va[q];
    }
    lds_barrier();
  }
  const float sy = a.conj_out ? -a.scale : a.scale;
  const int hm = a.shift ? L2 : 0, hn = a.shift ? half_n : 0;
  const bool lm = a.out_kind == OUT_LOGMAG;
  float mx = -__FLT_MAX__, mn = __FLT_MAX__;
  auto fin = [&](float2 y) { return make_float2(y.x * a.scale, y.y * sy); };
  auto mag = [&](float2 v) {
    const float m = spectrum_value(v, lm, a.mag_offset);
    mx = fmaxf(mx, m);
    mn = fminf(mn, m);
    return m;
  };
  auto row_ptr = [&](int r) { return (long long)((r + hm) & (M - 1)) * n; };
  // the half plane: the mirror columns hold the same magnitudes, so only
  // columns 0..N/2 are stored (and the partials are those of the whole plane)
  constexpr bool half = HALF;
  auto half_ptr = [&](int r, int c) { return gptr(a.half) + (long long)r * a.hpitch + c; };
  // a single column value (slot 0's columns N/2 and N-1)
  auto put1 = [&](int r, int c, float2 y) {
    const long long o = row_ptr(r) + ((c + hn) & (n - 1));
    const float2 v = fin(y);
    if constexpr (MAG) {
      if (half) *half_ptr(r, c) = mag(v);
      else gptr((float*)a.dst)[o] = mag(v);
    } else {
      gptr((float2*)a.dst)[o] = v;
    }
  };
  // columns c, c + 1 (c even): one 16-byte store (8 bytes of magnitudes)
  auto put2 = [&](int r, int c, float2 y0, float2 y1) {
    const long long o = row_ptr(r) + ((c + hn) & (n - 1));
    const float2 v0 = fin(y0), v1 = fin(y1);
    if constexpr (MAG) {
      auto* d = half ? half_ptr(r, c) : gptr((float*)a.dst) + o;
      *(GLSX_GLOBAL float2*)d = make_float2(mag(v0), mag(v1));
    } else {
      *(GLSX_GLOBAL float4*)(gptr((float2*)a.dst) + o) = make_float4(v0.x, v0.y, v1.x, v1.y);
    }
  };
  // Mirror columns: column N-c of row -r holds conj Y_c[r].  Complex output
  // from tiles of 8 or more slots stores them as aligned pairs (N-k-2, N-k-1)
  // -- slot k+1 of this lane and slot k+2, which lane b+2 holds for the same
  // rows and passes through LDS -- so every mirror store is a whole 16-byte
  // access; the tile's last slot pair stores its column N-k-1 alone, and its
  // first pair column N-k alone (the mirror of the tile's first slot, which
  // the previous tile cannot reach; in tile 0 that is column 0 itself).
  // 4096^2 forward: 78-82 against 84-86 us, write traffic 1.25x against
  // 1.29x the algorithmic bytes.  Elsewhere the LDS pass costs more than it
  // saves (8192^2, 4-slot tiles: 566-576 against 537-547 us; magnitude
  // planes, whose mirrors are 8 bytes: 81-83 against 78-80 us at 4096^2), and
  // the mirrors go one column off the grid.  (Passing slot k+2 in registers
  // spills: the 1024-thread tiles use 119-127 of their 128 registers.)
  constexpr bool kPairs = !MAG && P >= 8;
  auto put2m = [&](int r, int c, float2 y0, float2 y1) {  // columns c, c + 1, c odd
    const long long o = row_ptr(r) + ((c + hn) & (n - 1));
    const float2 v0 = fin(y0), v1 = fin(y1);
    if constexpr (MAG) {
      auto* d = gptr((float*)a.dst) + o;
      d[0] = mag(v0);
      d[1] = mag(v1);
    } else {
      *(GLSX_GLOBAL F4a8*)(gptr((float2*)a.dst) + o) = F4a8{v0.x, v0.y, v1.x, v1.y};
    }
  };
  constexpr bool mirrors_here = !HALF && !kPairs;  // mirror stores beside the column-k stores
  if (active && k == 0) {
    // Y0[r] = (G[r] + conj G[-r]) / 2, Y(N/2)[r] = (G[r] - conj G[-r]) / 2i;
    // slot 1 as any other
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int r = bin(q) + rofs, rm = (M - r) & (M - 1);
      const float2 g = va[q], m = lds[rm];
      put2(r, 0, make_float2(0.5f * (g.x + m.x), 0.5f * (g.y - m.y)), vb[q]);
      put1(r, half_n, make_float2(0.5f * (g.y + m.y), 0.5f * (m.x - g.x)));
      if constexpr (mirrors_here) put1(rm, n - 1, cconj(vb[q]));
    }
  } else if (active) {
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int r = bin(q) + rofs, rm = (M - r) & (M - 1);
      put2(r, k, va[q], vb[q]);
      if constexpr (mirrors_here) put2m(rm, n - k - 1, cconj(vb[q]), cconj(va[q]));
    }
  }
  if constexpr (kPairs && !HALF) {
    const int pr = b >> 1;  // slot pair of the tile
    const bool pair_ok = pr < P / 2 - 1 && k + 2 < half_n;
    const bool first_single = pr == 0 && k != 0;
    const int nt = blockDim.x;
    lds_barrier();  // the slot-0 reads of LDS are done
#pragma unroll
    for (int q = 0; q < RL; ++q) lds[q * nt + threadIdx.x] = va[q];
    lds_barrier();
    if (active) {
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int rm = (M - (bin(q) + rofs)) & (M - 1);
        if (pair_ok) put2(rm, n - k - 2, cconj(lds[q * nt + threadIdx.x + 2]), cconj(vb[q]));
        else put1(rm, n - k - 1, cconj(vb[q]));
        if (first_single) put1(rm, n - k, cconj(va[q]));
      }
    }
  }
  if constexpr (MAG) {
    if (a.partials) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        if (off >= (int)blockDim.x) continue;
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
        for (int v = 1; v < nw; ++v) {
          mx = fmaxf(mx, am_v[v]);
          mn = fminf(mn, __int_as_float(am_i[v]));
        }
        gptr(a.partials)[blockIdx.x] = make_int2(__float_as_int(mx), __float_as_int(mn));
      }
    }
  }
}

// Slots per tile: the widest tile (longer row segments) that still gives
// every CU a tile, within 1024 threads and the LDS.  The glsCam view's half
// plane takes at most 4 slots: its column pass stores half the bytes, and two
// 512-thread tiles share a CU instead of one 1024-thread tile running alone
// (4096^2 view 83.0-87.7 against 87.1-90.1 us).  (Whole magnitude planes
// with 4 slots: 52.0 against 56.1 us at 4096^2 on a half spectrum that no
// kernel has just written, but in the view, whose row pass writes it,
// 104.7-110.5 against 101.8-106.1 us; complex output 80.2 against 60.0 us.)
template <int L2>
int r2c_slots(const DeviceState* ds, int n, bool half) {
  constexpr int T = Sched<L2>::T;
  int p = half ? 4 : 16;
  while (p > 2 && (p * T > 1024 || r2c_lds_bytes<L2>(p) + 1024 > ds->lds_per_block)) p >>= 1;
  while (p > 2 && (n / 2) / p < ds->num_cu) p >>= 1;
  return p;
}

template <int L2, int P, bool MAG>
constexpr bool r2c_fits() {
  return P * Sched<L2>::T <= 1024;
}

template <int L2, int P, bool MAG, bool HALF = false>
int launch_r2c_p(const DeviceState* ds, const R2cArgs& a, long long grid, hipStream_t s) {
  constexpr int T = Sched<L2>::T;
  if constexpr (!r2c_fits<L2, P, MAG>()) {
    return fail(GLSX_ERR_STATE, "real-input column tile too large");
  } else {
    auto kern = fft_col_r2c_kernel<L2, P, MAG, HALF>;
    const size_t lds = r2c_lds_bytes<L2>(P);
    if (lds + 1024 > ds->lds_per_block) return fail(GLSX_ERR_STATE, "real-input column tile needs too much LDS");
    static LdsLimit lim;
    if (int rc = lim.raise((const void*)kern, ds->lds_per_block - 1024)) return rc;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(P * T), lds, s, a);
    GLSX_HIP(hipGetLastError());
    return GLSX_OK;
  }
}

// complex output, a whole magnitude plane, or the half plane
template <int L2, int P>
int launch_r2c_pk(const DeviceState* ds, const R2cArgs& a, bool mag, long long grid, hipStream_t s) {
  if (!mag) return launch_r2c_p<L2, P, false>(ds, a, grid, s);
  return a.half ? launch_r2c_p<L2, P, true, true>(ds, a, grid, s) : launch_r2c_p<L2, P, true>(ds, a, grid, s);
}

template <int L2>
int launch_r2c_l(const DeviceState* ds, const R2cArgs& a, bool mag, hipStream_t s, long long* grid_out) {
  int p = r2c_slots<L2>(ds, a.n, a.half != nullptr);
  const long long grid = ((long long)(a.n / 2) + p - 1) / p;
  if (grid_out) *grid_out = grid;
  if (!a.dst) return GLSX_OK;  // grid query only
  switch (p) {
    case 2: return launch_r2c_pk<L2, 2>(ds, a, mag, grid, s);
    case 4: return launch_r2c_pk<L2, 4>(ds, a, mag, grid, s);
    case 8: return launch_r2c_pk<L2, 8>(ds, a, mag, grid, s);
    default: return launch_r2c_pk<L2, 16>(ds, a, mag, grid, s);
  }
}

int launch_r2c(const DeviceState* ds, int rows, const R2cArgs& a, bool mag, hipStream_t s, long long* grid_out) {
  // (the lengths r2c_applies admits)
  switch (rows / 2) {
    case 2048: return launch_r2c_l<2048>(ds, a, mag, s, grid_out);
    case 4096: return launch_r2c_l<4096>(ds, a, mag, s, grid_out);
    case 8192: return launch_r2c_l<8192>(ds, a, mag, s, grid_out);
    default: return fail(GLSX_ERR_UNSUPPORTED, "real-input FFT: rows must be 4096, 8192 or 16384");
  }
}

}  // namespace

// Where the half-size spectra are faster than the complex transform of the
// real image (measured on MI355X, forward FFT / glsCam view, microseconds,
// half-size vs complex): 4096^2 86.6 vs 88.5 / 105.4 vs 121.8; 8192^2 543 vs
// 594 / 701 vs 738; 2048^2 34.2 vs 31.5 / 45.3 vs 40.2; 1024^2 21.4 vs 18.7;
// 256^2 15.6 vs 7.4.  Each column tile carries two transforms per lane and
// runs loads, stages and stores as one chain per CU, which shorter columns
// (fewer, smaller tiles) do not amortise.  A shifted 8192^2 FFT through the
// stream's scratch measured 604 vs 575: shifted transforms above 4096^2 keep
// the complex path.
bool r2c_applies(int rows, int cols, int src_type, int flags) {
  if (src_type != GLSX_CV_32FC1 || (flags & GLSX_FFT_SPLIT_MFMA)) return false;
  if (rows < 4096 || cols < 4096) return false;
  return !((flags & GLSX_FFT_SHIFT) && (long long)rows * cols > 4096LL * 4096);
}

// The magnitude planes follow the same rule, so that the fused forms keep the
// bits of the FFT followed by the separate magnitude call.
bool r2c_applies_mag(int rows, int cols, int src_type, int flags) { return r2c_applies(rows, cols, src_type, flags); }

long long r2c_partials(const DeviceState* ds, int rows, int cols) {
  R2cArgs a{};
  a.n = cols;
  long long g = 0;
  if (launch_r2c(ds, rows, a, true, nullptr, &g)) return -1;
  return g;
}

int fft2d_r2c(DeviceState* ds, const float* src, void* dst, int rows, int cols, int flags, int out_kind,
              float mag_offset, int2* partials, long long* n_partials, void* zbuf, hipStream_t stream,
              float* half, long long hpitch) {
  if (!r2c_applies(rows, cols, GLSX_CV_32FC1, flags)) return fail(GLSX_ERR_STATE, "real-input FFT: not applicable");
  const bool shf = flags & GLSX_FFT_SHIFT;
  // Z (M/2 rows, pitch 2N): in the destination's even rows, or in zbuf when
  // the destination is a real plane or the shifted stores would reach columns
  // other slots still have to read
  float2* z = (float2*)(zbuf ? zbuf : dst);
  if (!zbuf && (out_kind != OUT_C32 || shf)) return fail(GLSX_ERR_STATE, "real-input FFT: needs a buffer for Z");
  PassDesc r;
  r.L = cols;
  r.col = false;
  r.src = src;
  r.src2 = src + cols;
  r.dst = z;
  r.n_batch = rows / 2;
  r.es = 1;
  r.bs = 2LL * cols;
  r.in_kind = IN_R32PAIR;
  r.out_kind = OUT_C32;
  if (int rc = launch_pass(ds, r, stream)) return rc;
  R2cArgs a{};
  a.z = z;
  a.zpitch = 2LL * cols;
  a.dst = dst;
  a.tw = ds->twiddles;
  a.partials = partials;
  a.n = cols;
  a.out_kind = out_kind;
  a.conj_out = (flags & GLSX_FFT_INVERSE) != 0;
  a.shift = shf;
  a.scale = (flags & GLSX_FFT_SCALE) ? 1.0f / ((float)rows * (float)cols) : 1.0f;
  a.mag_offset = mag_offset;
  if (half) {
    if (out_kind != OUT_MAG && out_kind != OUT_LOGMAG)
      return fail(GLSX_ERR_STATE, "real-input FFT: half plane of a complex output");
    if (hpitch < cols / 2 + 1 || hpitch % 2 != 0 || (uintptr_t)half % 8 != 0)
      return fail(GLSX_ERR_STATE, "real-input FFT: half-plane pitch or alignment");
    a.half = half;
    a.hpitch = hpitch;
  }
  long long g = 0;
  const int rc = launch_r2c(ds, rows, a, out_kind == OUT_MAG || out_kind == OUT_LOGMAG, stream, &g);
  if (n_partials) *n_partials = g;
  return rc;
}

}  // namespace glsx
