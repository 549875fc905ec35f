// fft_kernels.hip -- twiddle generation, pass dispatch and the 2D FFT driver.
//
// A rows x cols 2D FFT is two sweeps over HBM:
//   row pass    : cols-point transforms of every row, src -> dst
//   column pass : rows-point transforms of every column, dst -> dst (in place)
// The reference runs one fragment pass per radix-4 stage plus tile copies
// (glsCV/glsFft.cpp:494-671: 80 draws and 14 sweeps at 4096^2); here every
// stage of a dimension stays on chip.  Flags follow glsFft.cpp:586-664:
// INVERSE conjugates the row pass input and the column pass output, SCALE
// multiplies by 1/len per transformed dimension, SHIFT adds len/2 (mod len)
// to the output index of every transformed dimension (the {2,3,0,1} tile
// permutation of :656-664).
#include <algorithm>
#include <utility>

#include "fft_pass.hpp"
#include "glsx_internal.hpp"
#include "mfma_fft.hpp"
#include "taskgraph.hpp"

namespace glsx {

// Entry i of table P is exp(-2*pi*i*I/P), computed in FP64 and rounded once.
__global__ void twiddle_kernel(float2* tw) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= kTwiddleEntries) return;
  // find P with P-2 <= gid < 2P-2
  int P = 2;
  while (tw_offset(2 * P) <= gid) P *= 2;
  const int i = gid - tw_offset(P);
  double s, c;
  sincospi(-2.0 * (double)i / (double)P, &s, &c);
  tw[gid] = make_float2((float)c, (float)s);
}

int init_twiddles(DeviceState* ds) {
  // the tables, then the split-precision passes' DFT fragments (mfma_fft.hpp)
  GLSX_HIP(hipMalloc(&ds->twiddles, sizeof(float2) * kTwiddleEntries + kMfmaFragBytes));
  const int threads = 256, blocks = (kTwiddleEntries + threads - 1) / threads;
  hipLaunchKernelGGL(twiddle_kernel, dim3(blocks), dim3(threads), 0, 0, ds->twiddles);
  GLSX_HIP(hipGetLastError());
  hipLaunchKernelGGL(mfma_fragments_kernel, dim3(1), dim3(64), 0, 0, (uint4*)(ds->twiddles + kTwiddleEntries));
  GLSX_HIP(hipGetLastError());
  GLSX_HIP(hipDeviceSynchronize());
  return GLSX_OK;
}

// ---------------------------------------------------------------------------
// pass geometry
// ---------------------------------------------------------------------------
namespace {

int log2i(long long x) {
  int r = 0;
  while ((1ll << r) < x) ++r;
  return r;
}

struct Geometry {
  int w = 1;            // transforms per workgroup
  int threads = 64;     // workgroup size
  size_t lds_bytes = 0; // dynamic LDS
  size_t lds_limit = 0; // device limit per workgroup
  int num_cu = 0;
  long long grid = 0;
};

int sched_T(int L) { return L < 16 ? 1 : L / 16; }
int sched_nstage(int L) {
  const int lg = log2i(L);
  return L < 16 ? 1 : (lg % 4 == 0 ? lg / 4 : lg / 4 + 1);
}

Geometry geometry(const DeviceState* ds, const PassDesc& p) {
  Geometry g;
  const int L = p.L, T = sched_T(L);
  const bool needs_lds = sched_nstage(L) > 1;
  if (!p.col) {
    // row mode: one wave per workgroup for short rows (more, smaller
    // workgroups measured faster than 256-thread groups up to 1024 points)
    g.w = T >= 64 ? 1 : 64 / T;
    while (g.w > 1 && g.w / 2 >= p.n_batch) g.w >>= 1;
    g.threads = g.w * T;
    // (lengths in wave_rows() exchange inside the wave and need no LDS)
    g.lds_bytes = needs_lds && !wave_rows(L) ? (size_t)g.w * (L + L / 16 + 1) * sizeof(float2) : 0;
  } else {
    // column mode: W neighbouring columns per workgroup.  Wider W means longer
    // contiguous row segments (W*8 bytes); bounded by 1024 threads and by LDS,
    // and narrowed (down to 4) while the grid would leave half the CUs idle.
    const size_t per_col = needs_lds ? (size_t)(L + L / 16) * sizeof(float2) : 0;
    const size_t lds_budget = ds->lds_per_block > 4096 ? ds->lds_per_block - 4096 : 0;
    int w = L <= 16 ? 64 : 16;
    while (w > 1 && (w * T > 1024 || (per_col && w * per_col > lds_budget))) w >>= 1;
    while (w > 1 && w / 2 >= p.n_batch) w >>= 1;
    while (w > 4 && (p.n_batch + w - 1) / w < ds->num_cu / 2) w >>= 1;
    g.w = w;
    g.threads = w * T;
    g.lds_bytes = (size_t)w * per_col;
  }
  g.grid = (p.n_batch + g.w - 1) / g.w;
  g.lds_limit = ds->lds_per_block;
  g.num_cu = ds->num_cu;
  return g;
}

// Column passes that run two columns per lane (fft_col_pair_kernel): 1024-,
// 2048-, 4096- and 8192-point complex columns in whole tiles of 2 * col_pairs(L)
// columns (8; 4 at 1024 and 8192 points) with aligned rows, storing
// complex values without partials, or magnitudes (OUT_MAG / OUT_LOGMAG,
// optional min/max partials).  Returns the column pairs per tile, 0 when the
// pass runs the one-column-per-lane kernels.  launch_pass and pass_grid both
// decide with it, so partial counts match the launch.
int pair_tile(const PassDesc& p, size_t lds_limit) {
  if (!p.col || p.split_mfma || (p.L != 1024 && p.L != 2048 && p.L != 4096 && p.L != 8192)) return 0;
  const int P = col_pairs(p.L);
  if (p.in_kind != IN_C32 || p.bs != 1 || p.es % 2 != 0 || p.n_batch <= 0 || p.n_batch % (2 * P) != 0) return 0;
  // complex output without partials, magnitudes (optional min/max partials)
  // or the real plane (optional argmax partials)
  const bool plane = p.out_kind == OUT_MAG || p.out_kind == OUT_LOGMAG || p.out_kind == OUT_R32;
  if (!plane && !(p.out_kind == OUT_C32 && !p.argmax)) return 0;
  // the real plane with argmax partials needs more than the 128 registers of a
  // 1024-thread tile (it spills at 4096 and 8192 points): those columns keep the
  // one-column-per-lane kernels
  if (p.out_kind == OUT_R32 && p.L > 2048) return 0;
  if ((uintptr_t)p.src % 16 != 0 || (uintptr_t)p.dst % (plane ? 8 : 16) != 0) return 0;
  if (col_pair_lds_bytes(p.L, P) + 1024 > lds_limit) return 0;
  return P;
}

// Column passes that run one column per lane with 16-byte accesses shared by
// lane pairs (fft_col_split_kernel): 1024-point columns, and 2048-point
// columns up to kSplitMaxCols2048 of them with complex or magnitude output,
// in whole tiles of kSplitCols columns, with the pair tiles' output kinds and
// alignment.  Returns the columns per tile, 0 when the pass runs another
// kernel; checked before pair_tile, and pass_grid decides with it too.
// (2048-point columns: 13.0 against 13.6 us for the pair tiles at 2048
// columns, 55.4 against 57.8 us at 8192, 109.7 against 107.5 us at 16384;
// 2048^2 step 53.1-53.7 against 54.5-54.7 us, glsCam view 38.4 against 39.3
// us; the real plane with argmax partials measured slower in the batched
// phase-correlation plan, 2.443 against 2.407-2.419 ms, and keeps the pair
// tiles at 2048 points.)
constexpr long long kSplitMaxCols2048 = 8192;
int split_tile(const PassDesc& p, size_t lds_limit) {
  const bool split2048 = p.L == 2048 && p.n_batch <= kSplitMaxCols2048 && p.out_kind != OUT_R32;
  if (!p.col || p.split_mfma || (p.L != 1024 && !split2048)) return 0;
  constexpr int C = kSplitCols;
  if (p.in_kind != IN_C32 || p.bs != 1 || p.es % 2 != 0 || p.n_batch <= 0 || p.n_batch % C != 0) return 0;
  const bool plane = p.out_kind == OUT_MAG || p.out_kind == OUT_LOGMAG || p.out_kind == OUT_R32;
  if (!plane && !(p.out_kind == OUT_C32 && !p.argmax)) return 0;
  if ((uintptr_t)p.src % 16 != 0 || (uintptr_t)p.dst % (plane ? 8 : 16) != 0) return 0;
  if ((size_t)C * (p.L + p.L / 16) * sizeof(float2) + 1024 > lds_limit) return 0;
  return C;
}

// The split-precision column pass over C-column tiles with 16-byte loads
// (fft_mfma_cols_kernel): 1024- and 2048-point complex columns in whole
// tiles with aligned rows, complex output.  Returns C, or 0 when it does not
// apply.  8 columns: 4-column tiles (two per CU) measured no faster at 2048^2
// (forward 32.7 against 32.3 us) and 2% faster at 1024^2.  (The pair tiles
// with matrix-core butterflies compute the same bits but measured slower:
// 2048 x 2048 15.5 against 15.2 us, 1024 x 1024 8.1 against 7.1 us.)
constexpr int kMfmaCols = 8;
int mfma_cols_tile(const PassDesc& p, size_t lds_limit) {
  if (!p.col || (p.L != 1024 && p.L != 2048)) return 0;
  if (p.in_kind != IN_C32 || p.out_kind != OUT_C32 || p.argmax || p.bs != 1 || p.es % 2 != 0 || p.n_batch <= 0 ||
      p.n_batch % kMfmaCols != 0)
    return 0;
  if ((uintptr_t)p.src % 16 != 0 || (uintptr_t)p.dst % 8 != 0) return 0;
  if ((size_t)kMfmaCols * (p.L + p.L / 16) * sizeof(float2) + 1024 > lds_limit) return 0;
  return kMfmaCols;
}

// The plain row passes: rows that one wave exchanges (32, 64, 512, 1024
// points) and whole contiguous rows of 2048 points and more, one per
// workgroup.  MF: radix-16 butterflies on the matrix cores.  Returns false
// when the pass is not such a row pass.
template <int L, bool MF>
bool launch_plain_rows(const Geometry& g, const PassArgs& a, hipStream_t stream, int* rc) {
  const bool loads = a.in_kind == IN_C32 || a.in_kind == IN_R32 || a.in_kind == IN_R32PAIR;
  if constexpr (wave_rows(L) && !MF) {
    if (loads && a.out_kind == OUT_C32 && !a.argmax && a.es == 1 && a.w >= 1 && a.n_batch % a.w == 0 &&
        g.threads == a.w * Sched<L>::T && g.threads <= 64) {
      hipLaunchKernelGGL(fft_row_wave_kernel<L>, dim3((unsigned)g.grid), dim3(g.threads), 0, stream, a);
      *rc = hipGetLastError() == hipSuccess ? GLSX_OK : fail(GLSX_ERR_HIP, "row pass launch failed");
      return true;
    }
  }
  if constexpr (L >= 2048) {
    if (a.w == 1 && a.es == 1 && loads && a.out_kind == OUT_C32 && !a.argmax) {
      auto kern = fft_row_kernel<L, MF>;
      static LdsLimit lim;
      if ((*rc = lim.raise((const void*)kern, g.lds_limit - 1024))) return true;
      hipLaunchKernelGGL(kern, dim3((unsigned)g.grid), dim3(Sched<L>::T), g.lds_bytes, stream, a);
      *rc = hipGetLastError() == hipSuccess ? GLSX_OK : fail(GLSX_ERR_HIP, "row pass launch failed");
      return true;
    }
  }
  return false;
}

template <int L, bool COL>
int launch_t(const Geometry& g, const PassArgs& a, bool split_mfma, int pair, int split, hipStream_t stream) {
  if constexpr (L >= 16) {
    if (split_mfma) {
      // C-column tiles with 16-byte loads (mfma_cols_tile() chose them)
      if constexpr (COL && (L == 1024 || L == 2048)) {
        if (pair) {
          constexpr int C = kMfmaCols;
          constexpr size_t lds = (size_t)C * (L + L / 16) * sizeof(float2);
          auto kern = fft_mfma_cols_kernel<L, C>;
          static LdsLimit lim;
          if (int rc = lim.raise((const void*)kern, g.lds_limit - 1024)) return rc;
          const long long grid = a.n_batch / C;
          if (grid > 0x7fffffffll) return fail(GLSX_ERR_UNSUPPORTED, "batch too large for one launch");
          hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(C * L / 16), lds, stream, a);
          GLSX_HIP(hipGetLastError());
          return GLSX_OK;
        }
      }
      // whole rows of 2048 points and more: the plain row kernel with
      // matrix-core butterflies (the same bits as the tile below; 2048 x 2048
      // 13.7 against 14.5 us, 4096 x 4096 47.2 against 54.4 us)
      if constexpr (!COL && L >= 2048) {
        int rc = GLSX_OK;
        if (launch_plain_rows<L, true>(g, a, stream, &rc)) return rc;
      }
      // every other shape: the tile of mfma_fft.hpp, rounded up to whole
      // waves; the data passes through LDS in the products' layout
      auto kern = fft_mfma_kernel<L, COL>;
      const int threads = (g.threads + 63) & ~63;
      const size_t lds = COL ? (size_t)a.w * (L + L / 16) * sizeof(float2)
                             : (size_t)a.w * Sched<L>::LP * sizeof(float2);
      if (lds + 1024 > g.lds_limit) return fail(GLSX_ERR_UNSUPPORTED, "MFMA pass needs too much LDS");
      static LdsLimit lim;
      if (int rc = lim.raise((const void*)kern, g.lds_limit - 1024)) return rc;
      hipLaunchKernelGGL(kern, dim3((unsigned)g.grid), dim3(threads), lds, stream, a);
      GLSX_HIP(hipGetLastError());
      return GLSX_OK;
    }
  }
  // one column per lane, 16-byte accesses shared by lane pairs (split_tile() chose it)
  if constexpr (COL && (L == 1024 || L == 2048)) {
    if (split) {
      constexpr int C = kSplitCols;
      constexpr size_t lds = (size_t)C * (L + L / 16) * sizeof(float2);
      const int kind = a.out_kind == OUT_C32 ? 0 : a.out_kind == OUT_R32 ? 2 : 1;
      auto kern = kind == 0   ? fft_col_split_kernel<L, C, false, false>
                  : kind == 1 ? fft_col_split_kernel<L, C, true, false>
                              : fft_col_split_kernel<L, C, false, true>;
      static LdsLimit lim[3];
      if (int rc = lim[kind].raise((const void*)kern, g.lds_limit - 1024)) return rc;
      const long long grid = a.n_batch / C;
      if (grid > 0x7fffffffll) return fail(GLSX_ERR_UNSUPPORTED, "batch too large for one launch");
      hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(C * Sched<L>::T), lds, stream, a);
      GLSX_HIP(hipGetLastError());
      return GLSX_OK;
    }
  }
  // two columns per lane (pair_tile() chose it)
  if constexpr (COL && (L == 1024 || L == 2048 || L == 4096 || L == 8192)) {
    if (pair) {
      constexpr int P = col_pairs(L);
      constexpr size_t lds = col_pair_lds_bytes(L, P);
      const int kind = a.out_kind == OUT_C32 ? 0 : a.out_kind == OUT_R32 ? 2 : 1;
      auto kern = kind == 0 ? fft_col_pair_kernel<L, P, false, false> : fft_col_pair_kernel<L, P, true, false>;
      if (kind == 2) {
        // (the real plane runs on pair tiles up to 2048 points, see pair_tile)
        if constexpr (L <= 2048) kern = fft_col_pair_kernel<L, P, false, true>;
        else return fail(GLSX_ERR_STATE, "real-plane pair tile above 2048 points");
      }
      static LdsLimit lim[3];
      if (int rc = lim[kind].raise((const void*)kern, g.lds_limit - 1024)) return rc;
      const long long grid = a.n_batch / (2 * P);
      if (grid > 0x7fffffffll) return fail(GLSX_ERR_UNSUPPORTED, "batch too large for one launch");
      hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(P * Sched<L>::T), lds, stream, a);
      GLSX_HIP(hipGetLastError());
      return GLSX_OK;
    }
  }
  if constexpr (!COL) {
    int rc = GLSX_OK;
    if (launch_plain_rows<L, false>(g, a, stream, &rc)) return rc;
  }
  // Long columns whose 16-element tile would fill a CU's LDS (1024 threads)
  // run as 512-thread tiles with 32 elements per thread and a two-round LDS
  // exchange: half the LDS, two workgroups per CU, same bits.  (Measured
  // faster from 4096 points; at 1024 and 2048 the 16-element tiles win.)
  if constexpr (COL && L >= 4096 && L <= 8192) {
    if (g.threads == 1024 && a.w >= 2) {
      auto kern = fft_col_half_kernel<L>;
      static LdsLimit lim;
      if (int rc = lim.raise((const void*)kern, g.lds_limit - 1024)) return rc;
      hipLaunchKernelGGL(kern, dim3((unsigned)g.grid), dim3(512), g.lds_bytes / 2, stream, a);
      GLSX_HIP(hipGetLastError());
      return GLSX_OK;
    }
  }
  auto kern = fft_pass_kernel<L, COL>;
  static LdsLimit lim;
  if (g.lds_bytes > 0)
    if (int rc = lim.raise((const void*)kern, g.lds_limit - 1024)) return rc;
  hipLaunchKernelGGL(kern, dim3((unsigned)g.grid), dim3(g.threads), g.lds_bytes, stream, a);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

template <bool COL>
int dispatch(int L, const Geometry& g, const PassArgs& a, bool m, int pr, int sp, hipStream_t s) {
  switch (L) {
    case 1: return launch_t<1, COL>(g, a, m, pr, sp, s);
    case 2: return launch_t<2, COL>(g, a, m, pr, sp, s);
    case 4: return launch_t<4, COL>(g, a, m, pr, sp, s);
    case 8: return launch_t<8, COL>(g, a, m, pr, sp, s);
    case 16: return launch_t<16, COL>(g, a, m, pr, sp, s);
    case 32: return launch_t<32, COL>(g, a, m, pr, sp, s);
    case 64: return launch_t<64, COL>(g, a, m, pr, sp, s);
    case 128: return launch_t<128, COL>(g, a, m, pr, sp, s);
    case 256: return launch_t<256, COL>(g, a, m, pr, sp, s);
    case 512: return launch_t<512, COL>(g, a, m, pr, sp, s);
    case 1024: return launch_t<1024, COL>(g, a, m, pr, sp, s);
    case 2048: return launch_t<2048, COL>(g, a, m, pr, sp, s);
    case 4096: return launch_t<4096, COL>(g, a, m, pr, sp, s);
    case 8192: return launch_t<8192, COL>(g, a, m, pr, sp, s);
    case 16384: return launch_t<16384, COL>(g, a, m, pr, sp, s);
    default: return fail(GLSX_ERR_UNSUPPORTED, "FFT length " + std::to_string(L) + " not supported");
  }
}

}  // namespace

int plain_shape(const DeviceState* ds, const PassDesc& p, PlainShape* s) {
  *s = PlainShape{};
  s->L = p.L;
  const int L = p.L, T = sched_T(L);
  if (p.n_batch <= 0) return fail(GLSX_ERR_ARG, "pass shape: empty pass");
  if (p.split_mfma) return fail(GLSX_ERR_UNSUPPORTED, "pass shape: split-precision passes are not grouped");
  if (!p.col && p.in_kind == IN_POCZ) {
    // row pairs (r, -r) of a phase correlation: pocz_row_kernel
    s->kernel = GK_POCZ;
    s->threads = std::max(pocz_threads(L), 64);
    s->lds = pocz_lds_bytes(L);
    s->n_tiles = p.n_batch / 2 + 1;
    s->w = 2;
  } else if (const int C = split_tile(p, ds->lds_per_block)) {
    s->kernel = p.out_kind == OUT_C32 ? GK_SPLIT : p.out_kind == OUT_R32 ? GK_SPLIT_REAL : GK_SPLIT_MAG;
    s->threads = C * T;
    s->lds = (size_t)C * (L + L / 16) * sizeof(float2);
    s->n_tiles = p.n_batch / C;
    s->w = C;
  } else if (const int P = pair_tile(p, ds->lds_per_block)) {
    s->kernel = p.out_kind == OUT_C32 ? GK_PAIR : p.out_kind == OUT_R32 ? GK_PAIR_REAL : GK_PAIR_MAG;
    s->threads = P * T;
    s->lds = col_pair_lds_bytes(L, P);
    s->n_tiles = p.n_batch / (2 * P);
    s->pairs = P;
    s->w = 2 * P;
  } else {
    const Geometry g = geometry(ds, p);
    s->w = g.w;
    s->n_tiles = g.grid;
    s->threads = g.threads;
    s->lds = g.lds_bytes;
    const bool loads = p.in_kind == IN_C32 || p.in_kind == IN_R32 || p.in_kind == IN_R32PAIR;
    const bool plain_out = loads && p.out_kind == OUT_C32 && !p.argmax && p.es == 1;
    // the same choices as launch_t
    if (!p.col && L >= 16 && wave_rows(L) && plain_out && p.n_batch % g.w == 0 && g.threads == g.w * T &&
        g.threads <= 64) {
      s->kernel = GK_ROW_WAVE;
      s->lds = 0;
    } else if (!p.col && L >= 2048 && plain_out && g.w == 1) {
      s->kernel = GK_ROW;
      s->threads = T;
    } else if (p.col && L >= 4096 && L <= 8192 && g.threads == 1024 && g.w >= 2) {
      s->kernel = GK_COL_HALF;
      s->threads = 512;
      s->lds = g.lds_bytes / 2;
    } else {
      s->kernel = p.col ? GK_PASS_COL : GK_PASS_ROW;
    }
  }
  if (s->lds > ds->lds_per_block) return fail(GLSX_ERR_UNSUPPORTED, "pass shape: tile needs too much LDS");
  if (s->n_tiles > 0x7fffffffll) return fail(GLSX_ERR_UNSUPPORTED, "pass shape: too many tiles");
  return GLSX_OK;
}

long long plain_tile(const PlainShape& s, long long wg) {
  switch (s.kernel) {
    case GK_PAIR:
    case GK_PAIR_MAG:
    case GK_PAIR_REAL:
      return pair_spread(s.L, s.pairs) ? pair_tile_index<true>(wg, s.n_tiles) : pair_tile_index<false>(wg, s.n_tiles);
    case GK_SPLIT:
    case GK_SPLIT_MAG:
    case GK_SPLIT_REAL:
      return pair_tile_index<false>(wg, s.n_tiles);
    case GK_PASS_COL:
    case GK_COL_HALF:
      return xcd_tile_index(wg, s.n_tiles);
    default:
      return wg;
  }
}

long long pass_grid(const DeviceState* ds, const PassDesc& p) {
  if (const int C = split_tile(p, ds->lds_per_block)) return p.n_batch / C;
  if (const int P = pair_tile(p, ds->lds_per_block)) return p.n_batch / (2 * P);
  return geometry(ds, p).grid;
}

int launch_pass(const DeviceState* ds, const PassDesc& p, hipStream_t stream) {
  if (p.n_batch <= 0) return GLSX_OK;
  // a separate output stride exists only in the pair column tiles
  if (p.es_out && p.es_out != p.es && (p.split_mfma || !pair_tile(p, ds->lds_per_block)))
    return fail(GLSX_ERR_STATE, "pass: an output stride of its own needs the pair column tiles");
  const Geometry g = geometry(ds, p);
  if (g.lds_bytes > ds->lds_per_block)
    return fail(GLSX_ERR_UNSUPPORTED, "pass needs " + std::to_string(g.lds_bytes) + " bytes of LDS");
  if (g.grid > 0x7fffffffll) return fail(GLSX_ERR_UNSUPPORTED, "batch too large for one launch");
  PassArgs a{};
  a.src = p.src;
  a.src2 = p.src2;
  a.win = p.win;
  a.dst = p.dst;
  a.tw = ds->twiddles;
  a.argmax = p.argmax;
  a.n_batch = p.n_batch;
  a.es = p.es;
  a.es_out = p.es_out;
  a.bs = p.bs;
  a.out_cols = p.out_cols;
  a.in_kind = p.in_kind;
  a.out_kind = p.out_kind;
  a.w = g.w;
  a.log_w = log2i(g.w);
  a.conj_in = p.conj_in;
  a.conj_out = p.conj_out;
  a.shift = p.shift;
  a.scale = p.scale;
  a.col_shift = p.col ? 1 : 0;
  a.mag_offset = p.mag_offset;
  const bool m = p.split_mfma && p.L >= 16 && p.out_kind == OUT_C32;
  const int pair = m ? mfma_cols_tile(p, ds->lds_per_block) : pair_tile(p, ds->lds_per_block);
  const int split = m ? 0 : split_tile(p, ds->lds_per_block);
  return p.col ? dispatch<true>(p.L, g, a, m, pair, split, stream)
               : dispatch<false>(p.L, g, a, m, pair, split, stream);
}

int validate_fft_shape(int rows, int cols) {
  auto ok = [](int n) { return n >= 1 && n <= kMaxLen && (n & (n - 1)) == 0; };
  if (!ok(rows) || !ok(cols))
    return fail(GLSX_ERR_UNSUPPORTED, "fft size " + std::to_string(rows) + "x" + std::to_string(cols) +
                                          ": rows and cols must be powers of two in [1, 16384]");
  return GLSX_OK;
}

namespace {

// Complex 2D FFT of 8192 rows, its column transforms decimated in time.  The
// 8192-point column pass holds 4 columns per 1024-thread tile (32-byte row
// segments), which reads and writes memory at about half the rate of longer
// segments when the array is larger than the last-level cache.  Split by the
// row index mod R instead: R column passes of M/R points transform the rows
// R*i + q of src into rows q*M/R.. of the stream's scratch, and the row pass
// reads the R rows of each frequency, finishes the column transform with the
// last radix-R stage and transforms the row into dst (fft_row_dit_kernel).
// R = 2: 4096-point halves in the plain 8-column pair tiles.  (R = 4, 2048-
// point quarters in 16-column tiles of 8 pairs, whole 128-byte lines per row
// segment, measured slower at 8192^2: quarters 4 x 83 us against halves 2 x
// 167 us, the 4-row combining pass 304 against about 230 us; step 1.23
// against 1.08 ms.)  The same algorithmic bytes, the same flags, in place or
// not; the column transform's stages are ordered differently, so the results
// match the row-column path within rounding, not bit for bit.
constexpr int kDitRows = 8192;
constexpr int kDitRadix = 2;

bool dit_applies(const void* src, const void* dst, int rows, int cols, int in_kind, bool mfma) {
  if (rows != kDitRows || cols < 2048 || cols > 8192 || in_kind != IN_C32 || mfma) return false;
  return (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0;
}

template <int N>
int launch_row_dit(const DeviceState* ds, const PassArgs& a, int M, hipStream_t stream) {
  auto kern = fft_row_dit_kernel<N, kDitRadix>;
  static LdsLimit lim;
  if (int rc = lim.raise((const void*)kern, ds->lds_per_block - 1024)) return rc;
  if (M % (8 * kDitRadix) != 0) return fail(GLSX_ERR_STATE, "fft (decimated columns): rows");
  const size_t lds = (size_t)Sched<N>::LP * sizeof(float2);
  hipLaunchKernelGGL(kern, dim3((unsigned)M), dim3(Sched<N>::T), lds, stream, a, M);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

// One column pass of the decimation: L-point columns, every R-th row of src
// from row q, into rows q*L.. of the scratch (contiguous).
int launch_dit_columns(DeviceState* ds, const void* src, void* eo, int rows, int cols, int q, bool inv, bool scl,
                       hipStream_t stream) {
  const int L = rows / kDitRadix;
  PassDesc c;
  c.L = L;
  c.col = true;
  c.src = (const char*)src + (size_t)q * cols * sizeof(float2);
  c.dst = (char*)eo + (size_t)q * L * cols * sizeof(float2);
  c.n_batch = cols;
  c.es = (long long)kDitRadix * cols;  // every R-th row of the image
  c.es_out = cols;                     // contiguous rows of the scratch
  c.bs = 1;
  c.in_kind = IN_C32;
  c.out_kind = OUT_C32;
  c.conj_in = inv;
  c.scale = scl ? 1.0f / (float)rows : 1.0f;
  if (!pair_tile(c, ds->lds_per_block)) return fail(GLSX_ERR_STATE, "fft (decimated columns): no pair tile");
  return launch_pass(ds, c, stream);
}

int fft2d_dit(DeviceState* ds, const void* src, void* dst, int rows, int cols, bool inv, bool scl, bool shf,
              hipStream_t stream) {
  const int M = rows;
  // the column transforms in the stream's scratch (the lock keeps another
  // host thread's chain on this stream from using it in between)
  std::unique_lock<std::mutex> lk = stream_lock(ds, stream);
  void* eo = nullptr;
  if (int rc = stream_scratch(ds, stream, 0, (size_t)rows * cols * sizeof(float2), &eo)) return rc;
  for (int q = 0; q < kDitRadix; ++q)
    if (int rc = launch_dit_columns(ds, src, eo, rows, cols, q, inv, scl, stream)) return rc;
  PassArgs a{};
  a.src = eo;
  a.dst = dst;
  a.tw = ds->twiddles;
  a.n_batch = M;
  a.es = 1;
  a.bs = cols;
  a.in_kind = IN_C32;
  a.out_kind = OUT_C32;
  a.w = 1;
  a.conj_out = inv;
  a.shift = shf;
  a.scale = scl ? 1.0f / (float)cols : 1.0f;
  switch (cols) {
    case 2048: return launch_row_dit<2048>(ds, a, M, stream);
    case 4096: return launch_row_dit<4096>(ds, a, M, stream);
    case 8192: return launch_row_dit<8192>(ds, a, M, stream);
    default: return fail(GLSX_ERR_STATE, "fft (decimated columns): row length");
  }
}

// Frames of at most 1024 points, both sides 16..64: fft2d_small_kernel, one
// launch.  Measured per call (back to back / call and synchronise, us): 16^2
// 3.3 / 13.2 against 7.2 / 16.5 for the two passes, 32^2 3.9 / 14.7 against
// 7.2 / 18.0; larger frames are slower in one workgroup than in two launches
// over many CUs (64^2 8.8 / 19.6 against 7.3 / 18.4, 128^2 18.6 / 29.5
// against 7.8 / 18.9).
constexpr int kSmallMaxPoints = 1024;
bool small_applies(int rows, int cols) {
  auto ok = [](int n) { return n == 16 || n == 32 || n == 64; };
  return ok(rows) && ok(cols) && rows * cols <= kSmallMaxPoints;
}

template <int R, int C>
int launch_small_rc(const DeviceState* ds, const SmallArgs& a, hipStream_t stream) {
  auto kern = fft2d_small_kernel<R, C>;
  constexpr size_t lds = (size_t)R * (C + 1) * sizeof(float2);
  static LdsLimit lim;
  if (int rc = lim.raise((const void*)kern, ds->lds_per_block - 1024)) return rc;
  if (lds + 1024 > ds->lds_per_block) return fail(GLSX_ERR_STATE, "fft (small frame): LDS");
  hipLaunchKernelGGL(kern, dim3(1), dim3(R * C / 16), lds, stream, a);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

template <int R>
int launch_small_r(const DeviceState* ds, int cols, const SmallArgs& a, hipStream_t stream) {
  switch (cols) {
    case 16: return launch_small_rc<R, 16>(ds, a, stream);
    case 32: return launch_small_rc<R, 32>(ds, a, stream);
    case 64: return launch_small_rc<R, 64>(ds, a, stream);
    default: return fail(GLSX_ERR_STATE, "fft (small frame): cols");
  }
}

int fft2d_small(const DeviceState* ds, const void* src, int in_kind, void* dst, int rows, int cols, bool inv,
                bool scl, bool shf, hipStream_t stream) {
  SmallArgs a;
  a.src = src;
  a.dst = dst;
  a.tw = ds->twiddles;
  a.in_kind = in_kind;
  a.conj = inv;
  a.shift = shf;
  a.scale_r = scl ? 1.0f / (float)cols : 1.0f;
  a.scale_c = scl ? 1.0f / (float)rows : 1.0f;
  switch (rows) {
    case 16: return launch_small_r<16>(ds, cols, a, stream);
    case 32: return launch_small_r<32>(ds, cols, a, stream);
    case 64: return launch_small_r<64>(ds, cols, a, stream);
    default: return fail(GLSX_ERR_STATE, "fft (small frame): rows");
  }
}

}  // namespace

int fft2d_device(DeviceState* ds, const void* src, int src_type, void* dst, int rows, int cols, int flags,
                 hipStream_t stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "fft: null pointer");
  if (src_type != GLSX_CV_32FC2 && src_type != GLSX_CV_32FC1)
    return fail(GLSX_ERR_ARG, "fft: source must be CV_32FC1 or CV_32FC2");
  if (flags & ~(GLSX_FFT_INVERSE | GLSX_FFT_SCALE | GLSX_FFT_SHIFT | GLSX_FFT_SPLIT_MFMA))
    return fail(GLSX_ERR_ARG, "fft: unknown flag bits");
  if (int rc = validate_fft_shape(rows, cols)) return rc;
  const bool inv = flags & GLSX_FFT_INVERSE, scl = flags & GLSX_FFT_SCALE, shf = flags & GLSX_FFT_SHIFT;
  const int in_kind = src_type == GLSX_CV_32FC1 ? IN_R32 : IN_C32;
  // split-precision passes only for lengths that hold the FP64-checked bound
  // (checked before anything is launched)
  bool mfma_rows = false, mfma_cols = false;
  if (flags & GLSX_FFT_SPLIT_MFMA) {
    if (int rc = split_mfma_use(ds, cols, &mfma_rows)) return rc;
    if (int rc = split_mfma_use(ds, rows, &mfma_cols)) return rc;
  }

  if (src_type == GLSX_CV_32FC1 && rows > 1 && cols > 1) {
    // a real image's rows are read while other rows' spectra (8 bytes per
    // pixel) are written: it may not share memory with the destination
    const char *s0 = (const char*)src, *s1 = s0 + (size_t)rows * cols * sizeof(float);
    const char *d0 = (const char*)dst, *d1 = d0 + (size_t)rows * cols * sizeof(float2);
    if (s0 < d1 && d0 < s1) return fail(GLSX_ERR_ARG, "fft: a CV_32FC1 source must not overlap the destination");
  }
  if (r2c_applies(rows, cols, src_type, flags) && (uintptr_t)dst % 16 == 0) {
    // real image: half-size spectra (20 instead of 28 bytes per pixel; the
    // column pass loads and stores 16-byte column pairs, so a destination
    // aligned to 8 bytes only takes the row and column passes below)
    if (!shf) return fft2d_r2c(ds, (const float*)src, dst, rows, cols, flags, OUT_C32, 0.0f, nullptr, nullptr, nullptr,
                               stream);
    // shifted stores reach columns other tiles still read: the half spectra
    // go to the stream's scratch (allocated on the first such call)
    std::unique_lock<std::mutex> lk = stream_lock(ds, stream);
    void* z = nullptr;
    if (int rc = stream_scratch(ds, stream, 0, (size_t)rows * cols * sizeof(float2), &z)) return rc;
    return fft2d_r2c(ds, (const float*)src, dst, rows, cols, flags, OUT_C32, 0.0f, nullptr, nullptr, z, stream);
  }

  if (rows == 1 || cols == 1) {
    // 1D: one contiguous transform of length max(rows, cols).  Both
    // conjugations are applied (the reference drops one of them for 1D
    // inputs; its 2D inverse is the specified behaviour).
    PassDesc p;
    p.L = rows * cols;
    p.col = false;
    p.src = src;
    p.dst = dst;
    p.n_batch = 1;
    p.es = 1;
    p.bs = p.L;
    p.in_kind = in_kind;
    p.out_kind = OUT_C32;
    p.conj_in = inv;
    p.conj_out = inv;
    p.shift = shf;
    p.scale = scl ? 1.0f / (float)p.L : 1.0f;
    p.split_mfma = rows == 1 ? mfma_rows : mfma_cols;
    return launch_pass(ds, p, stream);
  }
  if (!mfma_rows && !mfma_cols && small_applies(rows, cols))
    return fft2d_small(ds, src, in_kind, dst, rows, cols, inv, scl, shf, stream);
  if (dit_applies(src, dst, rows, cols, in_kind, mfma_rows || mfma_cols))
    return fft2d_dit(ds, src, dst, rows, cols, inv, scl, shf, stream);
  PassDesc r;
  r.L = cols;
  r.col = false;
  r.src = src;
  r.dst = dst;
  r.n_batch = rows;
  r.es = 1;
  r.bs = cols;
  r.in_kind = in_kind;
  r.out_kind = OUT_C32;
  r.conj_in = inv;
  r.shift = shf;
  r.scale = scl ? 1.0f / (float)cols : 1.0f;
  r.split_mfma = mfma_rows;
  if (int rc = launch_pass(ds, r, stream)) return rc;
  PassDesc c;
  c.L = rows;
  c.col = true;
  c.src = dst;
  c.dst = dst;
  c.n_batch = cols;
  c.es = cols;
  c.bs = 1;
  c.in_kind = IN_C32;
  c.out_kind = OUT_C32;
  c.conj_out = inv;
  c.shift = shf;
  c.scale = scl ? 1.0f / (float)rows : 1.0f;
  c.split_mfma = mfma_cols;
  return launch_pass(ds, c, stream);
}

}  // namespace glsx
