// group_kernels.hip -- the plain passes' kernels over a plan's item table
// (grouped levels, taskgraph.hpp).
//
// Workgroup b of a group launch runs item base + b of the table: its pass's
// arguments and tile come from there instead of the launch's own arguments
// and workgroup index.  The tile code, workgroup size and LDS are those of the
// one-launch-per-pass kernel the plain path would use for that pass
// (plain_shape), so a grouped plan computes the same bits as the plain path,
// one launch per dependency level and kernel instead of one per pass.  The
// item table lists every pass's tiles in the plain launch's workgroup order
// (plain_tile), so tiles that share 128-byte lines still land on one XCD.
#include <algorithm>

#include "taskgraph.hpp"

namespace glsx {

namespace {

template <int L>
struct GRowWave {
  static constexpr int kThreads = 64;
  static __device__ __forceinline__ void run(const TgPass& p, int tile, float2*, float*, int*) {
    const PassArgs a = p.args;
    row_wave_tile<L>(a, tile);
  }
};

template <int L>
struct GRow {
  static constexpr int kThreads = Sched<L>::T;
  static __device__ __forceinline__ void run(const TgPass& p, int tile, float2* lds, float*, int*) {
    const PassArgs a = p.args;
    row_tile<L>(a, tile, lds);
  }
};

template <int L, bool COL>
struct GPass {
  static constexpr int kThreads = PassLaunch<L, COL>::MAXT;
  static __device__ __forceinline__ void run(const TgPass& p, int tile, float2* lds, float* am_v, int* am_i) {
    const PassArgs a = p.args;
    pass_tile<L, COL>(a, tile, tile, lds, am_v, am_i);
  }
};

template <int L>
struct GColHalf {
  static constexpr int kThreads = 512;
  static __device__ __forceinline__ void run(const TgPass& p, int tile, float2* lds, float* am_v, int* am_i) {
    const PassArgs a = p.args;
    col_half_tile<L>(a, tile, tile, lds, am_v, am_i);
  }
};

template <int L, bool REAL>
struct GPair {
  static constexpr int P = col_pairs(L);
  static constexpr int kThreads = P * Sched<L>::T;
  static __device__ __forceinline__ void run(const TgPass& p, int tile, float2* lds, float* am_v, int* am_i) {
    const PassArgs a = p.args;
    col_pair_tile<L, P, false, REAL>(a, tile, tile, lds, am_v, am_i);
  }
};

template <int L, bool REAL>
struct GSplit {
  static constexpr int kThreads = kSplitCols * Sched<L>::T;
  static __device__ __forceinline__ void run(const TgPass& p, int tile, float2* lds, float* am_v, int* am_i) {
    const PassArgs a = p.args;
    col_split_tile<L, kSplitCols, false, REAL>(a, tile, tile, lds, am_v, am_i);
  }
};

template <int L>
struct GPocz {
  static constexpr int kThreads = pocz_threads(L) > 64 ? pocz_threads(L) : 64;
  static __device__ __forceinline__ void run(const TgPass& p, int tile, float2* lds, float* am_v, int* am_i) {
    const PassArgs a = p.args;
    run_tile_pocz<L>(a, tile, lds, am_v, am_i, threadIdx.x);
  }
};

struct GFinal {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const TgPass& p, int, float2*, float* am_v, int* am_i) {
    const PocFinalArgs f = p.fin;
    poc_finalize_block(f, am_v, am_i);
  }
};

template <class B>
__global__ void __launch_bounds__(B::kThreads) group_kernel(const TgPass* __restrict__ passes,
                                                            const TgItem* __restrict__ items, int base) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  const TgItem item = items[base + (int)blockIdx.x];
  const int pi = __builtin_amdgcn_readfirstlane(item.pass);
  const int tile = __builtin_amdgcn_readfirstlane(item.tile);
  B::run(passes[pi], tile, lds, am_v, am_i);
}

template <class B>
int launch_b(const DeviceState* ds, const PlainShape& s, const TgPass* d_passes, const TgItem* d_items, int base,
             int count, hipStream_t stream) {
  if (s.threads < 1 || s.threads > B::kThreads)
    return fail(GLSX_ERR_STATE, "grouped level: workgroup size beyond the kernel's bound");
  auto kern = group_kernel<B>;
  static LdsLimit lim;
  if (s.lds > 0)
    if (int rc = lim.raise((const void*)kern, ds->lds_per_block - 1024)) return rc;
  hipLaunchKernelGGL(kern, dim3((unsigned)count), dim3(s.threads), s.lds, stream, d_passes, d_items, base);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

// Calls f.template operator()<L>() for the power of two L (1..4096).
template <class F>
int by_length(int L, F&& f) {
  switch (L) {
    case 1: return f.template operator()<1>();
    case 2: return f.template operator()<2>();
    case 4: return f.template operator()<4>();
    case 8: return f.template operator()<8>();
    case 16: return f.template operator()<16>();
    case 32: return f.template operator()<32>();
    case 64: return f.template operator()<64>();
    case 128: return f.template operator()<128>();
    case 256: return f.template operator()<256>();
    case 512: return f.template operator()<512>();
    case 1024: return f.template operator()<1024>();
    case 2048: return f.template operator()<2048>();
    case 4096: return f.template operator()<4096>();
    default: return fail(GLSX_ERR_UNSUPPORTED, "grouped level: length " + std::to_string(L));
  }
}

}  // namespace

int launch_group(const DeviceState* ds, const PlainShape& s, const TgPass* d_passes, const TgItem* d_items, int base,
                 int count, hipStream_t stream) {
  if (count <= 0) return GLSX_OK;
  auto bad = [&]() { return fail(GLSX_ERR_UNSUPPORTED, "grouped level: no such kernel for length " +
                                                           std::to_string(s.L)); };
  auto go = [&]<class B>() { return launch_b<B>(ds, s, d_passes, d_items, base, count, stream); };
  switch (s.kernel) {
    case GK_ROW_WAVE:
      return by_length(s.L, [&]<int L>() {
        if constexpr (wave_rows(L)) return go.template operator()<GRowWave<L>>();
        else return bad();
      });
    case GK_ROW:
      return by_length(s.L, [&]<int L>() {
        if constexpr (L >= 2048) return go.template operator()<GRow<L>>();
        else return bad();
      });
    case GK_PASS_ROW:
      return by_length(s.L, [&]<int L>() { return go.template operator()<GPass<L, false>>(); });
    case GK_PASS_COL:
      return by_length(s.L, [&]<int L>() { return go.template operator()<GPass<L, true>>(); });
    case GK_COL_HALF:
      return by_length(s.L, [&]<int L>() {
        if constexpr (L == 4096) return go.template operator()<GColHalf<L>>();
        else return bad();
      });
    case GK_PAIR:
      return by_length(s.L, [&]<int L>() {
        if constexpr (L == 1024 || L == 2048 || L == 4096) return go.template operator()<GPair<L, false>>();
        else return bad();
      });
    case GK_PAIR_REAL:
      return by_length(s.L, [&]<int L>() {
        if constexpr (L == 1024 || L == 2048) return go.template operator()<GPair<L, true>>();
        else return bad();
      });
    case GK_SPLIT:
      return by_length(s.L, [&]<int L>() {
        if constexpr (L == 1024 || L == 2048) return go.template operator()<GSplit<L, false>>();
        else return bad();
      });
    case GK_SPLIT_REAL:
      return by_length(s.L, [&]<int L>() {
        if constexpr (L == 1024) return go.template operator()<GSplit<L, true>>();
        else return bad();
      });
    case GK_POCZ:
      return by_length(s.L, [&]<int L>() {
        if constexpr (L >= 2) return go.template operator()<GPocz<L>>();
        else return bad();
      });
    case GK_FINAL:
      return go.template operator()<GFinal>();
    default:
      return bad();
  }
}

}  // namespace glsx
