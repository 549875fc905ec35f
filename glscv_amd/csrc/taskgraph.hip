// taskgraph.hip -- the persistent task-graph kernel and its host-side builder.
#include <algorithm>
#include <string>

#include "taskgraph.hpp"

namespace glsx {

namespace {

constexpr unsigned kSpinLimit = 1u << 26;  // ~ seconds of s_sleep: far above any legitimate wait

template <bool COL>
__device__ __forceinline__ void tile_dispatch(int L, const PassArgs& a, long long tile, long long slot, float2* lds,
                                              float* am_v, int* am_i) {
  // A fresh copy of the lane index per item: without it the compiler hoists
  // every length's lane-dependent address arithmetic out of the item loop
  // and keeps it all live, which spills.
  int tid;
  asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
  if constexpr (!COL) {
    if (a.in_kind == IN_POCZ) {
      switch (L) {
        case 2: run_tile_pocz<2>(a, tile, lds, am_v, am_i, tid); break;
        case 4: run_tile_pocz<4>(a, tile, lds, am_v, am_i, tid); break;
        case 8: run_tile_pocz<8>(a, tile, lds, am_v, am_i, tid); break;
        case 16: run_tile_pocz<16>(a, tile, lds, am_v, am_i, tid); break;
        case 32: run_tile_pocz<32>(a, tile, lds, am_v, am_i, tid); break;
        case 64: run_tile_pocz<64>(a, tile, lds, am_v, am_i, tid); break;
        case 128: run_tile_pocz<128>(a, tile, lds, am_v, am_i, tid); break;
        case 256: run_tile_pocz<256>(a, tile, lds, am_v, am_i, tid); break;
        case 512: run_tile_pocz<512>(a, tile, lds, am_v, am_i, tid); break;
        case 1024: run_tile_pocz<1024>(a, tile, lds, am_v, am_i, tid); break;
        case 2048: run_tile_pocz<2048>(a, tile, lds, am_v, am_i, tid); break;
        case 4096: run_tile_pocz<4096>(a, tile, lds, am_v, am_i, tid); break;
        default: break;  // rejected on the host
      }
      return;
    }
  }
  switch (L) {
    case 1: run_tile<1, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 2: run_tile<2, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 4: run_tile<4, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 8: run_tile<8, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 16: run_tile<16, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 32: run_tile<32, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 64: run_tile<64, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 128: run_tile<128, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 256: run_tile<256, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 512: run_tile<512, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 1024: run_tile<1024, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 2048: run_tile<2048, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    case 4096: run_tile<4096, COL>(a, tile, slot, lds, am_v, am_i, tid); break;
    default: break;  // rejected on the host
  }
}

__global__ void __launch_bounds__(1024) taskgraph_kernel(const TgPass* __restrict__ passes,
                                                         const TgItem* __restrict__ items, int n_items,
                                                         unsigned* __restrict__ state, int n_passes) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  __shared__ int s_item;
  unsigned* head = state;
  unsigned* exited = state + 1;
  unsigned* err = state + 2;
  unsigned* done = state + kTgStateHeader;

  for (;;) {
    __syncthreads();  // the previous item is finished with LDS and s_item
    if (threadIdx.x == 0) s_item = (int)__hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) break;
    const TgItem item = items[it];
    const int pi = __builtin_amdgcn_readfirstlane(item.pass);
    const int tile = __builtin_amdgcn_readfirstlane(item.tile);
    const TgPass* p = passes + pi;
    const int kind = p->kind;

    // wait until the passes this one reads have finished all their items
    if (threadIdx.x == 0) {
      for (int d = 0; d < p->n_deps; ++d) {
        const int dp = p->deps[d];
        const unsigned need = (unsigned)passes[dp].n_items;
        unsigned spins = 0;
        while (__hip_atomic_load(&done[dp], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
          __builtin_amdgcn_s_sleep(4);
          if (++spins > kSpinLimit) {
            __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

    if (kind == PK_ROW) {
      const PassArgs a = p->args;
      tile_dispatch<false>(p->L, a, tile, tile, lds, am_v, am_i);
    } else if (kind == PK_COL) {
      const PassArgs a = p->args;
      tile_dispatch<true>(p->L, a, tile, tile, lds, am_v, am_i);
    } else {
      const PocFinalArgs f = p->fin;
      poc_finalize_block(f, am_v, am_i);
    }

    // publish this item's stores, then count it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(&done[pi], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }

  // The last workgroup out resets the counters, so the same plan (or a
  // replayed graph of it) starts from zero next time.
  if (threadIdx.x == 0) {
    const unsigned prev = __hip_atomic_fetch_add(exited, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == gridDim.x - 1) {
      for (int i = 0; i < n_passes; ++i) __hip_atomic_store(&done[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(head, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(exited, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Level mode: workgroup b runs item item_base + b of one dependency level.
__global__ void __launch_bounds__(1024) level_kernel(const TgPass* __restrict__ passes,
                                                     const TgItem* __restrict__ items, int item_base) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  const TgItem item = items[item_base + (int)blockIdx.x];
  const int pi = __builtin_amdgcn_readfirstlane(item.pass);
  const int tile = __builtin_amdgcn_readfirstlane(item.tile);
  const TgPass* p = passes + pi;
  const int kind = p->kind;
  if (kind == PK_ROW) {
    const PassArgs a = p->args;
    tile_dispatch<false>(p->L, a, tile, tile, lds, am_v, am_i);
  } else if (kind == PK_COL) {
    const PassArgs a = p->args;
    tile_dispatch<true>(p->L, a, tile, tile, lds, am_v, am_i);
  } else {
    const PocFinalArgs f = p->fin;
    poc_finalize_block(f, am_v, am_i);
  }
}

int sched_T(int L) { return L < 16 ? 1 : L / 16; }
bool needs_lds(int L) { return L >= 32; }
int ilog2ll(long long x) {
  int r = 0;
  while ((1ll << r) < x) ++r;
  return r;
}

}  // namespace

int tg_build(const DeviceState* ds, const std::vector<TgPassSpec>& specs, const TgConfig& cfg, TgPlanHost* out) {
  const int NT = cfg.threads;
  if (NT != 256 && NT != 512 && NT != 1024) return fail(GLSX_ERR_ARG, "task graph: threads must be 256/512/1024");
  // column tiles map lanes with tid & (w - 1): the cap must be a power of two
  if (cfg.col_wcap < 1 || cfg.col_wcap > 64 || (cfg.col_wcap & (cfg.col_wcap - 1)))
    return fail(GLSX_ERR_ARG, "task graph: col_wcap must be a power of two in [1, 64]");
  if (cfg.mode < 0 || cfg.mode > 2)
    return fail(GLSX_ERR_ARG, "task graph: mode must be 0 (levels), 1 (persistent) or 2 (grouped levels)");
  const size_t lds_budget = ds->lds_per_block > 2048 ? ds->lds_per_block - 2048 : 0;
  out->passes.clear();
  out->items.clear();
  out->lds_bytes = 0;
  out->level_start.clear();
  out->level_count.clear();
  out->level_lds.clear();
  out->groups.clear();
  const bool grouped = cfg.mode == 2;
  std::vector<PlainShape> shape(specs.size());
  std::vector<int> level(specs.size(), 0);
  for (size_t i = 0; i < specs.size(); ++i) {
    const TgPassSpec& s = specs[i];
    TgPass p{};
    p.kind = s.kind;
    p.L = s.L;
    p.args = s.args;
    p.fin = s.fin;
    if ((int)s.deps.size() > 3) return fail(GLSX_ERR_ARG, "task graph: at most 3 dependencies per pass");
    p.n_deps = (int)s.deps.size();
    for (int d = 0; d < p.n_deps; ++d) {
      if (s.deps[d] < 0 || s.deps[d] >= (int)i) return fail(GLSX_ERR_ARG, "task graph: dependency must precede");
      p.deps[d] = s.deps[d];
      level[i] = std::max(level[i], level[s.deps[d]] + 1);
    }
    if (s.kind == PK_POC_FINAL) {
      p.n_items = 1;
      shape[i].kernel = GK_FINAL;
      shape[i].threads = 256;
      shape[i].n_tiles = 1;
    } else if (grouped) {
      // the plain path's kernel and tiles for this pass
      if (s.L < 1 || s.L > 4096 || (s.L & (s.L - 1)))
        return fail(GLSX_ERR_UNSUPPORTED, "task graph: length must be 1..4096");
      PassDesc d;
      d.L = s.L;
      d.col = s.kind == PK_COL;
      d.src = s.args.src;
      d.src2 = s.args.src2;
      d.win = s.args.win;
      d.dst = s.args.dst;
      d.argmax = s.args.argmax;
      d.n_batch = s.args.n_batch;
      d.es = s.args.es;
      d.bs = s.args.bs;
      d.out_cols = s.args.out_cols;
      d.in_kind = s.args.in_kind;
      d.out_kind = s.args.out_kind;
      if (int rc = plain_shape(ds, d, &shape[i])) return rc;
      p.args.w = shape[i].w;
      p.args.log_w = ilog2ll(shape[i].w);
      p.args.col_shift = s.kind == PK_COL ? 1 : 0;
      p.n_items = (int)shape[i].n_tiles;
      out->lds_bytes = std::max(out->lds_bytes, shape[i].lds);
    } else {
      const int L = s.L;
      if (L < 1 || L > 4096 || (L & (L - 1))) return fail(GLSX_ERR_UNSUPPORTED, "task graph: length must be 1..4096");
      const int T = sched_T(L);
      if (T > NT) return fail(GLSX_ERR_UNSUPPORTED, "task graph: workgroup too small for length");
      const long long nb = s.args.n_batch;
      int w;
      size_t lds = 0;
      if (s.kind == PK_ROW && s.args.in_kind == IN_POCZ) {
        // row-pair tiles (run_tile_pocz): two rows, L/16 threads
        if (nb < 2 || (nb & (nb - 1))) return fail(GLSX_ERR_UNSUPPORTED, "task graph: POC rows must be a power of two");
        if (pocz_threads(L) > NT) return fail(GLSX_ERR_UNSUPPORTED, "task graph: workgroup too small for POC rows");
        w = 2;
        lds = pocz_lds_bytes(L);
      } else if (s.kind == PK_ROW) {
        w = std::max(1, NT / T);
        while (w > 1 && w / 2 >= nb) w >>= 1;
        const size_t per = needs_lds(L) ? (size_t)(L + L / 16 + 1) * sizeof(float2) : 0;
        while (w > 1 && per * w > lds_budget) w >>= 1;
        lds = per * w;
      } else {
        w = std::min(cfg.col_wcap, NT / T);
        w = std::max(w, 1);
        const size_t per = needs_lds(L) ? (size_t)(L + L / 16) * sizeof(float2) : 0;
        while (w > 1 && per * w > lds_budget) w >>= 1;
        while (w > 1 && w / 2 >= nb) w >>= 1;
        lds = per * w;
      }
      if (lds > lds_budget) return fail(GLSX_ERR_UNSUPPORTED, "task graph: tile does not fit in LDS");
      p.args.w = w;
      p.args.log_w = ilog2ll(w);
      p.args.col_shift = s.kind == PK_COL ? 1 : 0;
      p.n_items = s.args.in_kind == IN_POCZ && s.kind == PK_ROW ? (int)(nb / 2 + 1) : (int)((nb + w - 1) / w);
      out->lds_bytes = std::max(out->lds_bytes, lds);
    }
    out->passes.push_back(p);
  }
  // items in topological order: by dependency level, then by pass order
  std::vector<int> order(specs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return level[a] < level[b]; });
  if (grouped) {
    // level by level, one group per plain launch shape (kernel, length,
    // workgroup size, LDS: the wave-exchange rows assume exactly a.w rows per
    // workgroup) in order of first use; within a group, passes whose tile
    // count is a multiple of 16 first, so every such pass starts at a
    // workgroup index that is a multiple of 16 and keeps the plain launch's
    // tile-to-XCD order
    auto key = [&](int pi) {
      const PlainShape& sh = shape[pi];
      return ((unsigned long long)sh.kernel << 56) | ((unsigned long long)sh.L << 40) |
             ((unsigned long long)sh.threads << 24) | (unsigned long long)sh.lds;
    };
    for (size_t a = 0; a < order.size();) {
      size_t b = a;
      while (b < order.size() && level[order[b]] == level[order[a]]) ++b;
      std::vector<int> lv(order.begin() + a, order.begin() + b);
      std::vector<unsigned long long> keys;
      for (int pi : lv)
        if (std::find(keys.begin(), keys.end(), key(pi)) == keys.end()) keys.push_back(key(pi));
      out->level_start.push_back((int)out->items.size());
      out->level_count.push_back(0);
      out->level_lds.push_back(0);
      for (unsigned long long k : keys) {
        std::vector<int> ps;
        for (int pi : lv)
          if (key(pi) == k) ps.push_back(pi);
        std::stable_partition(ps.begin(), ps.end(), [&](int pi) { return out->passes[pi].n_items % 16 == 0; });
        TgGroup g;
        g.shape = shape[ps[0]];
        g.level = (int)out->level_start.size() - 1;
        g.start = (int)out->items.size();
        for (int pi : ps)
          for (int q = 0; q < out->passes[pi].n_items; ++q)
            out->items.push_back(TgItem{pi, (int)plain_tile(shape[pi], q)});
        g.count = (int)out->items.size() - g.start;
        out->level_count.back() += g.count;
        out->level_lds.back() = std::max(out->level_lds.back(), g.shape.lds);
        out->groups.push_back(g);
      }
      a = b;
    }
    return GLSX_OK;
  }
  int cur_level = -1;
  for (int pi : order) {
    if (level[pi] != cur_level) {
      cur_level = level[pi];
      out->level_start.push_back((int)out->items.size());
      out->level_count.push_back(0);
      out->level_lds.push_back(0);
    }
    const TgPass& ps = out->passes[pi];
    const int n = ps.n_items;
    // Column tiles narrower than a 128-byte line: give the tiles that share
    // lines the same workgroup index mod 8, i.e. the same XCD (workgroups are
    // dispatched round-robin over the 8 XCDs), as the one-pass kernel does.
    const bool xcd = ps.kind == PK_COL && ps.args.w < 16 && (n % 8) == 0;
    for (int q = 0; q < n; ++q) {
      const int t = xcd ? (q % 8) * (n / 8) + q / 8 : q;
      out->items.push_back(TgItem{pi, t});
    }
    out->level_count.back() += n;
    size_t lds = 0;
    if (ps.kind == PK_ROW && ps.args.in_kind == IN_POCZ) {
      lds = pocz_lds_bytes(ps.L);
    } else if (ps.kind != PK_POC_FINAL && needs_lds(ps.L)) {
      const int L = ps.L;
      lds = ps.kind == PK_ROW ? (size_t)ps.args.w * (L + L / 16 + 1) * sizeof(float2)
                              : (size_t)ps.args.w * (L + L / 16) * sizeof(float2);
    }
    out->level_lds.back() = std::max(out->level_lds.back(), lds);
  }
  return GLSX_OK;
}

int tg_fork_create(TgFork* f) {
  *f = TgFork{};
  GLSX_HIP(hipEventCreateWithFlags(&f->start, hipEventDisableTiming));
  for (int i = 0; i < kTgForkStreams; ++i) {
    GLSX_HIP(hipStreamCreateWithFlags(&f->side[i], hipStreamNonBlocking));
    GLSX_HIP(hipEventCreateWithFlags(&f->done[i], hipEventDisableTiming));
  }
  return GLSX_OK;
}

// (The owner has synchronised the device: nothing is in flight on the side
// streams; destruction errors are not actionable here.)
void tg_fork_destroy(TgFork* f) {
  for (int i = 0; i < kTgForkStreams; ++i) {
    if (f->side[i]) (void)hipStreamDestroy(f->side[i]);
    if (f->done[i]) (void)hipEventDestroy(f->done[i]);
  }
  if (f->start) (void)hipEventDestroy(f->start);
  *f = TgFork{};
}

int tg_launch_levels(const DeviceState* ds, const TgPass* d_passes, const TgItem* d_items, const TgPlanHost& host,
                     const TgConfig& cfg, hipStream_t stream, const TgFork* fork) {
  if (cfg.mode == 2) {
    // grouped levels: one launch per (level, launch shape); the launches of
    // a level are independent, the level's end orders the next one
    const auto& gs = host.groups;
    for (size_t a = 0; a < gs.size();) {
      size_t b = a;
      while (b < gs.size() && gs[b].level == gs[a].level) ++b;
      if (!fork || !fork->start || b - a == 1) {
        for (size_t i = a; i < b; ++i)
          if (int rc = launch_group(ds, gs[i].shape, d_passes, d_items, gs[i].start, gs[i].count, stream)) return rc;
      } else {
        // the largest group on the caller's stream, the others round robin
        // on the side streams
        size_t big = a;
        for (size_t i = a; i < b; ++i)
          if (gs[i].count * (long long)gs[i].shape.threads > gs[big].count * (long long)gs[big].shape.threads) big = i;
        const int nside = (int)std::min<size_t>(b - a - 1, kTgForkStreams);
        GLSX_HIP(hipEventRecord(fork->start, stream));
        for (int k = 0; k < nside; ++k) GLSX_HIP(hipStreamWaitEvent(fork->side[k], fork->start, 0));
        int k = 0;
        for (size_t i = a; i < b; ++i) {
          hipStream_t s = i == big ? stream : fork->side[k++ % nside];
          if (int rc = launch_group(ds, gs[i].shape, d_passes, d_items, gs[i].start, gs[i].count, s)) return rc;
        }
        for (int j = 0; j < nside; ++j) {
          GLSX_HIP(hipEventRecord(fork->done[j], fork->side[j]));
          GLSX_HIP(hipStreamWaitEvent(stream, fork->done[j], 0));
        }
      }
      a = b;
    }
    return GLSX_OK;
  }
  static LdsLimit lim;
  if (int rc = lim.raise((const void*)level_kernel, ds->lds_per_block - 1024)) return rc;
  for (size_t l = 0; l < host.level_start.size(); ++l) {
    if (host.level_count[l] <= 0) continue;
    hipLaunchKernelGGL(level_kernel, dim3(host.level_count[l]), dim3(cfg.threads), host.level_lds[l], stream,
                       d_passes, d_items, host.level_start[l]);
    GLSX_HIP(hipGetLastError());
  }
  return GLSX_OK;
}

int tg_grid(const DeviceState* ds, const TgConfig& cfg, size_t lds_bytes, int* grid) {
  int occ = 0;
  GLSX_HIP(hipFuncSetAttribute((const void*)taskgraph_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(ds->lds_per_block - 1024)));
  GLSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, taskgraph_kernel, cfg.threads, lds_bytes));
  if (occ < 1) return fail(GLSX_ERR_UNSUPPORTED, "task graph: kernel does not fit on a CU");
  *grid = occ * ds->num_cu;
  return GLSX_OK;
}

int tg_launch(const DeviceState* ds, const TgPass* d_passes, const TgItem* d_items, int n_items, unsigned* d_state,
              int n_passes, const TgConfig& cfg, size_t lds_bytes, int grid, hipStream_t stream) {
  (void)ds;
  if (n_items <= 0) return GLSX_OK;
  hipLaunchKernelGGL(taskgraph_kernel, dim3(grid), dim3(cfg.threads), lds_bytes, stream, d_passes, d_items, n_items,
                     d_state, n_passes);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

}  // namespace glsx
