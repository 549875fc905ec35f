// taskgraph.hpp -- a device-side executor for DAGs of FFT passes.
//
// A plan is a list of passes (one sweep over one buffer: a row pass, a column
// pass or a phase-correlation finalize) with dependencies, cut into tiles
// ("items").  One persistent kernel runs the whole plan: every workgroup takes
// the next item from a global counter, waits until the passes its pass depends
// on have completed all their items, runs the tile and bumps its pass's
// completion counter.  Items are queued in topological order, so an item only
// ever waits for items already taken by running workgroups, which never wait
// on later ones: the kernel cannot deadlock, and every wait is additionally
// bounded (an exhausted bound sets an error flag instead of hanging).
//
// The reference's equivalent is the host loop that issues one draw per radix
// stage per tile band (glsCV/glsFft.cpp:562-652) and the fixed call sequence
// of glsCV/glsPhaseCorrelate.cpp:178-219; here uneven FFTs and whole
// phase-correlation requests share one launch.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fft_pass.hpp"
#include "glsx_internal.hpp"
#include "poc_final.hpp"

namespace glsx {

enum PassKind { PK_ROW = 0, PK_COL = 1, PK_POC_FINAL = 2 };

struct TgPass {
  PassArgs args;          // FFT passes (w / log_w filled in)
  PocFinalArgs fin;       // PK_POC_FINAL
  int kind;               // PassKind
  int L;                  // transform length (FFT passes)
  int n_items;
  int n_deps;
  int deps[3];
  int pad_;
};

struct TgItem {
  int pass;
  int tile;
};

// Tunable launch shape of the persistent kernel.
struct TgConfig {
  int threads = 256;      // workgroup size
  int col_wcap = 16;      // most columns per column tile
  int mode = 0;           // 0: one launch per dependency level, 1: one persistent launch,
                          // 2: one launch per level and plain kernel (grouped levels)
};

// The kernels of the plain passes (fft_pass.hpp, poc.hip) and the tiling each
// pass gets from them.  Grouped levels run every pass with the kernel and tiles
// the plain path would launch for it, taking (pass, tile) from the plan's
// item table instead of the workgroup index, so each tile has the plain
// kernel's registers, LDS and workgroup size.
enum GroupKernel {
  GK_NONE = 0,
  GK_ROW_WAVE,   // fft_row_wave_kernel: rows that one wave exchanges, a.w rows per workgroup
  GK_ROW,        // fft_row_kernel: one whole row per workgroup (2048 points and more)
  GK_PASS_ROW,   // fft_pass_kernel<L, false>
  GK_PASS_COL,   // fft_pass_kernel<L, true>
  GK_COL_HALF,   // fft_col_half_kernel
  GK_PAIR,       // fft_col_pair_kernel, complex output
  GK_PAIR_MAG,   // fft_col_pair_kernel, magnitudes
  GK_PAIR_REAL,  // fft_col_pair_kernel, real plane with argmax partials
  GK_SPLIT,      // fft_col_split_kernel, complex output
  GK_SPLIT_MAG,  // fft_col_split_kernel, magnitudes
  GK_SPLIT_REAL, // fft_col_split_kernel, real plane with argmax partials
  GK_POCZ,       // pocz_row_kernel: row pairs of a phase correlation
  GK_FINAL,      // poc_finalize_kernel
};
struct PlainShape {
  int kernel = GK_NONE;
  int L = 0;
  int threads = 0;       // workgroup size
  size_t lds = 0;        // dynamic LDS bytes
  long long n_tiles = 0; // workgroups of the plain launch
  int w = 1;             // PassArgs::w (transforms per tile of the one-launch-per-pass kernel)
  int pairs = 0;         // GK_PAIR*: column pairs per tile
};
// The plain launch of pass p (fft_kernels.hip; not for split-precision passes).
int plain_shape(const DeviceState* ds, const PassDesc& p, PlainShape* s);
// The tile the plain launch's workgroup wg computes.
long long plain_tile(const PlainShape& s, long long wg);

// Launches `count` workgroups of the shape's kernel on items base.. of the
// table (group_kernels.hip).
int launch_group(const DeviceState* ds, const PlainShape& s, const TgPass* d_passes, const TgItem* d_items, int base,
                 int count, hipStream_t stream);

// One launch of a grouped level: items start.. start + count - 1, all of one
// plain launch shape.
struct TgGroup {
  PlainShape shape;
  int start = 0, count = 0;
  int level = 0;
};

// Side streams for the groups of one level (grouped levels): the level's
// largest group runs on the caller's stream, the others on side streams that
// wait for the level's start and are waited on before the next level, so the
// small groups run beside the large one.  Under stream capture the forks
// become branches of the graph.  Default-priority, non-blocking streams.
constexpr int kTgForkStreams = 3;
struct TgFork {
  hipStream_t side[kTgForkStreams] = {};
  hipEvent_t start = nullptr;
  hipEvent_t done[kTgForkStreams] = {};
};
int tg_fork_create(TgFork* f);
void tg_fork_destroy(TgFork* f);

// Device state: [0] head, [1] exited workgroups, [2] error flag, [3] spare,
// [4 + p] completed items of pass p.
constexpr int kTgStateHeader = 4;

struct TgPlanHost {
  std::vector<TgPass> passes;
  std::vector<TgItem> items;      // grouped by dependency level
  size_t lds_bytes = 0;
  std::vector<int> level_start;   // first item of each level
  std::vector<int> level_count;
  std::vector<size_t> level_lds;  // LDS bytes the level's tiles need
  std::vector<TgGroup> groups;    // grouped levels: the launches, level by level
};

// Host-side plan building.  The pass's PassArgs must be complete except for
// w / log_w, which are set from the config.
struct TgPassSpec {
  int kind = PK_ROW;
  int L = 0;
  PassArgs args{};
  PocFinalArgs fin{};
  std::vector<int> deps;
};

// Fills w/log_w/n_items of every pass and orders the items: passes are
// emitted in `order` (a topological order).  Returns the LDS bytes needed.
int tg_build(const DeviceState* ds, const std::vector<TgPassSpec>& specs, const TgConfig& cfg, TgPlanHost* out);

// Launches the persistent kernel for a plan already in device memory.
int tg_launch(const DeviceState* ds, const TgPass* d_passes, const TgItem* d_items, int n_items,
              unsigned* d_state, int n_passes, const TgConfig& cfg, size_t lds_bytes, int grid,
              hipStream_t stream);
int tg_grid(const DeviceState* ds, const TgConfig& cfg, size_t lds_bytes, int* grid);

// Level mode: one launch per dependency level, one workgroup per item; the
// kernel boundaries order the levels, so no counters or fences are needed.
int tg_launch_levels(const DeviceState* ds, const TgPass* d_passes, const TgItem* d_items, const TgPlanHost& host,
                     const TgConfig& cfg, hipStream_t stream, const TgFork* fork = nullptr);

// ---- phase-correlation batches (plans.hip) ---------------------------------
size_t align_up(size_t x);

// Per-request scratch inside one workspace: the packed spectrum of the image
// pair, the real correlation surface and the argmax partials; then one result
// per request.
struct PocLayout {
  std::vector<size_t> spec, surface, partials;
  size_t results = 0, total = 0;
};
int poc_layout(const glsx_poc_desc* d, int count, PocLayout* l);
// Smallest workgroup (256, 512 or 1024 threads) the batch's passes fit in.
int poc_min_threads(const glsx_poc_desc* d, int count);
// The five passes of each request (row/col of the packed spectrum, row-pair
// cross-power pass, column pass into the surface with argmax partials,
// finalize).
int poc_specs(const DeviceState* ds, const glsx_poc_desc* d, int count, char* ws, const PocLayout& l,
              std::vector<TgPassSpec>* specs);

}  // namespace glsx
