// plans.hip -- batched FFT and batched phase-correlation plans over the
// persistent task-graph kernel, with hipGraph capture and an autotuner whose
// choices persist in a small text database.
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "taskgraph.hpp"

using namespace glsx;

struct glsx_plan_s {
  int device = 0;
  int kind = 0;  // 0 FFT batch, 1 phase-correlation batch
  TgConfig cfg;
  TgPlanHost host;
  TgPass* d_passes = nullptr;
  TgItem* d_items = nullptr;
  unsigned* d_state = nullptr;
  int grid = 0;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture_stream = nullptr;
  TgFork fork;  // grouped levels: side streams for a level's groups
  glsx_poc_result* d_results = nullptr;
  int n_results = 0;
  int tuned = 0;
  double tuned_us = -1.0;
};

namespace {

// ---------------------------------------------------------------------------
// tuning database: lines "key threads col_wcap mode microseconds"
// ---------------------------------------------------------------------------
std::mutex g_db_mutex;

std::string db_path() {
  if (const char* e = std::getenv("GLSX_TUNE_DB")) return e;
  Dl_info info;
  if (dladdr((void*)&db_path, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t slash = p.rfind('/');
    return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/glsx_tune.db";
  }
  return "glsx_tune.db";
}

bool db_lookup(const std::string& key, TgConfig* cfg, double* us) {
  std::lock_guard<std::mutex> lk(g_db_mutex);
  std::ifstream f(db_path());
  std::string line;
  bool found = false;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string k;
    int t = 0, w = 0, m = 0;
    double u = 0;
    if (!(is >> k >> t >> w >> m >> u) || k != key) continue;
    // entries that no launch shape accepts (a damaged or hand-edited file) are skipped
    const bool valid = (t == 256 || t == 512 || t == 1024) && w >= 1 && w <= 64 && !(w & (w - 1)) &&
                       (m == 0 || m == 1);
    if (valid) {
      cfg->threads = t;
      cfg->col_wcap = w;
      cfg->mode = m;
      *us = u;
      found = true;  // the last entry for a key wins
    }
  }
  return found;
}

void db_store(const std::string& key, const TgConfig& cfg, double us) {
  std::lock_guard<std::mutex> lk(g_db_mutex);
  std::ofstream f(db_path(), std::ios::app);
  if (f) f << key << " " << cfg.threads << " " << cfg.col_wcap << " " << cfg.mode << " " << us << "\n";
}

std::string arch_name(int dev) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
  std::string a = p.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

// ---------------------------------------------------------------------------
// pass specs
// ---------------------------------------------------------------------------
PassArgs base_args(const DeviceState* ds) {
  PassArgs a{};
  a.tw = ds->twiddles;
  a.scale = 1.0f;
  return a;
}

int fft_specs(const DeviceState* ds, const glsx_fft_desc* d, int count, std::vector<TgPassSpec>* specs) {
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  // largest transforms first, so the long tiles start early
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (long long)d[a].rows * d[a].cols > (long long)d[b].rows * d[b].cols;
  });
  for (int i : order) {
    const glsx_fft_desc& f = d[i];
    if (!f.src || !f.dst) return fail(GLSX_ERR_ARG, "fft batch: null pointer in descriptor " + std::to_string(i));
    if (f.src_type != GLSX_CV_32FC1 && f.src_type != GLSX_CV_32FC2)
      return fail(GLSX_ERR_ARG, "fft batch: descriptor " + std::to_string(i) + " has a bad source type");
    if (f.flags & ~(GLSX_FFT_INVERSE | GLSX_FFT_SCALE | GLSX_FFT_SHIFT))
      return fail(GLSX_ERR_ARG, "fft batch: unknown flag bits");
    auto ok = [](int n) { return n >= 1 && n <= 4096 && (n & (n - 1)) == 0; };
    if (!ok(f.rows) || !ok(f.cols))
      return fail(GLSX_ERR_UNSUPPORTED, "fft batch: rows and cols must be powers of two in [1, 4096]");
    const bool inv = f.flags & GLSX_FFT_INVERSE, scl = f.flags & GLSX_FFT_SCALE, shf = f.flags & GLSX_FFT_SHIFT;
    const int in_kind = f.src_type == GLSX_CV_32FC1 ? IN_R32 : IN_C32;
    if (f.rows == 1 || f.cols == 1) {
      TgPassSpec s;
      s.kind = PK_ROW;
      s.L = f.rows * f.cols;
      s.args = base_args(ds);
      s.args.src = f.src;
      s.args.dst = f.dst;
      s.args.n_batch = 1;
      s.args.es = 1;
      s.args.bs = s.L;
      s.args.in_kind = in_kind;
      s.args.conj_in = inv;
      s.args.conj_out = inv;
      s.args.shift = shf;
      s.args.scale = scl ? 1.0f / (float)s.L : 1.0f;
      specs->push_back(s);
      continue;
    }
    TgPassSpec r;
    r.kind = PK_ROW;
    r.L = f.cols;
    r.args = base_args(ds);
    r.args.src = f.src;
    r.args.dst = f.dst;
    r.args.n_batch = f.rows;
    r.args.es = 1;
    r.args.bs = f.cols;
    r.args.in_kind = in_kind;
    r.args.conj_in = inv;
    r.args.shift = shf;
    r.args.scale = scl ? 1.0f / (float)f.cols : 1.0f;
    const int ri = (int)specs->size();
    specs->push_back(r);
    TgPassSpec c;
    c.kind = PK_COL;
    c.L = f.rows;
    c.args = base_args(ds);
    c.args.src = f.dst;
    c.args.dst = f.dst;
    c.args.n_batch = f.cols;
    c.args.es = f.cols;
    c.args.bs = 1;
    c.args.in_kind = IN_C32;
    c.args.conj_out = inv;
    c.args.shift = shf;
    c.args.scale = scl ? 1.0f / (float)f.rows : 1.0f;
    c.deps = {ri};
    specs->push_back(c);
  }
  return GLSX_OK;
}

}  // namespace

size_t glsx::align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int glsx::poc_layout(const glsx_poc_desc* d, int count, PocLayout* l) {
  size_t off = 0;
  for (int i = 0; i < count; ++i) {
    auto ok = [](int n) { return n >= 2 && n <= 4096 && (n & (n - 1)) == 0; };
    if (!ok(d[i].rows) || !ok(d[i].cols))
      return fail(GLSX_ERR_UNSUPPORTED, "poc batch: rows and cols must be powers of two in [2, 4096]");
    const size_t n = (size_t)d[i].rows * d[i].cols;
    l->spec.push_back(off);
    off = align_up(off + n * sizeof(float2));
    l->surface.push_back(off);
    off = align_up(off + n * sizeof(float));
    l->partials.push_back(off);
    off = align_up(off + (size_t)d[i].cols * sizeof(int2));
  }
  l->results = off;
  off = align_up(off + (size_t)count * sizeof(glsx_poc_result));
  l->total = off;
  return GLSX_OK;
}

int glsx::poc_min_threads(const glsx_poc_desc* d, int count) {
  int t = 256;
  for (int i = 0; i < count; ++i) t = std::max(t, pocz_threads(d[i].cols));
  return t;
}

int glsx::poc_specs(const DeviceState* ds, const glsx_poc_desc* d, int count, char* ws, const PocLayout& l,
              std::vector<TgPassSpec>* specs) {
  auto base_args_of = [](const DeviceState* ds) {
    PassArgs a{};
    a.tw = ds->twiddles;
    a.scale = 1.0f;
    return a;
  };
  for (int i = 0; i < count; ++i) {
    const glsx_poc_desc& q = d[i];
    if (!q.src1 || !q.src2) return fail(GLSX_ERR_ARG, "poc batch: null image in descriptor " + std::to_string(i));
    const int m = q.rows, n = q.cols;
    float2* Z = (float2*)(ws + l.spec[i]);
    float* S = (float*)(ws + l.surface[i]);
    int2* P = (int2*)(ws + l.partials[i]);
    glsx_poc_result* R = (glsx_poc_result*)(ws + l.results) + i;
    const int base = (int)specs->size();
    // the chain of poc.hip: Z = FFT2(a*w + i b*w), then the row-pair
    // cross-power pass, the column pass into the surface, the finalize
    TgPassSpec rz;
    rz.kind = PK_ROW;
    rz.L = n;
    rz.args = base_args_of(ds);
    rz.args.src = q.src1;
    rz.args.src2 = q.src2;
    rz.args.win = q.window;
    rz.args.in_kind = IN_R32PAIR;
    rz.args.dst = Z;
    rz.args.n_batch = m;
    rz.args.es = 1;
    rz.args.bs = n;
    TgPassSpec cz;
    cz.kind = PK_COL;
    cz.L = m;
    cz.args = base_args_of(ds);
    cz.args.src = Z;
    cz.args.dst = Z;
    cz.args.n_batch = n;
    cz.args.es = n;
    cz.args.bs = 1;
    cz.deps = {base + 0};
    TgPassSpec ri;
    ri.kind = PK_ROW;
    ri.L = n;
    ri.args = base_args_of(ds);
    ri.args.in_kind = IN_POCZ;
    ri.args.src = Z;
    ri.args.dst = Z;
    ri.args.n_batch = m;
    ri.args.es = 1;
    ri.args.bs = n;
    ri.args.conj_in = 1;
    ri.args.shift = 1;
    ri.deps = {base + 1};
    TgPassSpec ci;
    ci.kind = PK_COL;
    ci.L = m;
    ci.args = base_args_of(ds);
    ci.args.src = Z;
    ci.args.dst = S;
    ci.args.out_kind = OUT_R32;
    ci.args.conj_out = 1;
    ci.args.shift = 1;
    ci.args.argmax = P;
    ci.args.out_cols = n;
    ci.args.n_batch = n;
    ci.args.es = n;
    ci.args.bs = 1;
    ci.deps = {base + 2};
    TgPassSpec fin;
    fin.kind = PK_POC_FINAL;
    fin.fin = PocFinalArgs{P, 0, S, m, n, R};
    fin.deps = {base + 3};
    for (auto* s : {&rz, &cz, &ri, &ci, &fin}) specs->push_back(*s);
  }
  return GLSX_OK;
}

namespace {

void free_device(glsx_plan p) {
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  if (p->capture_stream) hipStreamDestroy(p->capture_stream);
  tg_fork_destroy(&p->fork);
  if (p->d_passes) hipFree(p->d_passes);
  if (p->d_items) hipFree(p->d_items);
  if (p->d_state) hipFree(p->d_state);
  p->exec = nullptr;
  p->graph = nullptr;
  p->capture_stream = nullptr;
  p->d_passes = nullptr;
  p->d_items = nullptr;
  p->d_state = nullptr;
}

// Build host tables for cfg, upload them and (optionally) capture the graph.
int realize(const DeviceState* ds, glsx_plan p, const std::vector<TgPassSpec>& specs) {
  free_device(p);
  if (int rc = tg_build(ds, specs, p->cfg, &p->host)) return rc;
  // the finalize of a phase correlation reduces as many partials as its
  // column pass has tiles
  for (auto& ps : p->host.passes)
    if (ps.kind == PK_POC_FINAL) ps.fin.n_partials = p->host.passes[ps.deps[0]].n_items;
  p->grid = 0;
  if (p->cfg.mode == 1) {
    if (int rc = tg_grid(ds, p->cfg, p->host.lds_bytes, &p->grid)) return rc;
    p->grid = std::max(1, std::min<int>(p->grid, (int)p->host.items.size()));
  }
  const size_t np = p->host.passes.size(), ni = p->host.items.size();
  GLSX_HIP(hipMalloc(&p->d_passes, sizeof(TgPass) * np));
  GLSX_HIP(hipMalloc(&p->d_items, sizeof(TgItem) * std::max<size_t>(ni, 1)));
  GLSX_HIP(hipMalloc(&p->d_state, sizeof(unsigned) * (kTgStateHeader + np)));
  GLSX_HIP(hipMemcpy(p->d_passes, p->host.passes.data(), sizeof(TgPass) * np, hipMemcpyHostToDevice));
  if (ni) GLSX_HIP(hipMemcpy(p->d_items, p->host.items.data(), sizeof(TgItem) * ni, hipMemcpyHostToDevice));
  GLSX_HIP(hipMemset(p->d_state, 0, sizeof(unsigned) * (kTgStateHeader + np)));
  if (p->cfg.mode == 2)
    if (int rc = tg_fork_create(&p->fork)) return rc;
  if (p->use_graph) {
    GLSX_HIP(hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking));
    GLSX_HIP(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal));
    int rc = p->cfg.mode == 1 ? tg_launch(ds, p->d_passes, p->d_items, (int)ni, p->d_state, (int)np, p->cfg,
                                          p->host.lds_bytes, p->grid, p->capture_stream)
                              : tg_launch_levels(ds, p->d_passes, p->d_items, p->host, p->cfg, p->capture_stream,
                                                 &p->fork);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->capture_stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(GLSX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    p->graph = g;
    GLSX_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  }
  GLSX_HIP(hipDeviceSynchronize());
  return GLSX_OK;
}

int launch_plan(const DeviceState* ds, glsx_plan p, hipStream_t s) {
  if (p->use_graph && p->exec) {
    GLSX_HIP(hipGraphLaunch(p->exec, s));
    return GLSX_OK;
  }
  if (p->cfg.mode != 1) return tg_launch_levels(ds, p->d_passes, p->d_items, p->host, p->cfg, s, &p->fork);
  return tg_launch(ds, p->d_passes, p->d_items, (int)p->host.items.size(), p->d_state, (int)p->host.passes.size(),
                   p->cfg, p->host.lds_bytes, p->grid, s);
}

// Time one configuration: 1 warm-up and `reps` timed launches, median in us.
int time_plan(const DeviceState* ds, glsx_plan p, int reps, double* us) {
  hipEvent_t e0, e1;
  GLSX_HIP(hipEventCreate(&e0));
  GLSX_HIP(hipEventCreate(&e1));
  std::vector<float> t;
  int rc = launch_plan(ds, p, nullptr);
  for (int r = 0; r < reps && !rc; ++r) {
    hipEventRecord(e0, nullptr);
    rc = launch_plan(ds, p, nullptr);
    hipEventRecord(e1, nullptr);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    t.push_back(ms);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  std::sort(t.begin(), t.end());
  *us = t[t.size() / 2] * 1000.0;
  return GLSX_OK;
}

std::string fft_key(const DeviceState* ds, const glsx_fft_desc* d, int count) {
  std::vector<std::string> parts;
  for (int i = 0; i < count; ++i) {
    char b[64];
    std::snprintf(b, sizeof(b), "%dx%d:%d:%d", d[i].rows, d[i].cols, d[i].src_type, d[i].flags);
    parts.push_back(b);
  }
  std::sort(parts.begin(), parts.end());
  unsigned long long h = 1469598103934665603ull;  // FNV-1a over the sorted shape list
  for (auto& s : parts)
    for (char c : s) h = (h ^ (unsigned char)c) * 1099511628211ull;
  char k[128];
  std::snprintf(k, sizeof(k), "fftbatch|%s|n%d|%016llx", arch_name(ds->id).c_str(), count, h);
  return k;
}

// Column tiles wider than 16 columns pay where a batch has many short columns
// (a 64-point column needs 4 threads, so 16 columns fill 64 of 512 threads):
// measured on the 256-FFT batch of shapes 64..2048, 512 threads, level mode:
// 0.815 ms at 16, 0.791 at 32, 0.783 at 64 columns.
// Grouped levels (mode 2) take every tile's shape from the plain path, so
// one candidate covers them.
const TgConfig kCandidates[] = {{256, 16, 2}, {256, 16, 0}, {256, 4, 0},  {256, 64, 0}, {512, 64, 0}, {512, 32, 0},
                                {512, 16, 0}, {512, 8, 0},  {512, 4, 0},  {1024, 32, 0}, {1024, 16, 0},
                                {1024, 8, 0}, {1024, 4, 0}, {256, 16, 1}, {512, 16, 1}, {1024, 8, 1}};

// Autotune an FFT batch: run every candidate on scratch buffers of the same
// shapes (the caller's buffers may be in place), keep the fastest.
int autotune_fft(const DeviceState* ds, glsx_plan p, const glsx_fft_desc* d, int count) {
  const std::string key = fft_key(ds, d, count);
  TgConfig cfg;
  double us = 0;
  if (db_lookup(key, &cfg, &us)) {
    p->cfg = cfg;
    p->tuned = 2;  // from the database
    p->tuned_us = us;
    return GLSX_OK;
  }
  size_t total = 0;
  std::vector<size_t> off(count);
  for (int i = 0; i < count; ++i) {
    off[i] = total;
    total = align_up(total + (size_t)d[i].rows * d[i].cols * sizeof(float2));
  }
  void* scratch = nullptr;
  GLSX_HIP(hipMalloc(&scratch, std::max<size_t>(total, 256)));
  GLSX_HIP(hipMemset(scratch, 0, std::max<size_t>(total, 256)));
  std::vector<glsx_fft_desc> sd(d, d + count);
  for (int i = 0; i < count; ++i) {
    sd[i].src = (char*)scratch + off[i];
    sd[i].dst = (char*)scratch + off[i];
    sd[i].src_type = GLSX_CV_32FC2;
  }
  std::vector<TgPassSpec> specs;
  int rc = fft_specs(ds, sd.data(), count, &specs);
  double best = 1e30;
  TgConfig best_cfg = p->cfg;
  const bool graph = p->use_graph;
  p->use_graph = false;
  for (const TgConfig& c : kCandidates) {
    if (rc) break;
    p->cfg = c;
    if (realize(ds, p, specs) != GLSX_OK) continue;  // shape does not fit this candidate
    double t = 0;
    if (time_plan(ds, p, 5, &t) != GLSX_OK) continue;
    if (t < best) {
      best = t;
      best_cfg = c;
    }
  }
  free_device(p);
  p->use_graph = graph;
  hipFree(scratch);
  if (rc) return rc;
  if (best >= 1e30) return fail(GLSX_ERR_UNSUPPORTED, "autotune: no candidate configuration fits");
  p->cfg = best_cfg;
  p->tuned = 1;
  p->tuned_us = best;
  db_store(key, best_cfg, best);
  return GLSX_OK;
}

std::string poc_key(const DeviceState* ds, const glsx_poc_desc* d, int count) {
  std::vector<std::string> parts;
  for (int i = 0; i < count; ++i) {
    char b[64];
    std::snprintf(b, sizeof(b), "%dx%d:%d", d[i].rows, d[i].cols, d[i].window ? 1 : 0);
    parts.push_back(b);
  }
  std::sort(parts.begin(), parts.end());
  unsigned long long h = 1469598103934665603ull;
  for (auto& s : parts)
    for (char c : s) h = (h ^ (unsigned char)c) * 1099511628211ull;
  char k[128];
  std::snprintf(k, sizeof(k), "pocbatch|%s|n%d|%016llx", arch_name(ds->id).c_str(), count, h);
  return k;
}

// Autotune a phase-correlation batch on its own buffers: the passes read the
// images and write only the plan's workspace, so candidates run in place.
int autotune_poc(const DeviceState* ds, glsx_plan p, const glsx_poc_desc* d, int count,
                 const std::vector<TgPassSpec>& specs) {
  const std::string key = poc_key(ds, d, count);
  TgConfig cfg;
  double us = 0;
  if (db_lookup(key, &cfg, &us)) {
    p->cfg = cfg;
    p->tuned = 2;
    p->tuned_us = us;
    return GLSX_OK;
  }
  double best = 1e30;
  TgConfig best_cfg = p->cfg;
  const bool graph = p->use_graph;
  p->use_graph = false;
  for (const TgConfig& c : kCandidates) {
    p->cfg = c;
    if (realize(ds, p, specs) != GLSX_OK) continue;  // a tile does not fit this candidate
    double t = 0;
    if (time_plan(ds, p, 5, &t) != GLSX_OK) continue;
    if (t < best) {
      best = t;
      best_cfg = c;
    }
  }
  free_device(p);
  p->use_graph = graph;
  if (best >= 1e30) return fail(GLSX_ERR_UNSUPPORTED, "autotune: no candidate configuration fits");
  p->cfg = best_cfg;
  p->tuned = 1;
  p->tuned_us = best;
  db_store(key, best_cfg, best);
  return GLSX_OK;
}

}  // namespace

#define GLSX_API extern "C" __attribute__((visibility("default")))

static int apply_config(glsx_plan p, const glsx_plan_config* c) {
  if (!c) return GLSX_OK;
  if (c->threads != 256 && c->threads != 512 && c->threads != 1024)
    return fail(GLSX_ERR_ARG, "plan config: threads must be 256, 512 or 1024");
  if (c->col_wcap < 1 || c->col_wcap > 64 || (c->col_wcap & (c->col_wcap - 1)))
    return fail(GLSX_ERR_ARG, "plan config: col_wcap must be a power of two in [1, 64]");
  if (c->persistent < 0 || c->persistent > 2) return fail(GLSX_ERR_ARG, "plan config: persistent must be 0, 1 or 2");
  p->cfg.threads = c->threads;
  p->cfg.col_wcap = c->col_wcap;
  p->cfg.mode = c->persistent;
  return GLSX_OK;
}

GLSX_API int glsx_plan_fft_batch(const glsx_fft_desc* descs, int count, const glsx_plan_config* config, int options,
                                 glsx_plan* out) {
  if (!descs || count <= 0 || !out) return fail(GLSX_ERR_ARG, "fft batch: bad arguments");
  *out = nullptr;
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  auto* p = new glsx_plan_s();
  p->device = ds->id;
  p->kind = 0;
  p->use_graph = (options & GLSX_PLAN_GRAPH) != 0;
  std::vector<TgPassSpec> specs;
  int rc = apply_config(p, config);
  if (!rc) rc = fft_specs(ds, descs, count, &specs);
  if (!rc && !config && (options & GLSX_PLAN_AUTOTUNE)) rc = autotune_fft(ds, p, descs, count);
  if (!rc) rc = realize(ds, p, specs);
  if (rc) {
    free_device(p);
    delete p;
    return rc;
  }
  ds->live_objects.fetch_add(1);
  *out = p;
  return GLSX_OK;
}

GLSX_API size_t glsx_poc_batch_workspace_bytes(const glsx_poc_desc* descs, int count) {
  if (!descs || count <= 0) return 0;
  PocLayout l;
  if (poc_layout(descs, count, &l)) return 0;
  return l.total;
}

GLSX_API int glsx_plan_poc_batch(const glsx_poc_desc* descs, int count, void* workspace, size_t workspace_bytes,
                                 const glsx_plan_config* config, int options, glsx_plan* out) {
  if (!descs || count <= 0 || !out || !workspace) return fail(GLSX_ERR_ARG, "poc batch: bad arguments");
  *out = nullptr;
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  PocLayout l;
  if (int rc = poc_layout(descs, count, &l)) return rc;
  if (workspace_bytes < l.total) return fail(GLSX_ERR_ARG, "poc batch: workspace too small");
  auto* p = new glsx_plan_s();
  p->device = ds->id;
  p->kind = 1;
  p->use_graph = (options & GLSX_PLAN_GRAPH) != 0;
  p->d_results = (glsx_poc_result*)((char*)workspace + l.results);
  p->n_results = count;
  std::vector<TgPassSpec> specs;
  int rc = apply_config(p, config);
  // default shape: grouped levels, every pass on the plain path's own kernel
  // (the 128-request serving mix: 2.49 ms against 3.03 ms for one generic
  // launch per level of 512-thread tiles); the workgroup size applies to the
  // other modes
  if (!config) {
    p->cfg.threads = std::max(512, poc_min_threads(descs, count));
    p->cfg.mode = 2;
  }
  if (!rc) rc = poc_specs(ds, descs, count, (char*)workspace, l, &specs);
  if (!rc && !config && (options & GLSX_PLAN_AUTOTUNE)) rc = autotune_poc(ds, p, descs, count, specs);
  if (!rc) rc = realize(ds, p, specs);
  if (rc) {
    free_device(p);
    delete p;
    return rc;
  }
  ds->live_objects.fetch_add(1);
  *out = p;
  return GLSX_OK;
}

GLSX_API int glsx_plan_execute(glsx_plan plan, void* stream) {
  if (!plan) return fail(GLSX_ERR_ARG, "null plan");
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (ds->id != plan->device) return fail(GLSX_ERR_STATE, "plan belongs to another device");
  return launch_plan(ds, plan, (hipStream_t)stream);
}

GLSX_API int glsx_plan_poc_results(glsx_plan plan, glsx_poc_result* host_out, void* stream) {
  if (!plan || !host_out || plan->kind != 1) return fail(GLSX_ERR_ARG, "poc results: bad arguments");
  GLSX_HIP(hipMemcpyAsync(host_out, plan->d_results, sizeof(glsx_poc_result) * plan->n_results,
                          hipMemcpyDeviceToHost, (hipStream_t)stream));
  GLSX_HIP(hipStreamSynchronize((hipStream_t)stream));
  return GLSX_OK;
}

GLSX_API int glsx_plan_check(glsx_plan plan) {
  if (!plan) return fail(GLSX_ERR_ARG, "null plan");
  GLSX_HIP(hipDeviceSynchronize());
  unsigned err = 0;
  GLSX_HIP(hipMemcpy(&err, plan->d_state + 2, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (err) {
    const unsigned zero = 0;
    GLSX_HIP(hipMemcpy(plan->d_state + 2, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
    return fail(GLSX_ERR_STATE, "task graph: a dependency wait exceeded its bound");
  }
  return GLSX_OK;
}

GLSX_API int glsx_plan_get_info(glsx_plan plan, glsx_plan_info* info) {
  if (!plan || !info) return fail(GLSX_ERR_ARG, "plan info: bad arguments");
  std::memset(info, 0, sizeof(*info));
  info->threads = plan->cfg.threads;
  info->col_wcap = plan->cfg.col_wcap;
  info->persistent = plan->cfg.mode;
  info->n_levels = (int)plan->host.level_start.size();
  info->grid = plan->grid;
  info->n_passes = (int)plan->host.passes.size();
  info->n_items = (int)plan->host.items.size();
  info->lds_bytes = plan->host.lds_bytes;
  info->uses_graph = plan->use_graph && plan->exec ? 1 : 0;
  info->tuned = plan->tuned;
  info->tuned_us = plan->tuned_us;
  return GLSX_OK;
}

GLSX_API int glsx_plan_destroy(glsx_plan plan) {
  if (!plan) return fail(GLSX_ERR_ARG, "null plan");
  hipDeviceSynchronize();
  free_device(plan);
  if (DeviceState* ds = device_slot(plan->device)) ds->live_objects.fetch_sub(1);
  delete plan;
  return GLSX_OK;
}
