// server.hip -- continuous-batching phase-correlation server.
//
// Requests (pairs of real images, any power-of-two size up to 4096) are
// submitted from any thread; a scheduler thread admits them against an HBM
// budget into batches and runs each batch as one persistent task-graph launch
// (glsx_plan machinery) on one of several streams, so uploads, compute and
// result copies of different batches overlap.  A batch is launched as soon as
// a stream slot is free -- requests that arrive while others are on the GPU
// join the next launch, and memory freed by a finished batch is reused by the
// next admission.  Spectra and correlation surfaces never leave the device;
// only a 40-byte result per request comes back.
//
// The reference serves one request at a time: gls::phaseCorrelate issues
// every FFT stage as its own draw and downloads the whole correlation surface
// for a CPU centroid (glsCV/glsPhaseCorrelate.cpp:178-230).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "host_copy.hpp"
#include "taskgraph.hpp"

using namespace glsx;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit sub-allocator over one device allocation (offsets, 256-byte
// granularity, neighbours coalesced on release).
class Arena {
 public:
  size_t size = 0, used = 0, peak = 0;
  bool alloc(size_t n, size_t* off) {
    n = align_up(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n;
        const size_t start = it->first + n;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used += n;
        peak = std::max(peak, used);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = align_up(n);
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t largest_free() const {
    size_t m = 0;
    for (auto& f : free_) m = std::max(m, f.second);
    return m;
  }
  void reset(size_t total) {
    size = total;
    used = peak = 0;
    free_.clear();
    free_[0] = total;
  }

 private:
  std::map<size_t, size_t> free_;
};

struct Request {
  long long id = 0;
  int rows = 0, cols = 0;
  bool has_win = false;
  float* host = nullptr;  // the server's pinned staging: img1 | img2 | window (nullptr: caller's memory)
  size_t host_bytes = 0;
  const float *src1 = nullptr, *src2 = nullptr, *srcw = nullptr;  // pinned images the copies read
  double t_submit = 0.0;
};

size_t image_bytes(const Request& r) { return (size_t)r.rows * r.cols * sizeof(float); }

// Upper bound of the device bytes one request adds to a batch block.
size_t request_device_bytes(const Request& r) {
  const size_t n = (size_t)r.rows * r.cols;
  const size_t images = align_up(n * 4) * (r.has_win ? 3 : 2);
  const size_t ws = align_up(n * 8) + align_up(n * 4) + align_up((size_t)r.cols * 8) + sizeof(glsx_poc_result);
  const size_t items = (size_t)(3 * r.rows + 3 * r.cols + 1) * sizeof(TgItem);
  return images + ws + items + 7 * sizeof(TgPass) + 7 * sizeof(unsigned) + 4096;
}

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Request> reqs;
  size_t block_off = 0, block_bytes = 0;
  unsigned char* host_tables = nullptr;  // pinned staging of the pass/item tables
  size_t host_tables_cap = 0;
  glsx_poc_result* host_results = nullptr;  // pinned
  unsigned* host_state = nullptr;           // pinned copy of the kernel's error flag
  int results_cap = 0;
  double t_launch = 0.0;
};

}  // namespace

struct glsx_server_s {
  glsx_server_config cfg{};
  DeviceState* ds = nullptr;
  char* dev_base = nullptr;
  Arena arena;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work;  // scheduler: new submission or stop
  std::condition_variable cv_done;  // pollers: completions available
  std::deque<Request> pending;
  std::deque<glsx_server_completion> completed;
  std::atomic<long long> next_id{1};
  bool stop = false;
  std::thread worker;
  glsx_server_stats stats{};
  std::map<size_t, int> grid_cache;  // lds bytes -> persistent grid
  std::mutex pinned_mu;
  std::multimap<size_t, float*> pinned_free;
};

namespace {

float* pinned_get(glsx_server_s* s, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(s->pinned_mu);
    auto it = s->pinned_free.lower_bound(bytes);
    if (it != s->pinned_free.end() && it->first <= 2 * bytes) {
      float* p = it->second;
      s->pinned_free.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
  return (float*)p;
}

void pinned_put(glsx_server_s* s, float* p, size_t bytes) {
  if (!p) return;  // the caller's own pinned memory
  std::lock_guard<std::mutex> lk(s->pinned_mu);
  s->pinned_free.emplace(bytes, p);
}

void complete(glsx_server_s* s, const Request& r, int status, const glsx_poc_result* res, int batch, double t) {
  glsx_server_completion c{};
  c.id = r.id;
  c.status = status;
  if (res) c.result = *res;
  c.latency_s = t - r.t_submit;
  c.batch_size = batch;
  s->completed.push_back(c);
  if (status == GLSX_OK) s->stats.completed++;
  else s->stats.rejected++;
}

// Launches one batch on `slot`; the requests are already admitted and the
// block [block_off, block_off + block_bytes) of the arena is theirs.
int launch_batch(glsx_server_s* s, Slot& slot) {
  const int count = (int)slot.reqs.size();
  char* block = s->dev_base + slot.block_off;
  size_t off = 0;
  std::vector<glsx_poc_desc> descs(count);
  for (int i = 0; i < count; ++i) {
    Request& r = slot.reqs[i];
    const size_t ib = image_bytes(r);
    float* d1 = (float*)(block + off);
    off = align_up(off + ib);
    float* d2 = (float*)(block + off);
    off = align_up(off + ib);
    float* dw = nullptr;
    if (r.has_win) {
      dw = (float*)(block + off);
      off = align_up(off + ib);
    }
    GLSX_HIP(hipMemcpyAsync(d1, r.src1, ib, hipMemcpyHostToDevice, slot.stream));
    GLSX_HIP(hipMemcpyAsync(d2, r.src2, ib, hipMemcpyHostToDevice, slot.stream));
    if (dw) GLSX_HIP(hipMemcpyAsync(dw, r.srcw, ib, hipMemcpyHostToDevice, slot.stream));
    descs[i] = glsx_poc_desc{d1, d2, dw, r.rows, r.cols};
  }
  PocLayout l;
  if (int rc = poc_layout(descs.data(), count, &l)) return rc;
  char* ws = block + off;
  off = align_up(off + l.total);
  std::vector<TgPassSpec> specs;
  if (int rc = poc_specs(s->ds, descs.data(), count, ws, l, &specs)) return rc;
  TgConfig cfg;
  cfg.threads = s->cfg.threads > 0 ? s->cfg.threads : std::max(512, poc_min_threads(descs.data(), count));
  cfg.mode = s->cfg.persistent;
  TgPlanHost host;
  if (int rc = tg_build(s->ds, specs, cfg, &host)) return rc;
  for (auto& ps : host.passes)
    if (ps.kind == PK_POC_FINAL) ps.fin.n_partials = host.passes[ps.deps[0]].n_items;
  const size_t np = host.passes.size(), ni = host.items.size();
  const size_t pass_bytes = np * sizeof(TgPass), item_bytes = ni * sizeof(TgItem);
  const size_t state_bytes = (kTgStateHeader + np) * sizeof(unsigned);
  TgPass* d_passes = (TgPass*)(block + off);
  off = align_up(off + pass_bytes);
  TgItem* d_items = (TgItem*)(block + off);
  off = align_up(off + item_bytes);
  unsigned* d_state = (unsigned*)(block + off);
  off = align_up(off + state_bytes);
  if (off > slot.block_bytes) return fail(GLSX_ERR_BUDGET, "server: batch block overflow");
  // tables through pinned staging (the slot's previous batch is finished)
  const size_t tb = pass_bytes + item_bytes;
  if (tb > slot.host_tables_cap) {
    if (slot.host_tables) hipHostFree(slot.host_tables);
    slot.host_tables = nullptr;
    GLSX_HIP(hipHostMalloc((void**)&slot.host_tables, tb * 2, hipHostMallocDefault));
    slot.host_tables_cap = tb * 2;
  }
  std::memcpy(slot.host_tables, host.passes.data(), pass_bytes);
  std::memcpy(slot.host_tables + pass_bytes, host.items.data(), item_bytes);
  GLSX_HIP(hipMemcpyAsync(d_passes, slot.host_tables, pass_bytes, hipMemcpyHostToDevice, slot.stream));
  GLSX_HIP(hipMemcpyAsync(d_items, slot.host_tables + pass_bytes, item_bytes, hipMemcpyHostToDevice, slot.stream));
  GLSX_HIP(hipMemsetAsync(d_state, 0, state_bytes, slot.stream));
  if (cfg.mode == 1) {
    auto g = s->grid_cache.find(host.lds_bytes);
    int grid = 0;
    if (g == s->grid_cache.end()) {
      if (int rc = tg_grid(s->ds, cfg, host.lds_bytes, &grid)) return rc;
      s->grid_cache[host.lds_bytes] = grid;
    } else {
      grid = g->second;
    }
    grid = std::max(1, std::min<int>(grid, (int)ni));
    if (int rc = tg_launch(s->ds, d_passes, d_items, (int)ni, d_state, (int)np, cfg, host.lds_bytes, grid,
                           slot.stream))
      return rc;
  } else {
    if (int rc = tg_launch_levels(s->ds, d_passes, d_items, host, cfg, slot.stream)) return rc;
  }
  if (count > slot.results_cap) {
    if (slot.host_results) hipHostFree(slot.host_results);
    slot.host_results = nullptr;
    GLSX_HIP(hipHostMalloc((void**)&slot.host_results, sizeof(glsx_poc_result) * count * 2, hipHostMallocDefault));
    slot.results_cap = count * 2;
  }
  GLSX_HIP(hipMemcpyAsync(slot.host_results, ws + l.results, sizeof(glsx_poc_result) * count,
                          hipMemcpyDeviceToHost, slot.stream));
  GLSX_HIP(hipMemcpyAsync(slot.host_state, d_state + 2, sizeof(unsigned), hipMemcpyDeviceToHost, slot.stream));
  GLSX_HIP(hipEventRecord(slot.done, slot.stream));
  return GLSX_OK;
}

// Finishes every slot whose batch is done.  Called with s->mu held.
bool reap(glsx_server_s* s) {
  bool any = false;
  for (Slot& slot : s->slots) {
    if (!slot.busy) continue;
    const hipError_t q = hipEventQuery(slot.done);
    if (q == hipErrorNotReady) continue;
    // any other error: the batch will never signal; its requests fail
    // instead of waiting forever
    if (q != hipSuccess) (void)hipGetLastError();
    const double t = now_s();
    const int status = q != hipSuccess ? GLSX_ERR_HIP : (*slot.host_state ? GLSX_ERR_STATE : GLSX_OK);
    for (size_t i = 0; i < slot.reqs.size(); ++i) {
      complete(s, slot.reqs[i], status, &slot.host_results[i], (int)slot.reqs.size(), t);
      pinned_put(s, slot.reqs[i].host, slot.reqs[i].host_bytes);
    }
    s->arena.release(slot.block_off, slot.block_bytes);
    slot.reqs.clear();
    slot.busy = false;
    any = true;
  }
  if (any) s->cv_done.notify_all();
  return any;
}

// Admits pending requests (FIFO) into a free slot and launches them.
// Called with s->mu held.
bool admit(glsx_server_s* s) {
  Slot* slot = nullptr;
  for (Slot& sl : s->slots)
    if (!sl.busy) {
      slot = &sl;
      break;
    }
  if (!slot || s->pending.empty()) return false;
  const size_t avail = s->arena.largest_free();
  size_t need = 0;
  std::vector<Request> batch;
  while (!s->pending.empty() && (int)batch.size() < s->cfg.max_batch) {
    const Request& r = s->pending.front();
    const size_t rb = request_device_bytes(r);
    if (rb > s->arena.size) {  // can never fit: reject
      complete(s, r, GLSX_ERR_BUDGET, nullptr, 0, now_s());
      pinned_put(s, r.host, r.host_bytes);
      s->pending.pop_front();
      s->cv_done.notify_all();
      continue;
    }
    if (need + rb > avail) break;  // wait for memory to free up
    need += rb;
    batch.push_back(r);
    s->pending.pop_front();
  }
  if (batch.empty()) return false;
  size_t off = 0;
  if (!s->arena.alloc(need, &off)) {  // cannot happen: need <= largest free block
    for (auto it = batch.rbegin(); it != batch.rend(); ++it) s->pending.push_front(*it);
    return false;
  }
  slot->reqs = std::move(batch);
  slot->block_off = off;
  slot->block_bytes = align_up(need);
  slot->t_launch = now_s();
  const int rc = launch_batch(s, *slot);
  if (rc) {
    for (const Request& r : slot->reqs) {
      complete(s, r, rc, nullptr, 0, now_s());
      pinned_put(s, r.host, r.host_bytes);
    }
    hipStreamSynchronize(slot->stream);
    s->arena.release(slot->block_off, slot->block_bytes);
    slot->reqs.clear();
    s->cv_done.notify_all();
    return true;
  }
  slot->busy = true;
  s->stats.batches++;
  s->stats.mean_batch += ((double)slot->reqs.size() - s->stats.mean_batch) / (double)s->stats.batches;
  int inflight = 0;
  for (Slot& sl : s->slots) inflight += sl.busy;
  s->stats.max_inflight_seen = std::max(s->stats.max_inflight_seen, inflight);
  s->stats.peak_bytes = std::max(s->stats.peak_bytes, s->arena.peak);
  return true;
}

void worker_main(glsx_server_s* s) {
  hipSetDevice(s->ds->id);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    reap(s);
    while (admit(s)) {
    }
    bool busy = false;
    for (Slot& sl : s->slots) busy |= sl.busy;
    if (s->stop && s->pending.empty() && !busy) break;
    if (busy) {
      // poll the in-flight batches; a new submission also wakes us
      s->cv_work.wait_for(lk, std::chrono::microseconds(20));
    } else {
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
    }
  }
}

}  // namespace

#define GLSX_API extern "C" __attribute__((visibility("default")))

GLSX_API int glsx_server_create(const glsx_server_config* cfg, glsx_server* out) {
  if (!cfg || !out) return fail(GLSX_ERR_ARG, "server: null argument");
  *out = nullptr;
  if (cfg->hbm_budget_bytes < (1u << 20)) return fail(GLSX_ERR_ARG, "server: budget below 1 MiB");
  if (cfg->max_batch < 1 || cfg->max_inflight < 1 || cfg->max_inflight > 8)
    return fail(GLSX_ERR_ARG, "server: max_batch >= 1 and max_inflight in [1, 8] required");
  if (cfg->threads != 0 && cfg->threads != 256 && cfg->threads != 512 && cfg->threads != 1024)
    return fail(GLSX_ERR_ARG, "server: threads must be 0, 256, 512 or 1024");
  if (cfg->persistent < 0 || cfg->persistent > 2) return fail(GLSX_ERR_ARG, "server: persistent must be 0, 1 or 2");
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  auto* s = new glsx_server_s();
  s->cfg = *cfg;
  s->ds = ds;
  s->stats.budget_bytes = cfg->hbm_budget_bytes;
  hipError_t e = hipMalloc((void**)&s->dev_base, cfg->hbm_budget_bytes);
  if (e != hipSuccess) {
    delete s;
    return fail(GLSX_ERR_NOMEM, "server: cannot allocate the HBM budget");
  }
  // counted until glsx_server_destroy (also on the failure path below)
  ds->live_objects.fetch_add(1);
  s->arena.reset(cfg->hbm_budget_bytes);
  s->slots.resize(cfg->max_inflight);
  for (Slot& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&sl.host_state, sizeof(unsigned), hipHostMallocDefault) != hipSuccess) {
      glsx_server_destroy(s);
      return fail(GLSX_ERR_HIP, "server: stream/event/pinned setup failed");
    }
    *sl.host_state = 0;
  }
  s->worker = std::thread(worker_main, s);
  *out = s;
  return GLSX_OK;
}

namespace {
int enqueue(glsx_server s, Request& r, long long* id_out);
}  // namespace

GLSX_API int glsx_server_submit(glsx_server s, const float* img1, const float* img2, const float* window, int rows,
                                int cols, long long* id_out) {
  if (!s || !img1 || !img2) return fail(GLSX_ERR_ARG, "server submit: null argument");
  auto ok = [](int n) { return n >= 2 && n <= 4096 && (n & (n - 1)) == 0; };
  if (!ok(rows) || !ok(cols)) return fail(GLSX_ERR_UNSUPPORTED, "server submit: sizes must be powers of two in [2, 4096]");
  Request r;
  r.rows = rows;
  r.cols = cols;
  r.has_win = window != nullptr;
  const size_t ib = (size_t)rows * cols * sizeof(float);
  r.host_bytes = ib * (r.has_win ? 3 : 2);
  r.host = pinned_get(s, r.host_bytes);
  if (!r.host) return fail(GLSX_ERR_NOMEM, "server submit: pinned staging allocation failed");
  staging_copy(r.host, img1, ib);
  staging_copy((char*)r.host + ib, img2, ib);
  if (window) staging_copy((char*)r.host + 2 * ib, window, ib);
  r.src1 = r.host;
  r.src2 = (const float*)((char*)r.host + ib);
  r.srcw = window ? (const float*)((char*)r.host + 2 * ib) : nullptr;
  return enqueue(s, r, id_out);
}

GLSX_API int glsx_server_submit_pinned(glsx_server s, const float* img1, const float* img2, const float* window,
                                       int rows, int cols, long long* id_out) {
  if (!s || !img1 || !img2) return fail(GLSX_ERR_ARG, "server submit: null argument");
  auto ok = [](int n) { return n >= 2 && n <= 4096 && (n & (n - 1)) == 0; };
  if (!ok(rows) || !ok(cols)) return fail(GLSX_ERR_UNSUPPORTED, "server submit: sizes must be powers of two in [2, 4096]");
  for (const float* p : {img1, img2, window}) {
    if (!p) continue;
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeHost) {
      (void)hipGetLastError();
      return fail(GLSX_ERR_ARG, "server submit_pinned: images must be pinned host memory (glsx_host_alloc)");
    }
  }
  Request r;
  r.rows = rows;
  r.cols = cols;
  r.has_win = window != nullptr;
  r.src1 = img1;
  r.src2 = img2;
  r.srcw = window;
  return enqueue(s, r, id_out);
}

namespace {

int enqueue(glsx_server s, Request& r, long long* id_out) {
  r.id = s->next_id.fetch_add(1);
  r.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      pinned_put(s, r.host, r.host_bytes);
      return fail(GLSX_ERR_STATE, "server submit: server is shutting down");
    }
    s->pending.push_back(r);
    s->stats.submitted++;
  }
  s->cv_work.notify_one();
  if (id_out) *id_out = r.id;
  return GLSX_OK;
}

}  // namespace

GLSX_API int glsx_server_poll(glsx_server s, glsx_server_completion* out, int max, int* n_out, int timeout_ms) {
  if (!s || !out || max < 1 || !n_out) return fail(GLSX_ERR_ARG, "server poll: bad arguments");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completed.empty() && timeout_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->completed.empty(); });
  int n = 0;
  while (n < max && !s->completed.empty()) {
    out[n++] = s->completed.front();
    s->completed.pop_front();
  }
  *n_out = n;
  return GLSX_OK;
}

GLSX_API int glsx_server_get_stats(glsx_server s, glsx_server_stats* st) {
  if (!s || !st) return fail(GLSX_ERR_ARG, "server stats: bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->stats;
  st->pending = (long long)s->pending.size();
  int inflight = 0;
  for (Slot& sl : s->slots) inflight += sl.busy;
  st->inflight = inflight;
  return GLSX_OK;
}

GLSX_API int glsx_server_destroy(glsx_server s) {
  if (!s) return fail(GLSX_ERR_ARG, "server: null handle");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();  // finishes every admitted and pending request
  for (Slot& sl : s->slots) {
    if (sl.stream) hipStreamSynchronize(sl.stream);
    if (sl.stream) hipStreamDestroy(sl.stream);
    if (sl.done) hipEventDestroy(sl.done);
    if (sl.host_tables) hipHostFree(sl.host_tables);
    if (sl.host_results) hipHostFree(sl.host_results);
    if (sl.host_state) hipHostFree(sl.host_state);
  }
  for (auto& p : s->pinned_free) hipHostFree(p.second);
  if (s->dev_base) hipFree(s->dev_base);
  s->ds->live_objects.fetch_sub(1);
  delete s;
  return GLSX_OK;
}
