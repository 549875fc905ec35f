// glsx_api.cpp -- the extern "C" entry points of libglsx (declared in
// include/glsx.h): context, GlsMat storage, FFT and phase correlation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "fft_pass.hpp"
#include "glsx_internal.hpp"

namespace glsx {

namespace {
thread_local std::string g_last_error;
std::mutex g_init_mutex;
DeviceState g_devices[16];
// set under g_init_mutex; read without it on every call (acquire/release order
// makes the initialised DeviceState visible to the reader)
std::atomic<bool> g_dev_ready[16] = {};
int g_default_device = -1;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg) {
  set_error(msg);
  return status;
}

static int init_device_locked(int dev) {
  int count = 0;
  GLSX_HIP(hipGetDeviceCount(&count));
  if (dev < 0 || dev >= count || dev >= 16)
    return fail(GLSX_ERR_ARG, "glsx_init: device " + std::to_string(dev) + " not present");
  if (g_dev_ready[dev].load(std::memory_order_acquire)) return GLSX_OK;
  GLSX_HIP(hipSetDevice(dev));
  hipDeviceProp_t prop;
  GLSX_HIP(hipGetDeviceProperties(&prop, dev));
  DeviceState& ds = g_devices[dev];
  ds.id = dev;
  ds.num_cu = prop.multiProcessorCount;
  ds.lds_per_block = prop.sharedMemPerBlock;
  if (int rc = init_twiddles(&ds)) return rc;
  if (int rc = init_static_workspace(&ds)) return rc;
  g_dev_ready[dev].store(true, std::memory_order_release);
  return GLSX_OK;
}

DeviceState* device_slot(int dev) { return dev >= 0 && dev < 16 ? &g_devices[dev] : nullptr; }

std::unique_lock<std::mutex> stream_lock(DeviceState* ds, hipStream_t stream) {
  std::mutex* m = nullptr;
  {
    std::lock_guard<std::mutex> lk(ds->scratch_mutex);
    m = &ds->stream_locks[(void*)stream];
  }
  return std::unique_lock<std::mutex>(*m);
}

int stream_scratch(DeviceState* ds, hipStream_t stream, int slot, size_t bytes, void** out) {
  std::lock_guard<std::mutex> lk(ds->scratch_mutex);
  auto& e = ds->scratch[{(void*)stream, slot}];
  if (e.second < bytes) {
    if (e.first) {
      // work queued on this stream may still use the old buffer
      GLSX_HIP(hipStreamSynchronize(stream));
      GLSX_HIP(hipFree(e.first));
      e = {nullptr, 0};
    }
    const size_t want = std::max<size_t>(bytes, 4096);
    GLSX_HIP(hipMalloc(&e.first, want));
    e.second = want;
  }
  *out = e.first;
  return GLSX_OK;
}

DeviceState* device_state(int* status) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *status = fail(GLSX_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    return nullptr;
  }
  if (dev < 0 || dev >= 16) {
    *status = fail(GLSX_ERR_ARG, "device index out of range");
    return nullptr;
  }
  if (!g_dev_ready[dev].load(std::memory_order_acquire)) {
    std::lock_guard<std::mutex> lk(g_init_mutex);
    int rc = init_device_locked(dev);
    if (rc) {
      *status = rc;
      return nullptr;
    }
  }
  *status = GLSX_OK;
  return &g_devices[dev];
}

DeviceState* device_state_if_ready(int* status) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *status = fail(GLSX_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    return nullptr;
  }
  if (dev < 0 || dev >= 16 || !g_dev_ready[dev].load(std::memory_order_acquire)) {
    *status = fail(GLSX_ERR_STATE, "device " + std::to_string(dev) + " not initialised: call glsx_init first");
    return nullptr;
  }
  *status = GLSX_OK;
  return &g_devices[dev];
}

// ---- the block cache of released GlsMat buffers ---------------------------------
// Callers hold ds->cache_mutex and have the block's device current.
static void free_cached_locked(DeviceState* ds, size_t i) {
  DeviceState::CachedBlock b = ds->cache[i];
  if (b.ready) {
    (void)hipEventSynchronize(b.ready);
    ds->spare_events.push_back(b.ready);
  }
  (void)hipFree(b.ptr);
  ds->cached_bytes -= b.bytes;
  ds->cache[i] = ds->cache.back();
  ds->cache.pop_back();
}

// A cached block of exactly `bytes` whose last use has completed, or nullptr.
static void* take_cached(DeviceState* ds, size_t bytes) {
  std::lock_guard<std::mutex> lk(ds->cache_mutex);
  for (size_t i = 0; i < ds->cache.size(); ++i) {
    DeviceState::CachedBlock& b = ds->cache[i];
    if (b.bytes != bytes) continue;
    if (b.ready) {
      if (hipEventQuery(b.ready) != hipSuccess) {
        (void)hipGetLastError();  // hipErrorNotReady is not an error of the caller's
        continue;
      }
      ds->spare_events.push_back(b.ready);
    }
    void* p = b.ptr;
    ds->cached_bytes -= b.bytes;
    ds->cache[i] = ds->cache.back();
    ds->cache.pop_back();
    return p;
  }
  return nullptr;
}

static hipEvent_t spare_event_locked(DeviceState* ds) {
  hipEvent_t ev = nullptr;
  if (!ds->spare_events.empty()) {
    ev = ds->spare_events.back();
    ds->spare_events.pop_back();
  } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    ev = nullptr;
  }
  return ev;
}

// Frees every cached block of the current device (each waits for its event).
static void trim_cache(DeviceState* ds) {
  std::lock_guard<std::mutex> lk(ds->cache_mutex);
  while (!ds->cache.empty()) free_cached_locked(ds, ds->cache.size() - 1);
}

// Puts a released block into the cache.  `stream` is where its last use was
// enqueued (ready once an event there completes); has_stream false means the
// block has no pending use.
// A block whose readiness is already recorded in `recorded` (an event of the
// device's pool) is passed with has_stream false.
static int cache_block(DeviceState* ds, void* ptr, size_t bytes, bool has_stream, hipStream_t stream,
                       hipEvent_t recorded = nullptr) {
  std::lock_guard<std::mutex> lk(ds->cache_mutex);
  hipEvent_t ev = recorded;
  if (has_stream) {
    ev = spare_event_locked(ds);
    if (!ev) return fail(GLSX_ERR_HIP, "glsx_buf_release: no event");
    hipError_t e = hipEventRecord(ev, stream);
    if (e != hipSuccess) {
      ds->spare_events.push_back(ev);
      return fail(GLSX_ERR_HIP, std::string("glsx_buf_release: hipEventRecord: ") + hipGetErrorString(e));
    }
  }
  ds->cache.push_back({ptr, bytes, ev});
  ds->cached_bytes += bytes;
  // over the cap: free the oldest blocks (each waits only for its own event)
  while (ds->cached_bytes > kBlockCacheCap && ds->cache.size() > 1) free_cached_locked(ds, 0);
  return GLSX_OK;
}

void note_use(glsx_buf_s* b, hipStream_t stream) {
  if (!b) return;
  std::lock_guard<std::mutex> lk(b->use_mutex);
  if (b->used && b->last_stream != stream) b->mixed = true;
  b->used = true;
  b->last_stream = stream;
  if (b->mixed || b->exported || stream == nullptr) return;
  DeviceState* ds = device_slot(b->device);
  if (!b->ev) {
    std::lock_guard<std::mutex> ck(ds->cache_mutex);
    b->ev = spare_event_locked(ds);
  }
  // without an event the release falls back to waiting for the device
  if (!b->ev || hipEventRecord(b->ev, stream) != hipSuccess) {
    (void)hipGetLastError();
    b->mixed = true;
  }
}

// Last reference gone: hand the block to the cache.  `stream` (when
// has_stream) orders every use of the buffer before the point reached now.
static int release_block(glsx_buf buf, bool has_stream, hipStream_t stream) {
  int cur = 0;
  GLSX_HIP(hipGetDevice(&cur));
  if (cur != buf->device) GLSX_HIP(hipSetDevice(buf->device));
  DeviceState* ds = device_slot(buf->device);
  int rc = GLSX_OK;
  hipEvent_t recorded = nullptr;  // the buffer's own event, recorded at its last use
  {
    std::lock_guard<std::mutex> lk(buf->use_mutex);
    if (!has_stream) {
      const bool wait_device = buf->exported || buf->mixed;
      if (wait_device) {
        // the device pointer was used by work the library did not see (or on
        // several streams): wait for the device before the block can be reused
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = fail(GLSX_ERR_HIP, std::string("glsx_buf_release: ") + hipGetErrorString(e));
      } else if (buf->used && buf->last_stream == nullptr) {
        has_stream = true;  // the null stream: record now
        stream = nullptr;
      } else if (buf->used) {
        recorded = buf->ev;  // recorded after the last use on its stream
        buf->ev = nullptr;
      }
    }
    if (buf->ev) {
      std::lock_guard<std::mutex> ck(ds->cache_mutex);
      ds->spare_events.push_back(buf->ev);
      buf->ev = nullptr;
    }
  }
  if (!rc) rc = cache_block(ds, buf->ptr, buf->block, has_stream, stream, recorded);
  if (rc) {
    // could not defer: wait for the device and free now
    (void)hipDeviceSynchronize();
    (void)hipFree(buf->ptr);
  }
  if (cur != buf->device) (void)hipSetDevice(cur);
  delete buf;
  return rc;
}

}  // namespace glsx

using namespace glsx;

#define GLSX_API extern "C" __attribute__((visibility("default")))

GLSX_API int glsx_init(int device) {
  std::lock_guard<std::mutex> lk(g_init_mutex);
  if (int rc = init_device_locked(device)) return rc;
  g_default_device = device;
  return GLSX_OK;
}

GLSX_API int glsx_shutdown(void) {
  std::lock_guard<std::mutex> lk(g_init_mutex);
  // plans and servers hold the twiddle tables: refuse while any is alive
  for (int d = 0; d < 16; ++d) {
    if (!g_dev_ready[d].load(std::memory_order_acquire)) continue;
    if (int n = g_devices[d].live_objects.load())
      return fail(GLSX_ERR_STATE, "glsx_shutdown: " + std::to_string(n) + " plan(s) or server(s) still alive on device " +
                                      std::to_string(d) + "; destroy them first");
  }
  for (int d = 0; d < 16; ++d) {
    DeviceState& ds = g_devices[d];
    const bool ready = g_dev_ready[d].load(std::memory_order_acquire);
    bool cached = false;
    {
      std::lock_guard<std::mutex> ck(ds.cache_mutex);
      cached = !ds.cache.empty() || !ds.spare_events.empty();
    }
    if (!ready && !cached) continue;
    GLSX_HIP(hipSetDevice(d));
    GLSX_HIP(hipDeviceSynchronize());
    trim_cache(&ds);
    {
      std::lock_guard<std::mutex> ck(ds.cache_mutex);
      for (hipEvent_t e : ds.spare_events) (void)hipEventDestroy(e);
      ds.spare_events.clear();
    }
    if (!ready) continue;
    if (ds.twiddles) GLSX_HIP(hipFree(ds.twiddles));
    {
      std::lock_guard<std::mutex> sk(ds.scratch_mutex);
      for (auto& kv : ds.scratch)
        if (kv.second.first) GLSX_HIP(hipFree(kv.second.first));
      ds.scratch.clear();
    }
    ds.twiddles = nullptr;
    g_dev_ready[d].store(false, std::memory_order_release);
  }
  if (g_default_device >= 0) GLSX_HIP(hipSetDevice(g_default_device));
  return GLSX_OK;
}

GLSX_API const char* glsx_last_error(void) { return g_last_error.c_str(); }

GLSX_API int glsx_get_device_info(glsx_device_info* out) {
  if (!out) return fail(GLSX_ERR_ARG, "null output");
  int dev = 0;
  GLSX_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  GLSX_HIP(hipGetDeviceProperties(&prop, dev));
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof(out->name), "%s", prop.name);
  std::snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
  out->compute_units = prop.multiProcessorCount;
  out->lds_per_block = prop.sharedMemPerBlock;
  out->global_mem_bytes = prop.totalGlobalMem;
  return GLSX_OK;
}

// ---------------------------------------------------------------------------
// GlsMat storage
// ---------------------------------------------------------------------------
GLSX_API int glsx_buf_create(int rows, int cols, int type, glsx_buf* out) {
  if (!out) return fail(GLSX_ERR_ARG, "glsx_buf_create: null output");
  *out = nullptr;
  if (rows <= 0 || cols <= 0) return fail(GLSX_ERR_ARG, "glsx_buf_create: rows and cols must be positive");
  if (type < 0 || type >= 512 || type_depth(type) == 7) return fail(GLSX_ERR_ARG, "glsx_buf_create: bad type");
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  auto b = std::make_unique<glsx_buf_s>();
  b->rows = rows;
  b->cols = cols;
  b->type = type;
  b->device = ds->id;
  b->bytes = (size_t)rows * cols * elem_size(type);
  b->block = (b->bytes + 255) & ~(size_t)255;
  b->ptr = take_cached(ds, b->block);
  if (!b->ptr) {
    hipError_t e = hipMalloc(&b->ptr, b->block);
    if (e != hipSuccess) {
      // give the cached blocks back and try once more
      (void)hipGetLastError();
      trim_cache(ds);
      e = hipMalloc(&b->ptr, b->block);
    }
    if (e != hipSuccess) {
      b->ptr = nullptr;
      return fail(GLSX_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
  }
  *out = b.release();
  return GLSX_OK;
}

GLSX_API int glsx_buf_retain(glsx_buf buf) {
  if (!buf) return fail(GLSX_ERR_ARG, "null buffer");
  buf->refs.fetch_add(1);
  return GLSX_OK;
}

GLSX_API int glsx_buf_release(glsx_buf buf) {
  if (!buf) return fail(GLSX_ERR_ARG, "null buffer");
  if (buf->refs.fetch_sub(1) == 1) return release_block(buf, false, nullptr);
  return GLSX_OK;
}

GLSX_API int glsx_buf_release_async(glsx_buf buf, void* stream) {
  if (!buf) return fail(GLSX_ERR_ARG, "null buffer");
  if (buf->refs.fetch_sub(1) == 1) return release_block(buf, true, (hipStream_t)stream);
  return GLSX_OK;
}

GLSX_API int glsx_buf_info(glsx_buf buf, int* rows, int* cols, int* type, void** dev_ptr, size_t* bytes) {
  if (!buf) return fail(GLSX_ERR_ARG, "null buffer");
  if (rows) *rows = buf->rows;
  if (cols) *cols = buf->cols;
  if (type) *type = buf->type;
  if (dev_ptr) {
    *dev_ptr = buf->ptr;
    std::lock_guard<std::mutex> lk(buf->use_mutex);
    buf->exported = true;
  }
  if (bytes) *bytes = buf->bytes;
  return GLSX_OK;
}

GLSX_API int glsx_buf_upload(glsx_buf buf, const void* host, size_t host_pitch, void* stream) {
  if (!buf || !host) return fail(GLSX_ERR_ARG, "glsx_buf_upload: null argument");
  const size_t row_bytes = (size_t)buf->cols * elem_size(buf->type);
  if (host_pitch == 0) host_pitch = row_bytes;
  if (host_pitch < row_bytes) return fail(GLSX_ERR_ARG, "glsx_buf_upload: host pitch smaller than a row");
  GLSX_HIP(hipMemcpy2DAsync(buf->ptr, row_bytes, host, host_pitch, row_bytes, buf->rows, hipMemcpyHostToDevice,
                            (hipStream_t)stream));
  // pageable host memory: the copy has consumed the source when the call returns
  GLSX_HIP(hipStreamSynchronize((hipStream_t)stream));
  return GLSX_OK;
}

GLSX_API int glsx_buf_download(glsx_buf buf, void* host, size_t host_pitch, void* stream) {
  if (!buf || !host) return fail(GLSX_ERR_ARG, "glsx_buf_download: null argument");
  const size_t row_bytes = (size_t)buf->cols * elem_size(buf->type);
  if (host_pitch == 0) host_pitch = row_bytes;
  if (host_pitch < row_bytes) return fail(GLSX_ERR_ARG, "glsx_buf_download: host pitch smaller than a row");
  GLSX_HIP(hipMemcpy2DAsync(host, host_pitch, buf->ptr, row_bytes, row_bytes, buf->rows, hipMemcpyDeviceToHost,
                            (hipStream_t)stream));
  GLSX_HIP(hipStreamSynchronize((hipStream_t)stream));
  return GLSX_OK;
}

// Pinned host memory and stream-ordered transfers: the asynchronous GlsMat
// upload/download the reference leaves disabled (its PBO paths,
// glsMat.cpp:55-56, 311-552).
GLSX_API int glsx_host_alloc(size_t bytes, void** out) {
  if (!out) return fail(GLSX_ERR_ARG, "glsx_host_alloc: null output");
  *out = nullptr;
  if (bytes == 0) return fail(GLSX_ERR_ARG, "glsx_host_alloc: zero bytes");
  if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return fail(GLSX_ERR_NOMEM, "glsx_host_alloc: pinned allocation failed");
  }
  return GLSX_OK;
}

GLSX_API int glsx_host_free(void* p) {
  if (!p) return GLSX_OK;
  GLSX_HIP(hipHostFree(p));
  return GLSX_OK;
}

GLSX_API int glsx_buf_upload_async(glsx_buf buf, const void* host, size_t host_pitch, void* stream) {
  if (!buf || !host) return fail(GLSX_ERR_ARG, "glsx_buf_upload_async: null argument");
  const size_t row_bytes = (size_t)buf->cols * elem_size(buf->type);
  if (host_pitch == 0) host_pitch = row_bytes;
  if (host_pitch < row_bytes) return fail(GLSX_ERR_ARG, "glsx_buf_upload_async: host pitch smaller than a row");
  GLSX_HIP(hipMemcpy2DAsync(buf->ptr, row_bytes, host, host_pitch, row_bytes, buf->rows, hipMemcpyHostToDevice,
                            (hipStream_t)stream));
  note_use(buf, (hipStream_t)stream);
  return GLSX_OK;
}

GLSX_API int glsx_buf_download_async(glsx_buf buf, void* host, size_t host_pitch, void* stream) {
  if (!buf || !host) return fail(GLSX_ERR_ARG, "glsx_buf_download_async: null argument");
  const size_t row_bytes = (size_t)buf->cols * elem_size(buf->type);
  if (host_pitch == 0) host_pitch = row_bytes;
  if (host_pitch < row_bytes) return fail(GLSX_ERR_ARG, "glsx_buf_download_async: host pitch smaller than a row");
  GLSX_HIP(hipMemcpy2DAsync(host, host_pitch, buf->ptr, row_bytes, row_bytes, buf->rows, hipMemcpyDeviceToHost,
                            (hipStream_t)stream));
  note_use(buf, (hipStream_t)stream);
  return GLSX_OK;
}

GLSX_API int glsx_stream_sync(void* stream) {
  GLSX_HIP(hipStreamSynchronize((hipStream_t)stream));
  return GLSX_OK;
}

GLSX_API int glsx_buf_copy(glsx_buf src, glsx_buf dst, void* stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "glsx_buf_copy: null buffer");
  if (src->rows != dst->rows || src->cols != dst->cols || src->type != dst->type)
    return fail(GLSX_ERR_ARG, "glsx_buf_copy: size or type mismatch");
  if (src == dst) return GLSX_OK;
  GLSX_HIP(hipMemcpyAsync(dst->ptr, src->ptr, src->bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  note_use(src, (hipStream_t)stream);
  note_use(dst, (hipStream_t)stream);
  return GLSX_OK;
}

// ---------------------------------------------------------------------------
// FFT
// ---------------------------------------------------------------------------
GLSX_API int glsx_fft2d_dev(const void* src, int src_type, void* dst, int rows, int cols, int flags,
                            void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  return fft2d_device(ds, src, src_type, dst, rows, cols, flags, (hipStream_t)stream);
}

GLSX_API int glsx_fft1d_dev(const void* src, int src_type, void* dst, int len, long long count,
                            long long elem_stride, long long batch_stride, int flags, void* stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "fft1d: null pointer");
  if (src_type != GLSX_CV_32FC2 && src_type != GLSX_CV_32FC1)
    return fail(GLSX_ERR_ARG, "fft1d: source must be CV_32FC1 or CV_32FC2");
  if (flags & ~(GLSX_FFT_INVERSE | GLSX_FFT_SCALE | GLSX_FFT_SHIFT | GLSX_FFT_SPLIT_MFMA))
    return fail(GLSX_ERR_ARG, "fft1d: unknown flag bits");
  if (int rc = validate_fft_shape(len, 1)) return rc;
  if (count < 0 || elem_stride < 1 || batch_stride < 0)
    return fail(GLSX_ERR_ARG, "fft1d: bad count or strides");
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  const bool inv = flags & GLSX_FFT_INVERSE;
  PassDesc p;
  p.L = len;
  p.col = elem_stride != 1;
  p.src = src;
  p.dst = dst;
  p.n_batch = count;
  p.es = elem_stride;
  p.bs = batch_stride;
  p.in_kind = src_type == GLSX_CV_32FC1 ? IN_R32 : IN_C32;
  p.out_kind = OUT_C32;
  p.conj_in = inv;
  p.conj_out = inv;
  p.shift = (flags & GLSX_FFT_SHIFT) != 0;
  p.scale = (flags & GLSX_FFT_SCALE) ? 1.0f / (float)len : 1.0f;
  if (flags & GLSX_FFT_SPLIT_MFMA) {
    if (int rc = split_mfma_use(ds, len, &p.split_mfma)) return rc;
  }
  return launch_pass(ds, p, (hipStream_t)stream);
}

GLSX_API int glsx_fft2d(glsx_buf src, glsx_buf dst, int flags, void* stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "fft: null buffer");
  if (src->type != GLSX_CV_32FC1 && src->type != GLSX_CV_32FC2)
    return fail(GLSX_ERR_ARG, "fft: source must be CV_32FC1 or CV_32FC2");
  if (dst->type != GLSX_CV_32FC2 || dst->rows != src->rows || dst->cols != src->cols)
    return fail(GLSX_ERR_ARG, "fft: destination must be CV_32FC2 of the source size");
  if (src->device != dst->device) return fail(GLSX_ERR_ARG, "fft: buffers on different devices");
  const int rc = glsx_fft2d_dev(src->ptr, src->type, dst->ptr, src->rows, src->cols, flags, stream);
  if (rc == GLSX_OK) {
    note_use(src, (hipStream_t)stream);
    note_use(dst, (hipStream_t)stream);
  }
  return rc;
}

// ---------------------------------------------------------------------------
// phase correlation
// ---------------------------------------------------------------------------
GLSX_API int glsx_phase_correlate_dev(const float* src1, const float* src2, const float* window, int rows,
                                      int cols, glsx_poc_result* out, void* stream) {
  if (!src1 || !src2 || !out) return fail(GLSX_ERR_ARG, "phaseCorrelate: null argument");
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (int rc = validate_fft_shape(rows, cols)) return rc;
  hipStream_t s = (hipStream_t)stream;
  {
    // the workspace is the stream's own scratch: calls on different streams
    // run concurrently; on one stream the lock keeps two host threads' chains
    // (and each chain's result copy) from interleaving.  Stream order then
    // keeps the next chain from overwriting the result before it is copied.
    std::unique_lock<std::mutex> lk = stream_lock(ds, s);
    const size_t ws_bytes = phase_correlate_workspace_bytes(ds, rows, cols);
    void* ws = nullptr;
    if (int rc = stream_scratch(ds, s, 2, ws_bytes + 256, &ws)) return rc;
    glsx_poc_result* dres = (glsx_poc_result*)((char*)ws + ws_bytes);
    if (int rc = phase_correlate_device(ds, src1, src2, window, rows, cols, ws, ws_bytes, dres, s)) return rc;
    GLSX_HIP(hipMemcpyAsync(out, dres, sizeof(glsx_poc_result), hipMemcpyDeviceToHost, s));
  }
  GLSX_HIP(hipStreamSynchronize(s));
  return GLSX_OK;
}

GLSX_API int glsx_phase_correlate(glsx_buf src1, glsx_buf src2, glsx_buf window, glsx_poc_result* out,
                                  void* stream) {
  if (!src1 || !src2 || !out) return fail(GLSX_ERR_ARG, "phaseCorrelate: null argument");
  // glsPhaseCorrelate.cpp:144-150
  if (src1->type != GLSX_CV_32FC1 || src2->type != GLSX_CV_32FC1)
    return fail(GLSX_ERR_ARG, "phaseCorrelate: inputs must be CV_32FC1");
  if (src1->rows != src2->rows || src1->cols != src2->cols)
    return fail(GLSX_ERR_ARG, "phaseCorrelate: inputs must have the same size");
  if (window && (window->type != GLSX_CV_32FC1 || window->rows != src1->rows || window->cols != src1->cols))
    return fail(GLSX_ERR_ARG, "phaseCorrelate: window must be CV_32FC1 of the input size");
  // (synchronous: every use has completed when the call returns)
  return glsx_phase_correlate_dev((const float*)src1->ptr, (const float*)src2->ptr,
                                  window ? (const float*)window->ptr : nullptr, src1->rows, src1->cols, out, stream);
}
