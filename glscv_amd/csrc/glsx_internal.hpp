// glsx_internal.hpp -- shared declarations of the libglsx implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/glsx.h"

namespace glsx {

// ---- error handling ---------------------------------------------------------
// Every entry point returns a glsx_status; the message of the last failure on
// the calling thread is kept for glsx_last_error().
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);

#define GLSX_HIP(call)                                                               \
  do {                                                                               \
    hipError_t _e = (call);                                                          \
    if (_e != hipSuccess)                                                            \
      return ::glsx::fail(GLSX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

// ---- kernel attributes ----------------------------------------------------------
// A kernel's dynamic LDS limit is an attribute of the function on the device
// current at the call.  One LdsLimit per kernel raises it once per device (a
// bit per device id; two threads may both set it, with the same value).
class LdsLimit {
 public:
  int raise(const void* kern, size_t bytes) {
    int dev = 0;
    GLSX_HIP(hipGetDevice(&dev));
    const uint64_t bit = dev >= 0 && dev < 64 ? 1ull << dev : 0;
    if (bit && (done_.load(std::memory_order_acquire) & bit)) return GLSX_OK;
    GLSX_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done_.fetch_or(bit, std::memory_order_release);
    return GLSX_OK;
  }

 private:
  std::atomic<uint64_t> done_{0};
};

// ---- cv type codes (OpenCV's CV_MAKETYPE encoding) ---------------------------
inline int type_depth(int type) { return type & 7; }
inline int type_channels(int type) { return (type >> 3) + 1; }
inline int depth_size(int depth) {
  static const int sz[8] = {1, 1, 2, 2, 4, 4, 8, 2};
  return sz[depth & 7];
}
inline size_t elem_size(int type) { return (size_t)depth_size(type_depth(type)) * type_channels(type); }

// ---- per-device state ---------------------------------------------------------
struct DeviceState {
  int id = -1;
  int num_cu = 0;
  size_t lds_per_block = 0;
  float2* twiddles = nullptr;   // concatenated tables exp(-2*pi*i*k/P), P = 2..16384
  // plans and servers alive on this device: they hold the twiddle pointer,
  // so glsx_shutdown refuses while this is non-zero
  std::atomic<int> live_objects{0};
  // device address of this device's copy of the allocation-free entry
  // point's static workspace (hipGetSymbolAddress at initialisation)
  void* static_ws = nullptr;
  // split-precision MFMA passes, per transform length (index log2 L):
  // 0 not checked yet, 1 within the error bound, -1 not within it
  std::mutex mfma_mutex;
  int mfma_state[15] = {};
  double mfma_err[15] = {};
  // per-stream scratch buffers (stream, slot) -> (pointer, bytes), kept until shutdown
  std::mutex scratch_mutex;
  std::map<std::pair<void*, int>, std::pair<void*, size_t>> scratch;
  // one mutex per stream: held while a call enqueues work that uses the
  // stream's scratch, so chains of two host threads on one stream never
  // interleave (guarded by scratch_mutex; nodes of a map never move)
  std::map<void*, std::mutex> stream_locks;
  // device blocks of released GlsMat buffers, reused by later glsx_buf_create
  // calls of the same size once `ready` (an event recorded on the stream that
  // last used the buffer; nullptr = reusable now) has completed.  hipFree
  // synchronises the whole device; the cache keeps it off the release path.
  struct CachedBlock {
    void* ptr;
    size_t bytes;
    hipEvent_t ready;
  };
  std::mutex cache_mutex;
  std::vector<CachedBlock> cache;
  size_t cached_bytes = 0;
  std::vector<hipEvent_t> spare_events;
};

// Most bytes of released blocks a device keeps for reuse; a release beyond it
// waits for the buffer's last stream and frees the block.
constexpr size_t kBlockCacheCap = (size_t)4 << 30;

// Returns the state of the current device, initialising it on first use.
DeviceState* device_state(int* status);
// The same without initialising: GLSX_ERR_STATE when glsx_init (or a first
// call) has not set the current device up.
DeviceState* device_state_if_ready(int* status);
// The state of device `dev` (0..15) whether or not it is initialised.
DeviceState* device_slot(int dev);

// Lock of `stream` on this device (see DeviceState::stream_locks).
std::unique_lock<std::mutex> stream_lock(DeviceState* ds, hipStream_t stream);

// Scratch owned by one stream: work queued on `stream` may use it without
// racing other streams.  Grown (after synchronising the stream) when too
// small, kept until glsx_shutdown.  Slots let one call chain hold several.
int stream_scratch(DeviceState* ds, hipStream_t stream, int slot, size_t bytes, void** out);

// ---- FFT driver (fft_kernels.hip) --------------------------------------------
int init_twiddles(DeviceState* ds);

// One pass over a batch of 1D transforms, see fft_pass.hpp.
struct PassDesc {
  int L = 0;
  bool col = false;
  const void* src = nullptr;
  const void* src2 = nullptr;
  const float* win = nullptr;
  void* dst = nullptr;
  int2* argmax = nullptr;
  long long n_batch = 0, es = 1, bs = 0, out_cols = 0;
  long long es_out = 0;     // output element stride of a pair column tile, 0 = es
  int in_kind = 0, out_kind = 0;
  int conj_in = 0, conj_out = 0, shift = 0;
  float scale = 1.0f;
  float mag_offset = 0.0f;  // OUT_LOGMAG
  bool split_mfma = false;  // radix-16 stages on MFMA (3-term BF16 split), OUT_C32 only
};
// Number of workgroups the pass would launch (for sizing argmax partials).
long long pass_grid(const DeviceState* ds, const PassDesc& p);
int launch_pass(const DeviceState* ds, const PassDesc& p, hipStream_t stream);

// 2D (or 1D) FFT with gls::fft semantics on device pointers.
//   src_type: GLSX_CV_32FC2 (interleaved complex) or GLSX_CV_32FC1 (real)
//   dst: interleaved complex rows x cols; src == dst allowed.
int fft2d_device(DeviceState* ds, const void* src, int src_type, void* dst, int rows, int cols, int flags,
                 hipStream_t stream);
int validate_fft_shape(int rows, int cols);

// ---- real-input 2D FFT through half-size spectra (fft_r2c.hip) -------------------
// Used for CV_32FC1 sources where it measured faster than the complex
// transform (rows, cols >= 4096; not with GLSX_FFT_SPLIT_MFMA): r2c_applies for
// complex output, r2c_applies_mag for the magnitude planes.
bool r2c_applies(int rows, int cols, int src_type, int flags);
bool r2c_applies_mag(int rows, int cols, int src_type, int flags);
// dst: M x N complex (out_kind OUT_C32) or a float plane (OUT_MAG / OUT_LOGMAG,
// optional (max, min) partials, *n_partials of them).  zbuf: M x N complex of
// scratch for the half spectra, or null to keep them in dst's even rows
// (OUT_C32 without SHIFT only).  src must not overlap dst.  half (magnitude
// planes only): instead of dst, the magnitudes of columns 0..cols/2 alone,
// unshifted, row r at half + r * hpitch (the other columns are their mirrors).
int fft2d_r2c(DeviceState* ds, const float* src, void* dst, int rows, int cols, int flags, int out_kind,
              float mag_offset, int2* partials, long long* n_partials, void* zbuf, hipStream_t stream,
              float* half = nullptr, long long hpitch = 0);
// Workgroups (= partials) of the real-input column pass.
long long r2c_partials(const DeviceState* ds, int rows, int cols);

// ---- spectrum display chain (spectrum.hip) ---------------------------------------
// 2D FFT whose last pass stores |z| or log(|z| + offset) as float (dst rows x
// cols CV_32FC1); the intermediate spectrum lives in a stream-ordered buffer.
int fft2d_mag_device(DeviceState* ds, const void* src, int src_type, float* dst, int rows, int cols, int flags,
                     int log_mag, float offset, hipStream_t stream);

// ---- split-precision MFMA passes (split_mfma.hip) ------------------------------
// FP64 check of length L (run once per device and length, then kept).
int split_mfma_check(DeviceState* ds, int L, double* rel_rms, int* enabled);
// Should a pass of length L honour GLSX_FFT_SPLIT_MFMA?  Runs the check on
// first use; a length outside the bound (or L < 16) answers false.
int split_mfma_use(DeviceState* ds, int L, bool* use);

// ---- phase correlation (poc.hip) ---------------------------------------------
int phase_correlate_device(DeviceState* ds, const float* a, const float* b, const float* win, int rows,
                           int cols, void* workspace, size_t workspace_bytes, glsx_poc_result* dev_result,
                           hipStream_t stream);
size_t phase_correlate_workspace_bytes(const DeviceState* ds, int rows, int cols);
// Bytes of the static workspace of glsx_phase_correlate_dev_noalloc.
constexpr size_t kStaticWorkspaceBytes = GLSX_NOALLOC_WORKSPACE_BYTES;
// Looks up this device's copy of that workspace (device current).
int init_static_workspace(DeviceState* ds);

}  // namespace glsx

// opaque GlsMat storage handle
struct glsx_buf_s {
  std::atomic<int> refs{1};
  int rows = 0, cols = 0, type = 0;
  int device = 0;
  void* ptr = nullptr;
  size_t bytes = 0;  // rows * cols * elem_size
  size_t block = 0;  // bytes of the device block behind ptr (bytes rounded up to 256)
  // What the release must wait for: the stream of the library calls that
  // used the buffer.  `mixed` once calls used two different streams,
  // `exported` once the device pointer was handed out (glsx_buf_info): the
  // library then cannot know which work still reads the buffer, and the last
  // glsx_buf_release waits for the device.
  // On a stream other than the null stream, `ev` is recorded right after
  // each use (the stream exists then; it may be gone by the release); uses
  // on the null stream are covered by an event recorded at the release.
  std::mutex use_mutex;
  hipStream_t last_stream = nullptr;
  hipEvent_t ev = nullptr;
  bool used = false, mixed = false, exported = false;
};

namespace glsx {
// Records that a library call has just enqueued work on `stream` that uses
// `b` (call it after the enqueue).
void note_use(glsx_buf_s* b, hipStream_t stream);
}  // namespace glsx
