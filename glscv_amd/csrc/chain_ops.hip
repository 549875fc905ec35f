// chain_ops.hip -- the standalone operations of glsCV's phase-correlation
// chain (glsCV/glsPhaseCorrelate.cpp:174-210), for callers that compose the
// chain themselves or use the operations elsewhere:
//   gls::multiply(src0, src1, dst)           glsBasicOperation.cpp:265-272, shader case 2 (:169)
//   gls::mulSpectrums(src0, src1, dst, conj) glsBasicOperation.cpp:320-327, shader cases 6-7 (:182-193)
//   gls::mulSpectrumsPhaseOnly               glsBasicOperation.cpp:329-335, shader case 8 (:195-206)
//   gls::split / gls::merge                  glsSplit.cpp:192-210, glsMerge.cpp:193-211
//   gls::minMaxLoc                           glsMinMaxLoc.cpp:149-199, shader :65-118
// gls::phaseCorrelate itself runs them fused into the FFT passes; these are
// the one-operation-per-launch forms, each a single HBM sweep (minMaxLoc: a
// sweep plus one small folding launch).  Elementwise kernels move 16-byte
// vectors when every pointer is 16-byte aligned (glsx buffers always are) and
// fall back to scalar accesses otherwise; every form gives the same bits.
#include <algorithm>
#include <cfloat>
#include <initializer_list>
#include <string>

#include "fft_core.hpp"
#include "glsx_internal.hpp"

#define GLSX_API extern "C" __attribute__((visibility("default")))

namespace glsx {

namespace {

constexpr int kThreads = 256;

int grid_for(const DeviceState* ds, long long n) {
  const long long want = (n + kThreads - 1) / kThreads;
  return (int)std::max<long long>(1, std::min<long long>(want, (long long)ds->num_cu * 8));
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// [a, a + abytes) and [b, b + bbytes) share a byte
bool overlaps(const void* a, size_t abytes, const void* b, size_t bbytes) {
  const uintptr_t a0 = (uintptr_t)a, b0 = (uintptr_t)b;
  return a0 < b0 + bbytes && b0 < a0 + abytes;
}

// An elementwise destination may be one of its sources exactly (each element
// is read before it is written) but must not overlap one partly.
bool bad_alias(const void* d, const void* x, size_t bytes) { return d != x && overlaps(d, bytes, x, bytes); }

// ---- multiply: dst = src0 * src1 per float (channels are independent) -------
// dst may be src0 or src1 (each element is read before it is written).
template <bool VEC>
__global__ void __launch_bounds__(kThreads) multiply_kernel(const float* a, const float* b, float* d, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if constexpr (VEC) {
    const long long n4 = n >> 2;
    const auto* a4 = (const GLSX_GLOBAL float4*)a;
    const auto* b4 = (const GLSX_GLOBAL float4*)b;
    auto* d4 = (GLSX_GLOBAL float4*)d;
    for (long long i = t0; i < n4; i += stride) {
      const float4 x = a4[i], y = b4[i];
      d4[i] = make_float4(x.x * y.x, x.y * y.y, x.z * y.z, x.w * y.w);
    }
    for (long long i = (n4 << 2) + t0; i < n; i += stride) gptr(d)[i] = gptr(a)[i] * gptr(b)[i];
  } else {
    for (long long i = t0; i < n; i += stride) gptr(d)[i] = gptr(a)[i] * gptr(b)[i];
  }
}

// ---- spectrum products on interleaved complex (re, im) ----------------------
// MODE 0: a*b (case 6), 1: a*conj(b) (case 7), 2: a*conj(b)/|a*conj(b)|
// (case 8).  The shader divides by sqrt(re^2 + im^2) without a guard (0/0 is
// NaN); here a magnitude that is not above 1e-30 (zero, or NaN) gives 0.  The
// arithmetic is the shader's, operation for operation (no contraction), with
// IEEE division and square root.
template <int MODE>
__device__ __forceinline__ float2 spec_op(float ar, float ai, float br, float bi) {
  if constexpr (MODE == 0) {
    return make_float2(ar * br - ai * bi, ar * bi + ai * br);
  } else {
    const float re = ar * br + ai * bi;
    const float im = -ar * bi + ai * br;
    if constexpr (MODE == 1) {
      return make_float2(re, im);
    } else {
      const float mag = sqrtf(re * re + im * im);
      if (mag > 1e-30f) return make_float2(re / mag, im / mag);
      return make_float2(0.0f, 0.0f);
    }
  }
}

template <int MODE, bool VEC>
__global__ void __launch_bounds__(kThreads) spectrum_product_kernel(const float2* a, const float2* b, float2* d,
                                                                    long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long done = 0;
  if constexpr (VEC) {
    const long long n2 = n >> 1;  // two complex values per 16 bytes
    const auto* a4 = (const GLSX_GLOBAL float4*)a;
    const auto* b4 = (const GLSX_GLOBAL float4*)b;
    auto* d4 = (GLSX_GLOBAL float4*)d;
    for (long long i = t0; i < n2; i += stride) {
      const float4 x = a4[i], y = b4[i];
      const float2 p = spec_op<MODE>(x.x, x.y, y.x, y.y);
      const float2 q = spec_op<MODE>(x.z, x.w, y.z, y.w);
      d4[i] = make_float4(p.x, p.y, q.x, q.y);
    }
    done = n2 << 1;
  }
  for (long long i = done + t0; i < n; i += stride) {
    const float2 x = gptr(a)[i], y = gptr(b)[i];
    gptr(d)[i] = spec_op<MODE>(x.x, x.y, y.x, y.y);
  }
}

// ---- split / merge of a two-channel float image ---------------------------
// split: p0[i] = src[2i], p1[i] = src[2i+1] (p1 may be null: channel 0 only).
template <bool VEC>
__global__ void __launch_bounds__(kThreads) split2_kernel(const float* s, float* p0, float* p1, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long done = 0;
  if constexpr (VEC) {
    // four pixels per thread: two 16-byte loads, one 16-byte store per plane
    const long long n4 = n >> 2;
    const auto* s4 = (const GLSX_GLOBAL float4*)s;
    for (long long i = t0; i < n4; i += stride) {
      const float4 x = s4[2 * i], y = s4[2 * i + 1];
      ((GLSX_GLOBAL float4*)p0)[i] = make_float4(x.x, x.z, y.x, y.z);
      if (p1) ((GLSX_GLOBAL float4*)p1)[i] = make_float4(x.y, x.w, y.y, y.w);
    }
    done = n4 << 2;
  }
  for (long long i = done + t0; i < n; i += stride) {
    gptr(p0)[i] = gptr(s)[2 * i];
    if (p1) gptr(p1)[i] = gptr(s)[2 * i + 1];
  }
}

// merge: dst[2i] = p0[i], dst[2i+1] = p1[i].
template <bool VEC>
__global__ void __launch_bounds__(kThreads) merge2_kernel(const float* p0, const float* p1, float* d, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long done = 0;
  if constexpr (VEC) {
    const long long n4 = n >> 2;
    auto* d4 = (GLSX_GLOBAL float4*)d;
    for (long long i = t0; i < n4; i += stride) {
      const float4 x = ((const GLSX_GLOBAL float4*)p0)[i], y = ((const GLSX_GLOBAL float4*)p1)[i];
      d4[2 * i] = make_float4(x.x, y.x, x.y, y.y);
      d4[2 * i + 1] = make_float4(x.z, y.z, x.w, y.w);
    }
    done = n4 << 2;
  }
  for (long long i = done + t0; i < n; i += stride) {
    gptr(d)[2 * i] = gptr(p0)[i];
    gptr(d)[2 * i + 1] = gptr(p1)[i];
  }
}

// ---- minMaxLoc --------------------------------------------------------------
// The reference scans each row for its first strict minimum and maximum
// (starting from FLT_MAX / -FLT_MAX, location 0), then the row results in row
// order the same way: the result is the first extremum in row-major order,
// NaN never compares and so is never chosen, and when no value passes the
// start value the result is (FLT_MAX or -FLT_MAX, location (0, 0)).  Here
// every thread scans its elements in increasing index order with the same
// strict comparisons and start values, and the (value, index) pairs fold by
// "more extreme value, then lower index": the start pairs (+-FLT_MAX, 0)
// carry index 0, so the fold reproduces the start-value cases too.
struct MinMaxPart {
  float minv, maxv;
  long long mini, maxi;
};

__device__ __forceinline__ void mm_scan(MinMaxPart& p, float v, long long i) {
  if (v < p.minv) {
    p.minv = v;
    p.mini = i;
  }
  if (v > p.maxv) {
    p.maxv = v;
    p.maxi = i;
  }
}

__device__ __forceinline__ void mm_fold(MinMaxPart& p, float minv, long long mini, float maxv, long long maxi) {
  if (minv < p.minv || (minv == p.minv && mini < p.mini)) {
    p.minv = minv;
    p.mini = mini;
  }
  if (maxv > p.maxv || (maxv == p.maxv && maxi < p.maxi)) {
    p.maxv = maxv;
    p.maxi = maxi;
  }
}

// fold over the block (64-lane shuffles, then the waves through LDS); the
// result is valid in thread 0
__device__ MinMaxPart mm_block(MinMaxPart p) {
  __shared__ MinMaxPart sh[kThreads / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float minv = __shfl_xor(p.minv, off), maxv = __shfl_xor(p.maxv, off);
    const long long mini = __shfl_xor(p.mini, off), maxi = __shfl_xor(p.maxi, off);
    mm_fold(p, minv, mini, maxv, maxi);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = p;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) mm_fold(p, sh[k].minv, sh[k].mini, sh[k].maxv, sh[k].maxi);
  }
  return p;
}

// CN channels per pixel, channel 0 compared (the shader reads .r).
template <int CN, bool VEC>
__global__ void __launch_bounds__(kThreads) minmax_partials_kernel(const float* s, long long n,
                                                                   MinMaxPart* __restrict__ part) {
  MinMaxPart p{FLT_MAX, -FLT_MAX, 0, 0};
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long done = 0;
  if constexpr (VEC) {
    constexpr int PX = 4 / CN;  // pixels per 16 bytes
    const long long nv = n / PX;
    const auto* s4 = (const GLSX_GLOBAL float4*)s;
    for (long long i = t0; i < nv; i += stride) {
      const float4 x = s4[i];
      mm_scan(p, x.x, i * PX);
      if constexpr (CN == 1) {
        mm_scan(p, x.y, i * PX + 1);
        mm_scan(p, x.z, i * PX + 2);
        mm_scan(p, x.w, i * PX + 3);
      } else {
        mm_scan(p, x.z, i * PX + 1);
      }
    }
    done = nv * PX;
  }
  // the remaining pixels have higher indices than every vector one, so each
  // thread still visits its pixels in increasing order
  for (long long i = done + t0; i < n; i += stride) mm_scan(p, gptr(s)[i * CN], i);
  p = mm_block(p);
  if (threadIdx.x == 0) part[blockIdx.x] = p;
}

__global__ void __launch_bounds__(kThreads) minmax_final_kernel(const MinMaxPart* __restrict__ part, int n_part,
                                                                long long cols, glsx_minmax_result* out) {
  MinMaxPart p{FLT_MAX, -FLT_MAX, 0, 0};
  for (int k = threadIdx.x; k < n_part; k += blockDim.x) mm_fold(p, part[k].minv, part[k].mini, part[k].maxv, part[k].maxi);
  p = mm_block(p);
  if (threadIdx.x == 0) {
    glsx_minmax_result r;
    r.min_val = p.minv;
    r.max_val = p.maxv;
    r.min_x = (int)(p.mini % cols);
    r.min_y = (int)(p.mini / cols);
    r.max_x = (int)(p.maxi % cols);
    r.max_y = (int)(p.maxi / cols);
    *out = r;
  }
}

int launch_error(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(GLSX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return GLSX_OK;
}

int multiply_device(DeviceState* ds, const float* a, const float* b, float* d, long long n, hipStream_t s) {
  if (!a || !b || !d) return fail(GLSX_ERR_ARG, "multiply: null pointer");
  if ((uintptr_t)a % 4 || (uintptr_t)b % 4 || (uintptr_t)d % 4)
    return fail(GLSX_ERR_ARG, "multiply: pointers must be 4-byte aligned");
  if (n <= 0) return GLSX_OK;
  if (bad_alias(d, a, 4 * (size_t)n) || bad_alias(d, b, 4 * (size_t)n))
    return fail(GLSX_ERR_ARG, "multiply: the destination overlaps a source partly");
  const bool vec = aligned16(a) && aligned16(b) && aligned16(d);
  const int g = grid_for(ds, vec ? (n + 3) / 4 : n);
  if (vec) hipLaunchKernelGGL(multiply_kernel<true>, dim3(g), dim3(kThreads), 0, s, a, b, d, n);
  else hipLaunchKernelGGL(multiply_kernel<false>, dim3(g), dim3(kThreads), 0, s, a, b, d, n);
  return launch_error("multiply");
}

template <int MODE>
int spectrum_product_launch(DeviceState* ds, const float2* a, const float2* b, float2* d, long long n, hipStream_t s) {
  const bool vec = aligned16(a) && aligned16(b) && aligned16(d);
  const int g = grid_for(ds, vec ? (n + 1) / 2 : n);
  if (vec) hipLaunchKernelGGL((spectrum_product_kernel<MODE, true>), dim3(g), dim3(kThreads), 0, s, a, b, d, n);
  else hipLaunchKernelGGL((spectrum_product_kernel<MODE, false>), dim3(g), dim3(kThreads), 0, s, a, b, d, n);
  return launch_error("mulSpectrums");
}

int spectrum_product_device(DeviceState* ds, const float2* a, const float2* b, float2* d, long long n, int mode,
                            hipStream_t s) {
  if (!a || !b || !d) return fail(GLSX_ERR_ARG, "mulSpectrums: null pointer");
  if (mode < GLSX_SPECTRUM_MUL || mode > GLSX_SPECTRUM_PHASE_ONLY)
    return fail(GLSX_ERR_ARG, "mulSpectrums: mode must be GLSX_SPECTRUM_MUL, _MUL_CONJ or _PHASE_ONLY");
  if ((uintptr_t)a % 8 || (uintptr_t)b % 8 || (uintptr_t)d % 8)
    return fail(GLSX_ERR_ARG, "mulSpectrums: pointers must be 8-byte aligned");
  if (n <= 0) return GLSX_OK;
  if (bad_alias(d, a, 8 * (size_t)n) || bad_alias(d, b, 8 * (size_t)n))
    return fail(GLSX_ERR_ARG, "mulSpectrums: the destination overlaps a source partly");
  if (mode == GLSX_SPECTRUM_MUL) return spectrum_product_launch<0>(ds, a, b, d, n, s);
  if (mode == GLSX_SPECTRUM_MUL_CONJ) return spectrum_product_launch<1>(ds, a, b, d, n, s);
  return spectrum_product_launch<2>(ds, a, b, d, n, s);
}

int split_device(DeviceState* ds, const float* src, float* p0, float* p1, long long n, hipStream_t s) {
  if (!src || !p0) return fail(GLSX_ERR_ARG, "split: null pointer");
  if ((uintptr_t)src % 4 || (uintptr_t)p0 % 4 || (uintptr_t)p1 % 4)
    return fail(GLSX_ERR_ARG, "split: pointers must be 4-byte aligned");
  if (n <= 0) return GLSX_OK;
  const size_t pb = 4 * (size_t)n, sb = 8 * (size_t)n;
  if (overlaps(p0, pb, src, sb) || (p1 && (overlaps(p1, pb, src, sb) || overlaps(p0, pb, p1, pb))))
    return fail(GLSX_ERR_ARG, "split: the planes must not overlap the source or each other");
  const bool vec = aligned16(src) && aligned16(p0) && (!p1 || aligned16(p1));
  const int g = grid_for(ds, vec ? (n + 3) / 4 : n);
  if (vec) hipLaunchKernelGGL(split2_kernel<true>, dim3(g), dim3(kThreads), 0, s, src, p0, p1, n);
  else hipLaunchKernelGGL(split2_kernel<false>, dim3(g), dim3(kThreads), 0, s, src, p0, p1, n);
  return launch_error("split");
}

int merge_device(DeviceState* ds, const float* p0, const float* p1, float* dst, long long n, hipStream_t s) {
  if (!p0 || !p1 || !dst) return fail(GLSX_ERR_ARG, "merge: null pointer");
  if ((uintptr_t)p0 % 4 || (uintptr_t)p1 % 4 || (uintptr_t)dst % 4)
    return fail(GLSX_ERR_ARG, "merge: pointers must be 4-byte aligned");
  if (n <= 0) return GLSX_OK;
  if (overlaps(dst, 8 * (size_t)n, p0, 4 * (size_t)n) || overlaps(dst, 8 * (size_t)n, p1, 4 * (size_t)n))
    return fail(GLSX_ERR_ARG, "merge: the destination must not overlap the planes");
  const bool vec = aligned16(p0) && aligned16(p1) && aligned16(dst);
  const int g = grid_for(ds, vec ? (n + 3) / 4 : n);
  if (vec) hipLaunchKernelGGL(merge2_kernel<true>, dim3(g), dim3(kThreads), 0, s, p0, p1, dst, n);
  else hipLaunchKernelGGL(merge2_kernel<false>, dim3(g), dim3(kThreads), 0, s, p0, p1, dst, n);
  return launch_error("merge");
}

// Partials and the result live in the stream's scratch slot 3 (held under
// the stream lock by the caller while this enqueues).
int min_max_loc_device(DeviceState* ds, const float* src, int channels, int rows, int cols,
                       glsx_minmax_result* dev_out, hipStream_t s) {
  if (!src || !dev_out) return fail(GLSX_ERR_ARG, "minMaxLoc: null pointer");
  if (channels != 1 && channels != 2) return fail(GLSX_ERR_ARG, "minMaxLoc: 1 or 2 channels");
  if (rows <= 0 || cols <= 0) return fail(GLSX_ERR_ARG, "minMaxLoc: empty image");
  if ((uintptr_t)src % 4 || (uintptr_t)dev_out % 4) return fail(GLSX_ERR_ARG, "minMaxLoc: pointers must be 4-byte aligned");
  const long long n = (long long)rows * cols;
  const bool vec = aligned16(src);
  const int g = (int)std::min<long long>(grid_for(ds, vec ? (n * channels + 3) / 4 : n), 1024);
  MinMaxPart* part = nullptr;
  if (int rc = stream_scratch(ds, s, 3, sizeof(MinMaxPart) * g, (void**)&part)) return rc;
  if (channels == 1) {
    if (vec) hipLaunchKernelGGL((minmax_partials_kernel<1, true>), dim3(g), dim3(kThreads), 0, s, src, n, part);
    else hipLaunchKernelGGL((minmax_partials_kernel<1, false>), dim3(g), dim3(kThreads), 0, s, src, n, part);
  } else {
    if (vec) hipLaunchKernelGGL((minmax_partials_kernel<2, true>), dim3(g), dim3(kThreads), 0, s, src, n, part);
    else hipLaunchKernelGGL((minmax_partials_kernel<2, false>), dim3(g), dim3(kThreads), 0, s, src, n, part);
  }
  if (int rc = launch_error("minMaxLoc")) return rc;
  hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(kThreads), 0, s, part, g, (long long)cols, dev_out);
  return launch_error("minMaxLoc");
}

bool float_type(int type) { return type == GLSX_CV_32FC1 || type == GLSX_CV_32FC2; }

int check_same(glsx_buf a, glsx_buf b, const char* what) {
  if (a->rows != b->rows || a->cols != b->cols || a->type != b->type)
    return fail(GLSX_ERR_ARG, std::string(what) + ": size or type mismatch");
  return GLSX_OK;
}

void note_all(std::initializer_list<glsx_buf> bufs, hipStream_t s) {
  for (glsx_buf b : bufs)
    if (b) note_use(b, s);
}

}  // namespace

}  // namespace glsx

using namespace glsx;

// ---- multiply -----------------------------------------------------------------
GLSX_API int glsx_multiply_dev(const float* src0, const float* src1, float* dst, long long n, void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (n < 0) return fail(GLSX_ERR_ARG, "multiply: negative size");
  return multiply_device(ds, src0, src1, dst, n, (hipStream_t)stream);
}

GLSX_API int glsx_multiply(glsx_buf src0, glsx_buf src1, glsx_buf dst, void* stream) {
  if (!src0 || !src1 || !dst) return fail(GLSX_ERR_ARG, "multiply: null buffer");
  if (!float_type(src0->type)) return fail(GLSX_ERR_ARG, "multiply: sources must be CV_32FC1 or CV_32FC2");
  if (int rc = check_same(src0, src1, "multiply")) return rc;
  if (int rc = check_same(src0, dst, "multiply")) return rc;
  const long long n = (long long)src0->rows * src0->cols * type_channels(src0->type);
  const int rc = glsx_multiply_dev((const float*)src0->ptr, (const float*)src1->ptr, (float*)dst->ptr, n, stream);
  if (rc == GLSX_OK) note_all({src0, src1, dst}, (hipStream_t)stream);
  return rc;
}

// ---- mulSpectrums / mulSpectrumsPhaseOnly ------------------------------------------
GLSX_API int glsx_mul_spectrums_dev(const void* src0, const void* src1, void* dst, long long n, int mode,
                                    void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (n < 0) return fail(GLSX_ERR_ARG, "mulSpectrums: negative size");
  return spectrum_product_device(ds, (const float2*)src0, (const float2*)src1, (float2*)dst, n, mode,
                                 (hipStream_t)stream);
}

static int spectrum_product_buf(glsx_buf src0, glsx_buf src1, glsx_buf dst, int mode, void* stream,
                                const char* what) {
  if (!src0 || !src1 || !dst) return fail(GLSX_ERR_ARG, std::string(what) + ": null buffer");
  if (src0->type != GLSX_CV_32FC2) return fail(GLSX_ERR_ARG, std::string(what) + ": sources must be CV_32FC2");
  if (int rc = check_same(src0, src1, what)) return rc;
  if (int rc = check_same(src0, dst, what)) return rc;
  const int rc = glsx_mul_spectrums_dev(src0->ptr, src1->ptr, dst->ptr, (long long)src0->rows * src0->cols, mode,
                                        stream);
  if (rc == GLSX_OK) note_all({src0, src1, dst}, (hipStream_t)stream);
  return rc;
}

GLSX_API int glsx_mul_spectrums(glsx_buf src0, glsx_buf src1, glsx_buf dst, int conj, void* stream) {
  return spectrum_product_buf(src0, src1, dst, conj ? GLSX_SPECTRUM_MUL_CONJ : GLSX_SPECTRUM_MUL, stream,
                              "mulSpectrums");
}

GLSX_API int glsx_mul_spectrums_phase_only(glsx_buf src0, glsx_buf src1, glsx_buf dst, void* stream) {
  return spectrum_product_buf(src0, src1, dst, GLSX_SPECTRUM_PHASE_ONLY, stream, "mulSpectrumsPhaseOnly");
}

// ---- split / merge ---------------------------------------------------------------
GLSX_API int glsx_split_dev(const float* src, float* plane0, float* plane1, long long n, void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (n < 0) return fail(GLSX_ERR_ARG, "split: negative size");
  return split_device(ds, src, plane0, plane1, n, (hipStream_t)stream);
}

GLSX_API int glsx_merge_dev(const float* plane0, const float* plane1, float* dst, long long n, void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (n < 0) return fail(GLSX_ERR_ARG, "merge: negative size");
  return merge_device(ds, plane0, plane1, dst, n, (hipStream_t)stream);
}

GLSX_API int glsx_split(glsx_buf src, const glsx_buf* planes, int n_planes, void* stream) {
  if (!src || !planes) return fail(GLSX_ERR_ARG, "split: null buffer");
  if (!float_type(src->type)) return fail(GLSX_ERR_ARG, "split: source must be CV_32FC1 or CV_32FC2");
  const int cn = type_channels(src->type);
  if (n_planes != cn) return fail(GLSX_ERR_ARG, "split: one plane per channel");
  for (int c = 0; c < cn; ++c) {
    if (!planes[c]) return fail(GLSX_ERR_ARG, "split: null plane");
    if (planes[c]->type != GLSX_CV_32FC1 || planes[c]->rows != src->rows || planes[c]->cols != src->cols)
      return fail(GLSX_ERR_ARG, "split: planes must be CV_32FC1 of the source's size");
  }
  const long long n = (long long)src->rows * src->cols;
  int rc;
  if (cn == 1) {
    rc = planes[0] == src ? GLSX_OK : glsx_buf_copy(src, planes[0], stream);
  } else {
    if (planes[0] == planes[1]) return fail(GLSX_ERR_ARG, "split: the two planes must be distinct");
    rc = glsx_split_dev((const float*)src->ptr, (float*)planes[0]->ptr, (float*)planes[1]->ptr, n, stream);
    if (rc == GLSX_OK) note_all({src, planes[0], planes[1]}, (hipStream_t)stream);
  }
  return rc;
}

GLSX_API int glsx_merge(const glsx_buf* planes, int n_planes, glsx_buf dst, void* stream) {
  if (!planes || !dst) return fail(GLSX_ERR_ARG, "merge: null buffer");
  if (n_planes != 1 && n_planes != 2) return fail(GLSX_ERR_ARG, "merge: 1 or 2 planes");
  for (int c = 0; c < n_planes; ++c) {
    if (!planes[c]) return fail(GLSX_ERR_ARG, "merge: null plane");
    if (planes[c]->type != GLSX_CV_32FC1 || planes[c]->rows != planes[0]->rows || planes[c]->cols != planes[0]->cols)
      return fail(GLSX_ERR_ARG, "merge: planes must be CV_32FC1 of one size");
  }
  const int want = n_planes == 1 ? GLSX_CV_32FC1 : GLSX_CV_32FC2;
  if (dst->type != want || dst->rows != planes[0]->rows || dst->cols != planes[0]->cols)
    return fail(GLSX_ERR_ARG, "merge: destination must have the planes' size and their count of channels");
  if (n_planes == 1) return planes[0] == dst ? GLSX_OK : glsx_buf_copy(planes[0], dst, stream);
  const long long n = (long long)dst->rows * dst->cols;
  const int rc = glsx_merge_dev((const float*)planes[0]->ptr, (const float*)planes[1]->ptr, (float*)dst->ptr, n, stream);
  if (rc == GLSX_OK) note_all({planes[0], planes[1], dst}, (hipStream_t)stream);
  return rc;
}

// ---- minMaxLoc ---------------------------------------------------------------------
GLSX_API int glsx_min_max_loc_dev(const void* src, int src_type, int rows, int cols, glsx_minmax_result* dev_out,
                                  void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (!float_type(src_type)) return fail(GLSX_ERR_ARG, "minMaxLoc: source must be CV_32FC1 or CV_32FC2");
  // the partials live in the stream's scratch: one call at a time per stream
  std::unique_lock<std::mutex> lk = stream_lock(ds, (hipStream_t)stream);
  return min_max_loc_device(ds, (const float*)src, type_channels(src_type), rows, cols, dev_out, (hipStream_t)stream);
}

GLSX_API int glsx_min_max_loc(glsx_buf src, glsx_minmax_result* out, void* stream) {
  if (!src || !out) return fail(GLSX_ERR_ARG, "minMaxLoc: null argument");
  if (!float_type(src->type)) return fail(GLSX_ERR_ARG, "minMaxLoc: source must be CV_32FC1 or CV_32FC2");
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  const hipStream_t s = (hipStream_t)stream;
  std::unique_lock<std::mutex> lk = stream_lock(ds, s);
  // the result goes to the end of slot 4 of the stream's scratch, then to the host
  void* res = nullptr;
  if (int rc = stream_scratch(ds, s, 4, sizeof(glsx_minmax_result), &res)) return rc;
  if (int rc = min_max_loc_device(ds, (const float*)src->ptr, type_channels(src->type), src->rows, src->cols,
                                  (glsx_minmax_result*)res, s))
    return rc;
  note_use(src, s);
  GLSX_HIP(hipMemcpyAsync(out, res, sizeof(glsx_minmax_result), hipMemcpyDeviceToHost, s));
  GLSX_HIP(hipStreamSynchronize(s));
  return GLSX_OK;
}
