// spectrum.hip -- the FFT display chain of glsCam (glsCam/glsCam.cpp:196-202):
//   gls::fft(frame, frame, GLS_FFT_SHIFT)
//   gls::logMagSpectrums(frame, frame, 1.0)      glsBasicOperation.cpp:345-351
//   gls::normalize(frame, frame, 0, 1, NORM_MINMAX)   glsNormalize.cpp:52-75
// The reference runs the FFT's 14 sweeps, one more sweep for the log
// magnitude, gls::minMaxLoc (two passes and a download), then a multiply and
// an add sweep.  Here:
//   * magSpectrums / logMagSpectrums are one elementwise kernel (8 B in, 4 B
//     out per pixel), and fft2d_mag_device fuses them into the FFT's column
//     pass store (OUT_MAG / OUT_LOGMAG), so the complex spectrum is written
//     and read once instead of twice;
//   * normalize is two launches: per-block partials (min/max, or the L1/L2
//     sums in FP64) and one apply kernel whose blocks each fold the partials
//     and then scale their share of the plane.
#include <algorithm>

#include "fft_pass.hpp"
#include "glsx_internal.hpp"

#define GLSX_API extern "C" __attribute__((visibility("default")))

namespace glsx {

namespace {

constexpr int kThreads = 256;

// The elementwise kernels move 16 bytes per access where both pointers are
// 16-byte aligned, with kUnroll accesses of a thread in flight before its
// stores, and stride over the plane with at most 8 workgroups per CU; the
// rest (unaligned pointers, the last n % 4 elements) goes element by element.
// 4096 x 4096 minmax map in place after the view's column pass: 23.7 us,
// against 27.1 us with one 4-byte access per thread and iteration.  Source
// and destination may be the same plane: a thread
// writes only elements it has read.
constexpr int kUnroll = 4;

__device__ __forceinline__ bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// dst[i] = f(src[i]) for n floats
template <class F>
__device__ __forceinline__ void map_floats(const float* src, float* dst, long long n, F f) {
  const long long nth = (long long)gridDim.x * blockDim.x;
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long head = 0;
  if (aligned16(src) && aligned16(dst)) {
    const long long n4 = n >> 2;
    const auto* s = (const float4*)src;
    auto* d = (float4*)dst;
    auto g = [&](float4 x) { return make_float4(f(x.x), f(x.y), f(x.z), f(x.w)); };
    long long i = tid;
    for (; i + (kUnroll - 1) * nth < n4; i += kUnroll * nth) {
      float4 x[kUnroll];
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) x[k] = s[i + k * nth];
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) d[i + k * nth] = g(x[k]);
    }
    for (; i < n4; i += nth) d[i] = g(s[i]);
    head = n4 << 2;
  }
  for (long long i = head + tid; i < n; i += nth) dst[i] = f(src[i]);
}

// f(src[i]) for n floats, every element once
template <class F>
__device__ __forceinline__ void visit_floats(const float* src, long long n, F f) {
  const long long nth = (long long)gridDim.x * blockDim.x;
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long head = 0;
  if (aligned16(src)) {
    const long long n4 = n >> 2;
    const auto* s = (const float4*)src;
    auto g = [&](float4 x) { f(x.x); f(x.y); f(x.z); f(x.w); };
    long long i = tid;
    for (; i + (kUnroll - 1) * nth < n4; i += kUnroll * nth) {
      float4 x[kUnroll];
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) x[k] = s[i + k * nth];
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) g(x[k]);
    }
    for (; i < n4; i += nth) g(s[i]);
    head = n4 << 2;
  }
  for (long long i = head + tid; i < n; i += nth) f(src[i]);
}

// magSpectrums / logMagSpectrums: 4 complex values in (two 16-byte loads), 4
// floats out
__global__ void __launch_bounds__(kThreads) spectrum_kernel(const float2* __restrict__ src, float* __restrict__ dst,
                                                            long long n, int log_mag, float offset) {
  const long long nth = (long long)gridDim.x * blockDim.x;
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const bool lm = log_mag != 0;
  long long head = 0;
  if (aligned16(src) && aligned16(dst)) {
    const long long n4 = n >> 2;
    const auto* s = (const float4*)src;
    auto* d = (float4*)dst;
    for (long long i = tid; i < n4; i += nth) {
      const float4 x = s[2 * i], y = s[2 * i + 1];
      d[i] = make_float4(spectrum_value(make_float2(x.x, x.y), lm, offset),
                         spectrum_value(make_float2(x.z, x.w), lm, offset),
                         spectrum_value(make_float2(y.x, y.y), lm, offset),
                         spectrum_value(make_float2(y.z, y.w), lm, offset));
    }
    head = n4 << 2;
  }
  for (long long i = head + tid; i < n; i += nth) dst[i] = spectrum_value(src[i], lm, offset);
}

struct NormPartial {
  double a, b;  // NORM_MINMAX: min, max; NORM_L1: sum |x|; NORM_L2: sum x^2
};

__device__ __forceinline__ void fold(NormPartial& p, const NormPartial& q, int type) {
  if (type == GLSX_NORM_MINMAX) {
    p.a = fmin(p.a, q.a);
    p.b = fmax(p.b, q.b);
  } else {
    p.a += q.a;
  }
}

__device__ __forceinline__ NormPartial identity(int type) {
  return type == GLSX_NORM_MINMAX ? NormPartial{INFINITY, -INFINITY} : NormPartial{0.0, 0.0};
}

// tree over the block in LDS; every thread returns the block's value
__device__ NormPartial block_fold(NormPartial v, int type, NormPartial* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      NormPartial p = sh[threadIdx.x];
      fold(p, sh[threadIdx.x + s], type);
      sh[threadIdx.x] = p;
    }
    __syncthreads();
  }
  const NormPartial r = sh[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kThreads) norm_partials_kernel(const float* __restrict__ src, long long n, int type,
                                                                 NormPartial* __restrict__ part) {
  __shared__ NormPartial sh[kThreads];
  NormPartial v = identity(type);
  if (type == GLSX_NORM_MINMAX) {
    visit_floats(src, n, [&](float x) {
      v.a = fmin(v.a, (double)x);  // fmin/fmax skip NaN
      v.b = fmax(v.b, (double)x);
    });
  } else if (type == GLSX_NORM_L1) {
    visit_floats(src, n, [&](float x) { v.a += fabs((double)x); });
  } else {
    visit_floats(src, n, [&](float x) { v.a += (double)x * (double)x; });
  }
  v = block_fold(v, type, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Every block folds all partials (a few hundred), derives the affine map of
// glsNormalize.cpp:56-71 and applies it to its share of the plane:
//   NORM_MINMAX: scl = (float)(beta - alpha) / (float)(max - min),
//                dst = scl*src + (float)(alpha - scl*min)  (0 scale if max == min)
//   NORM_L1/L2:  scl = (float)alpha / (float)norm, dst = scl*src
__global__ void __launch_bounds__(kThreads) normalize_apply_kernel(const float* src,
                                                                   float* dst, long long n, int type,
                                                                   const NormPartial* __restrict__ part, int n_part,
                                                                   double alpha, double beta) {
  __shared__ NormPartial sh[kThreads];
  __shared__ float s_scl, s_off;
  NormPartial v = identity(type);
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) fold(v, part[i], type);
  v = block_fold(v, type, sh);
  if (threadIdx.x == 0) {
    float scl, off = 0.0f;
    if (type == GLSX_NORM_MINMAX) {
      scl = v.b > v.a ? (float)(beta - alpha) / (float)(v.b - v.a) : 0.0f;
      off = (float)(alpha - (double)scl * v.a);
    } else {
      const double norm = type == GLSX_NORM_L2 ? sqrt(v.a) : v.a;
      scl = norm > 0.0 ? (float)alpha / (float)norm : 0.0f;
    }
    s_scl = scl;
    s_off = off;
  }
  __syncthreads();
  const float scl = s_scl, off = s_off;
  if (type == GLSX_NORM_MINMAX) map_floats(src, dst, n, [&](float x) { return scl * x + off; });
  else map_floats(src, dst, n, [&](float x) { return scl * x; });
}

// The glsCam view's normalisation from the column pass's per-workgroup
// (max, min) partials: same map as normalize_apply_kernel (NORM_MINMAX).
__global__ void __launch_bounds__(kThreads) minmax_apply_kernel(float* dst, long long n,
                                                                const int2* __restrict__ part, int n_part,
                                                                double alpha, double beta) {
  __shared__ NormPartial sh[kThreads];
  __shared__ float s_scl, s_off;
  NormPartial v = identity(GLSX_NORM_MINMAX);
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) {
    const int2 p = part[i];
    fold(v, NormPartial{(double)__int_as_float(p.y), (double)__int_as_float(p.x)}, GLSX_NORM_MINMAX);
  }
  v = block_fold(v, GLSX_NORM_MINMAX, sh);
  if (threadIdx.x == 0) {
    const float scl = v.b > v.a ? (float)(beta - alpha) / (float)(v.b - v.a) : 0.0f;
    s_scl = scl;
    s_off = (float)(alpha - (double)scl * v.a);
  }
  __syncthreads();
  const float scl = s_scl, off = s_off;
  map_floats(dst, dst, n, [&](float x) { return scl * x + off; });
}

// The glsCam view's last pass after a real frame's column pass, which stored
// the log magnitudes of columns 0..N/2 only, unshifted (half plane h, row
// pitch hp): the plane is Hermitian, |Y(-r, -c)| = |Y(r, c)|, so column c >
// N/2 of row r is h[-r][N-c].  Workgroup w takes the rows r = w and rm = -w
// (mod M, 0 <= w <= M/2) and writes both output rows, normalized with the
// map of minmax_apply_kernel and moved by SHIFT: each h element is read once
// and stored twice (8 + 4 bytes per pixel of the view instead of 4 + 4 + 8),
// every store a 16-byte aligned group.  N >= 8: the 16-byte groups of columns
// k..k+3 (k a multiple of 4) do not wrap under the shift by N/2.
__global__ void __launch_bounds__(kThreads) view_expand_kernel(const float* __restrict__ h, long long hp,
                                                               float* __restrict__ dst, int M, int N, int shift,
                                                               const int2* __restrict__ part, int n_part,
                                                               double alpha, double beta) {
  __shared__ NormPartial sh[kThreads];
  __shared__ float s_scl, s_off;
  NormPartial v = identity(GLSX_NORM_MINMAX);
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) {
    const int2 p = part[i];
    fold(v, NormPartial{(double)__int_as_float(p.y), (double)__int_as_float(p.x)}, GLSX_NORM_MINMAX);
  }
  v = block_fold(v, GLSX_NORM_MINMAX, sh);
  if (threadIdx.x == 0) {
    const float scl = v.b > v.a ? (float)(beta - alpha) / (float)(v.b - v.a) : 0.0f;
    s_scl = scl;
    s_off = (float)(alpha - (double)scl * v.a);
  }
  __syncthreads();
  const float scl = s_scl, off = s_off;
  auto f = [&](float x) { return scl * x + off; };
  const int r = blockIdx.x, rm = (M - r) & (M - 1);
  const bool two = rm != r;  // rows 0 and M/2 are their own mirrors
  const int hm = shift ? M / 2 : 0, hn = shift ? N / 2 : 0, half_n = N / 2;
  float* o_r = dst + (long long)((r + hm) & (M - 1)) * N;
  float* o_rm = dst + (long long)((rm + hm) & (M - 1)) * N;
  const float* h_r = h + (long long)r * hp;
  const float* h_rm = h + (long long)rm * hp;
  for (int k = 4 * threadIdx.x; k < half_n; k += 4 * blockDim.x) {
    const float4 a = *(const float4*)(h_r + k), b = *(const float4*)(h_rm + k);
    // the next group's first column (k + 4 <= N/2): the next lane's a.x and
    // b.x; a wave's last lane, and a lane whose successor left the loop, load it
    // (4096^2: 18.0 against 18.3-18.4 us with a load in every lane)
    float a4 = __shfl_down(a.x, 1), b4 = __shfl_down(b.x, 1);
    if ((threadIdx.x & 63) == 63 || k + 4 >= half_n) {
      a4 = h_r[k + 4];
      b4 = h_rm[k + 4];
    }
    const int ck = (k + hn) & (N - 1);
    *(float4*)(o_r + ck) = make_float4(f(a.x), f(a.y), f(a.z), f(a.w));
    if (two) *(float4*)(o_rm + ck) = make_float4(f(b.x), f(b.y), f(b.z), f(b.w));
    // column N - c of row rm is column c of row r, and the other way round:
    // columns N-k-4 .. N-k-1 (a 16-byte aligned group) mirror columns k+4 ..
    // k+1.  The groups of k = 0, 4, .., N/2 - 4 cover N/2 .. N-1; column N/2
    // mirrors itself (row -r of it holds the same magnitude as row r, bit for
    // bit: its values are each other's conjugates), and column 0 is its own.
    const int cm = (N - k - 4 + hn) & (N - 1);
    *(float4*)(o_rm + cm) = make_float4(f(a4), f(a.w), f(a.z), f(a.y));
    if (two) *(float4*)(o_r + cm) = make_float4(f(b4), f(b.w), f(b.z), f(b.y));
  }
}

int grid_for(const DeviceState* ds, long long n) {
  const long long want = (n + kThreads * 4 * kUnroll - 1) / (kThreads * 4 * kUnroll);
  return (int)std::max<long long>(1, std::min<long long>(want, (long long)ds->num_cu * 8));
}

int spectrum_device(DeviceState* ds, const float2* src, float* dst, long long n, int log_mag, float offset,
                    hipStream_t stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "magSpectrums: null pointer");
  if (n <= 0) return GLSX_OK;
  hipLaunchKernelGGL(spectrum_kernel, dim3(grid_for(ds, n)), dim3(kThreads), 0, stream, src, dst, n, log_mag, offset);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

int normalize_device(DeviceState* ds, const float* src, float* dst, long long n, double alpha, double beta,
                     int type, hipStream_t stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "normalize: null pointer");
  if (type != GLSX_NORM_MINMAX && type != GLSX_NORM_L1 && type != GLSX_NORM_L2)
    return fail(GLSX_ERR_ARG, "normalize: norm type must be NORM_MINMAX, NORM_L1 or NORM_L2");
  if (n <= 0) return GLSX_OK;
  const int g = grid_for(ds, n);
  NormPartial* part = nullptr;
  if (int rc = stream_scratch(ds, stream, 1, sizeof(NormPartial) * g, (void**)&part)) return rc;
  hipLaunchKernelGGL(norm_partials_kernel, dim3(g), dim3(kThreads), 0, stream, src, n, type, part);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(normalize_apply_kernel, dim3(g), dim3(kThreads), 0, stream, src, dst, n, type, part, g,
                       alpha, beta);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(GLSX_ERR_HIP, std::string("normalize: ") + hipGetErrorString(e));
  return GLSX_OK;
}

}  // namespace

namespace {

// fft + (log) magnitude; with `partials`, the last pass also writes per-
// workgroup (max, min) partials of the stored values (*n_partials of them).
int fft2d_mag_impl(DeviceState* ds, const void* src, int src_type, float* dst, int rows, int cols, int flags,
                   int log_mag, float offset, int2* partials, long long* n_partials, hipStream_t stream);

}  // namespace

int fft2d_mag_device(DeviceState* ds, const void* src, int src_type, float* dst, int rows, int cols, int flags,
                     int log_mag, float offset, hipStream_t stream) {
  return fft2d_mag_impl(ds, src, src_type, dst, rows, cols, flags, log_mag, offset, nullptr, nullptr, stream);
}

int spectrum_view_device(DeviceState* ds, const void* src, int src_type, float* dst, int rows, int cols, int flags,
                         float offset, double alpha, double beta, hipStream_t stream) {
  if (int rc = validate_fft_shape(rows, cols)) return rc;
  // partial count of the last pass (at most one per column or row tile)
  const long long max_part = std::max(rows, cols);
  int2* part = nullptr;
  if (int rc = stream_scratch(ds, stream, 1, sizeof(int2) * max_part, (void**)&part)) return rc;
  long long n_part = 0;
  if (r2c_applies_mag(rows, cols, src_type, flags) && (uintptr_t)dst % 16 == 0) {
    // real frame: the column pass stores the half plane, the last pass
    // expands it with 16-byte stores (the half spectra and the half plane in
    // stream scratch)
    float2* tmp = nullptr;
    float* half = nullptr;
    const long long hp = cols / 2 + 4;  // 16-byte aligned rows
    if (int rc = stream_scratch(ds, stream, 0, (size_t)rows * cols * sizeof(float2), (void**)&tmp)) return rc;
    if (int rc = stream_scratch(ds, stream, 5, (size_t)rows * hp * sizeof(float), (void**)&half)) return rc;
    if (int rc = fft2d_r2c(ds, (const float*)src, dst, rows, cols, flags, OUT_LOGMAG, offset, part, &n_part, tmp,
                           stream, half, hp))
      return rc;
    hipLaunchKernelGGL(view_expand_kernel, dim3(rows / 2 + 1), dim3(kThreads), 0, stream, half, hp, (float*)dst,
                       rows, cols, (flags & GLSX_FFT_SHIFT) ? 1 : 0, part, (int)n_part, alpha, beta);
    GLSX_HIP(hipGetLastError());
    return GLSX_OK;
  }
  if (int rc = fft2d_mag_impl(ds, src, src_type, dst, rows, cols, flags, 1, offset, part, &n_part, stream))
    return rc;
  const long long n = (long long)rows * cols;
  hipLaunchKernelGGL(minmax_apply_kernel, dim3(grid_for(ds, n)), dim3(kThreads), 0, stream, dst, n, part,
                     (int)n_part, alpha, beta);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

namespace {

int fft2d_mag_impl(DeviceState* ds, const void* src, int src_type, float* dst, int rows, int cols, int flags,
                   int log_mag, float offset, int2* partials, long long* n_partials, hipStream_t stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "fft mag: null pointer");
  if (src_type != GLSX_CV_32FC2 && src_type != GLSX_CV_32FC1)
    return fail(GLSX_ERR_ARG, "fft mag: source must be CV_32FC1 or CV_32FC2");
  if (flags & ~(GLSX_FFT_INVERSE | GLSX_FFT_SCALE | GLSX_FFT_SHIFT | GLSX_FFT_SPLIT_MFMA))
    return fail(GLSX_ERR_ARG, "fft mag: unknown flag bits");
  if (int rc = validate_fft_shape(rows, cols)) return rc;
  const bool inv = flags & GLSX_FFT_INVERSE, scl = flags & GLSX_FFT_SCALE, shf = flags & GLSX_FFT_SHIFT;
  const int in_kind = src_type == GLSX_CV_32FC1 ? IN_R32 : IN_C32;
  const int out_kind = log_mag ? OUT_LOGMAG : OUT_MAG;
  if (rows == 1 || cols == 1) {
    PassDesc p;
    p.L = rows * cols;
    p.src = src;
    p.dst = dst;
    p.n_batch = 1;
    p.es = 1;
    p.bs = p.L;
    p.in_kind = in_kind;
    p.out_kind = out_kind;
    p.conj_in = inv;
    p.conj_out = inv;
    p.shift = shf;
    p.scale = scl ? 1.0f / (float)p.L : 1.0f;
    p.mag_offset = offset;
    p.argmax = partials;
    if (n_partials) *n_partials = pass_grid(ds, p);
    return launch_pass(ds, p, stream);
  }
  bool mfma_rows = false;
  if (flags & GLSX_FFT_SPLIT_MFMA) {
    if (int rc = split_mfma_use(ds, cols, &mfma_rows)) return rc;
  }
  // the complex spectrum between the passes: the stream's scratch
  float2* tmp = nullptr;
  if (int rc = stream_scratch(ds, stream, 0, (size_t)rows * cols * sizeof(float2), (void**)&tmp)) return rc;
  if (r2c_applies_mag(rows, cols, src_type, flags) && (uintptr_t)dst % 8 == 0) {
    // real frame: half-size spectra in the scratch, magnitudes (and partials)
    // stored by the real-input column pass (8-byte stores of column pairs; a
    // plane aligned to 4 bytes only takes the row and column passes below)
    return fft2d_r2c(ds, (const float*)src, dst, rows, cols, flags, out_kind, offset, partials, n_partials, tmp,
                     stream);
  }
  int rc = GLSX_OK;
  {
    PassDesc r;
    r.L = cols;
    r.src = src;
    r.dst = tmp;
    r.n_batch = rows;
    r.es = 1;
    r.bs = cols;
    r.in_kind = in_kind;
    r.out_kind = OUT_C32;
    r.conj_in = inv;
    r.shift = shf;
    r.scale = scl ? 1.0f / (float)cols : 1.0f;
    r.split_mfma = mfma_rows;
    rc = launch_pass(ds, r, stream);
  }
  if (!rc) {
    // the magnitude store runs in the FP32 column kernels
    PassDesc c;
    c.L = rows;
    c.col = true;
    c.src = tmp;
    c.dst = dst;
    c.n_batch = cols;
    c.es = cols;
    c.bs = 1;
    c.in_kind = IN_C32;
    c.out_kind = out_kind;
    c.conj_out = inv;
    c.shift = shf;
    c.scale = scl ? 1.0f / (float)rows : 1.0f;
    c.mag_offset = offset;
    c.argmax = partials;
    if (n_partials) *n_partials = pass_grid(ds, c);
    rc = launch_pass(ds, c, stream);
  }
  return rc;
}

}  // namespace

}  // namespace glsx

using namespace glsx;

namespace {

int check_plane(glsx_buf b, int type, const char* what) {
  if (!b) return fail(GLSX_ERR_ARG, std::string(what) + ": null buffer");
  if (b->type != type)
    return fail(GLSX_ERR_ARG, std::string(what) + (type == GLSX_CV_32FC2 ? ": expected CV_32FC2" : ": expected CV_32FC1"));
  return GLSX_OK;
}

}  // namespace

GLSX_API int glsx_spectrum_dev(const void* src, void* dst, long long n, int log_mag, float offset, void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (n < 0) return fail(GLSX_ERR_ARG, "magSpectrums: negative size");
  return spectrum_device(ds, (const float2*)src, (float*)dst, n, log_mag, offset, (hipStream_t)stream);
}

GLSX_API int glsx_mag_spectrums(glsx_buf src, glsx_buf dst, void* stream) {
  if (int rc = check_plane(src, GLSX_CV_32FC2, "magSpectrums")) return rc;
  if (int rc = check_plane(dst, GLSX_CV_32FC1, "magSpectrums")) return rc;
  if (dst->rows != src->rows || dst->cols != src->cols) return fail(GLSX_ERR_ARG, "magSpectrums: size mismatch");
  const int rc = glsx_spectrum_dev(src->ptr, dst->ptr, (long long)src->rows * src->cols, 0, 0.0f, stream);
  if (rc == GLSX_OK) {
    note_use(src, (hipStream_t)stream);
    note_use(dst, (hipStream_t)stream);
  }
  return rc;
}

GLSX_API int glsx_log_mag_spectrums(glsx_buf src, glsx_buf dst, float offset, void* stream) {
  if (int rc = check_plane(src, GLSX_CV_32FC2, "logMagSpectrums")) return rc;
  if (int rc = check_plane(dst, GLSX_CV_32FC1, "logMagSpectrums")) return rc;
  if (dst->rows != src->rows || dst->cols != src->cols) return fail(GLSX_ERR_ARG, "logMagSpectrums: size mismatch");
  const int rc = glsx_spectrum_dev(src->ptr, dst->ptr, (long long)src->rows * src->cols, 1, offset, stream);
  if (rc == GLSX_OK) {
    note_use(src, (hipStream_t)stream);
    note_use(dst, (hipStream_t)stream);
  }
  return rc;
}

GLSX_API int glsx_normalize_dev(const float* src, float* dst, long long n, double alpha, double beta, int norm_type,
                                void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  if (n < 0) return fail(GLSX_ERR_ARG, "normalize: negative size");
  // the partials live in the stream's scratch: one chain at a time per stream
  std::unique_lock<std::mutex> lk = stream_lock(ds, (hipStream_t)stream);
  return normalize_device(ds, src, dst, n, alpha, beta, norm_type, (hipStream_t)stream);
}

GLSX_API int glsx_normalize(glsx_buf src, glsx_buf dst, double alpha, double beta, int norm_type, void* stream) {
  if (int rc = check_plane(src, GLSX_CV_32FC1, "normalize")) return rc;
  if (int rc = check_plane(dst, GLSX_CV_32FC1, "normalize")) return rc;
  if (dst->rows != src->rows || dst->cols != src->cols) return fail(GLSX_ERR_ARG, "normalize: size mismatch");
  const int rc = glsx_normalize_dev((const float*)src->ptr, (float*)dst->ptr, (long long)src->rows * src->cols, alpha, beta,
                            norm_type, stream);
  if (rc == GLSX_OK) {
    note_use(src, (hipStream_t)stream);
    note_use(dst, (hipStream_t)stream);
  }
  return rc;
}

GLSX_API int glsx_fft2d_mag_dev(const void* src, int src_type, float* dst, int rows, int cols, int flags, int log_mag,
                                float offset, void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  std::unique_lock<std::mutex> lk = stream_lock(ds, (hipStream_t)stream);
  return fft2d_mag_device(ds, src, src_type, dst, rows, cols, flags, log_mag, offset, (hipStream_t)stream);
}

GLSX_API int glsx_fft2d_mag(glsx_buf src, glsx_buf dst, int flags, int log_mag, float offset, void* stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "fft mag: null buffer");
  if (src->type != GLSX_CV_32FC1 && src->type != GLSX_CV_32FC2)
    return fail(GLSX_ERR_ARG, "fft mag: source must be CV_32FC1 or CV_32FC2");
  if (int rc = check_plane(dst, GLSX_CV_32FC1, "fft mag")) return rc;
  if (dst->rows != src->rows || dst->cols != src->cols) return fail(GLSX_ERR_ARG, "fft mag: size mismatch");
  if (src->device != dst->device) return fail(GLSX_ERR_ARG, "fft mag: buffers on different devices");
  const int rc = glsx_fft2d_mag_dev(src->ptr, src->type, (float*)dst->ptr, src->rows, src->cols, flags, log_mag, offset,
                            stream);
  if (rc == GLSX_OK) {
    note_use(src, (hipStream_t)stream);
    note_use(dst, (hipStream_t)stream);
  }
  return rc;
}

GLSX_API int glsx_spectrum_view_dev(const void* src, int src_type, float* dst, int rows, int cols, int flags,
                                    float offset, double alpha, double beta, void* stream) {
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  std::unique_lock<std::mutex> lk = stream_lock(ds, (hipStream_t)stream);
  return spectrum_view_device(ds, src, src_type, dst, rows, cols, flags, offset, alpha, beta, (hipStream_t)stream);
}

GLSX_API int glsx_spectrum_view(glsx_buf src, glsx_buf dst, int flags, float offset, double alpha, double beta,
                                void* stream) {
  if (!src || !dst) return fail(GLSX_ERR_ARG, "spectrum view: null buffer");
  if (src->type != GLSX_CV_32FC1 && src->type != GLSX_CV_32FC2)
    return fail(GLSX_ERR_ARG, "spectrum view: source must be CV_32FC1 or CV_32FC2");
  if (int rc = check_plane(dst, GLSX_CV_32FC1, "spectrum view")) return rc;
  if (dst->rows != src->rows || dst->cols != src->cols) return fail(GLSX_ERR_ARG, "spectrum view: size mismatch");
  if (src->device != dst->device) return fail(GLSX_ERR_ARG, "spectrum view: buffers on different devices");
  const int rc = glsx_spectrum_view_dev(src->ptr, src->type, (float*)dst->ptr, src->rows, src->cols, flags, offset, alpha,
                                beta, stream);
  if (rc == GLSX_OK) {
    note_use(src, (hipStream_t)stream);
    note_use(dst, (hipStream_t)stream);
  }
  return rc;
}
