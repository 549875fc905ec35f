// split_mfma.hip -- the FP64 check that gates the split-precision MFMA passes.
//
// The MFMA passes (mfma_fft.hpp) trade the FP32 butterflies for 3-term BF16
// products.  A transform length uses them only after this check: a batch of
// pseudo-random complex inputs (uniform in [-1, 1)) goes through the MFMA pass
// and through a direct DFT computed in FP64 on the device,
//   y[q] = sum_m x[m] * exp(-2*pi*i*((q*m) mod L)/L),
// and the relative RMS error sqrt(sum |y32 - y64|^2 / sum |y64|^2) must be
// within GLSX_SPLIT_MFMA_TOL.  The result is kept per device and length.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fft_pass.hpp"
#include "glsx_internal.hpp"

#define GLSX_API extern "C" __attribute__((visibility("default")))

namespace glsx {
namespace {

__device__ __forceinline__ unsigned mix32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__global__ void check_input_kernel(float2* x, long long n, unsigned seed) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned a = mix32((unsigned)(2 * i) ^ seed), b = mix32((unsigned)(2 * i + 1) ^ seed);
  constexpr float k = 2.0f / 16777216.0f;  // 24 random bits -> [-1, 1)
  x[i] = make_float2((float)(a >> 8) * k - 1.0f, (float)(b >> 8) * k - 1.0f);
}

__global__ void roots64_kernel(double2* w, int L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  double s, c;
  sincospi(-2.0 * (double)i / (double)L, &s, &c);
  w[i] = make_double2(c, s);
}

// one output bin per thread; the root index q*m mod L is stepped exactly
__global__ void dft64_kernel(const float2* x, const double2* w, double2* y, int L, int B) {
  const long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (gid >= (long long)L * B) return;
  const int b = (int)(gid / L), q = (int)(gid % L);
  const float2* xb = x + (long long)b * L;
  double re = 0.0, im = 0.0;
  int idx = 0;
  for (int m = 0; m < L; ++m) {
    const float2 v = xb[m];
    const double2 t = w[idx];
    re = fma((double)v.x, t.x, fma(-(double)v.y, t.y, re));
    im = fma((double)v.x, t.y, fma((double)v.y, t.x, im));
    idx = (idx + q) & (L - 1);
  }
  y[gid] = make_double2(re, im);
}

int log2_of(int L) {
  int r = 0;
  while ((1 << r) < L) ++r;
  return r;
}

int run_check(DeviceState* ds, int L, double* err) {
  const int B = std::max(4, 65536 / L);
  const long long n = (long long)L * B;
  float2 *x = nullptr, *y32 = nullptr;
  double2 *w = nullptr, *y64 = nullptr;
  auto release = [&] {
    if (x) (void)hipFree(x);
    if (y32) (void)hipFree(y32);
    if (w) (void)hipFree(w);
    if (y64) (void)hipFree(y64);
  };
  int rc = GLSX_OK;
  do {
    if (hipMalloc(&x, n * sizeof(float2)) != hipSuccess || hipMalloc(&y32, n * sizeof(float2)) != hipSuccess ||
        hipMalloc(&w, L * sizeof(double2)) != hipSuccess || hipMalloc(&y64, n * sizeof(double2)) != hipSuccess) {
      rc = fail(GLSX_ERR_NOMEM, "split MFMA check: device allocation failed");
      break;
    }
    hipLaunchKernelGGL(check_input_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, x, n, 0x5eedu + L);
    hipLaunchKernelGGL(roots64_kernel, dim3((L + 255) / 256), dim3(256), 0, 0, w, L);
    hipLaunchKernelGGL(dft64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, x, w, y64, L, B);
    if (hipGetLastError() != hipSuccess) {
      rc = fail(GLSX_ERR_HIP, "split MFMA check: kernel launch failed");
      break;
    }
    PassDesc p;
    p.L = L;
    p.col = false;
    p.src = x;
    p.dst = y32;
    p.n_batch = B;
    p.es = 1;
    p.bs = L;
    p.in_kind = IN_C32;
    p.out_kind = OUT_C32;
    p.split_mfma = true;
    if ((rc = launch_pass(ds, p, 0))) break;
    std::vector<float2> h32(n);
    std::vector<double2> h64(n);
    if (hipMemcpy(h32.data(), y32, n * sizeof(float2), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h64.data(), y64, n * sizeof(double2), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(GLSX_ERR_HIP, "split MFMA check: copy failed");
      break;
    }
    double num = 0.0, den = 0.0;
    for (long long i = 0; i < n; ++i) {
      const double dr = (double)h32[i].x - h64[i].x, di = (double)h32[i].y - h64[i].y;
      num += dr * dr + di * di;
      den += h64[i].x * h64[i].x + h64[i].y * h64[i].y;
    }
    *err = den > 0.0 ? std::sqrt(num / den) : 0.0;
    if (!std::isfinite(*err)) *err = INFINITY;
  } while (false);
  release();
  return rc;
}

}  // namespace

int split_mfma_check(DeviceState* ds, int L, double* rel_rms, int* enabled) {
  if (int rc = validate_fft_shape(L, 1)) return rc;
  if (L < 16) {
    // no radix-16 stage: the FP32 kernels always run
    if (rel_rms) *rel_rms = 0.0;
    if (enabled) *enabled = 0;
    return GLSX_OK;
  }
  const int lg = log2_of(L);
  std::lock_guard<std::mutex> lk(ds->mfma_mutex);
  if (ds->mfma_state[lg] == 0) {
    double err = INFINITY;
    if (int rc = run_check(ds, L, &err)) return rc;
    ds->mfma_err[lg] = err;
    ds->mfma_state[lg] = err <= GLSX_SPLIT_MFMA_TOL ? 1 : -1;
  }
  if (rel_rms) *rel_rms = ds->mfma_err[lg];
  if (enabled) *enabled = ds->mfma_state[lg] == 1;
  return GLSX_OK;
}

int split_mfma_use(DeviceState* ds, int L, bool* use) {
  int enabled = 0;
  *use = false;
  if (int rc = split_mfma_check(ds, L, nullptr, &enabled)) return rc;
  *use = enabled == 1;
  return GLSX_OK;
}

}  // namespace glsx

GLSX_API int glsx_fft_split_mfma_check(int len, double* rel_rms, int* enabled) {
  using namespace glsx;
  int st = 0;
  DeviceState* ds = device_state(&st);
  if (!ds) return st;
  return split_mfma_check(ds, len, rel_rms, enabled);
}
