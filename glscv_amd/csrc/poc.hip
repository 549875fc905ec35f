// poc.hip -- gls::phaseCorrelate on the device.
//
// glsCV/glsPhaseCorrelate.cpp:142-231 runs: window multiply x2, fft x2,
// mulSpectrumsPhaseOnly, fft(INVERSE|SHIFT), split, minMaxLoc (2 passes + a
// download), then downloads the whole surface for a CPU 5x5 centroid.  Here
// the chain is four FFT passes plus one tiny kernel, and only the 40-byte
// result leaves the device:
//   1. row pass  img1 (real) * window      -> spec1      (merge + multiply fused)
//   2. col pass  spec1 in place
//   3. row pass  img2 (real) * window      -> spec2
//   4. col pass  spec2 in place
//   5. row pass  POC(spec1, spec2), conj   -> spec1      (mulSpectrumsPhaseOnly fused)
//   6. col pass  spec1 -> surface (real part, fft-shifted, unscaled) + argmax
//      partials per workgroup                             (split + minMaxLoc pass 0)
//   7. finalize: first maximum over the partials, 5x5 FP64 weighted centroid
//      with border clamping, response / (rows*cols), center - centroid.
#include "fft_pass.hpp"
#include "glsx_internal.hpp"
#include "poc_final.hpp"

namespace glsx {

namespace {

struct Layout {
  size_t spec1 = 0, spec2 = 0, surface = 0, partials = 0, total = 0;
  long long n_partials = 0;
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

Layout layout(const DeviceState* ds, int rows, int cols) {
  Layout l;
  const size_t n = (size_t)rows * cols;
  PassDesc c;
  c.L = rows;
  c.col = true;
  c.n_batch = cols;
  l.n_partials = pass_grid(ds, c);
  l.spec1 = 0;
  l.spec2 = align_up(l.spec1 + n * sizeof(float2));
  l.surface = align_up(l.spec2 + n * sizeof(float2));
  l.partials = align_up(l.surface + n * sizeof(float));
  l.total = align_up(l.partials + (size_t)l.n_partials * sizeof(int2));
  return l;
}

__global__ void __launch_bounds__(256) poc_finalize_kernel(PocFinalArgs f) {
  __shared__ float sv[16];
  __shared__ int si[16];
  poc_finalize_block(f, sv, si);
}

}  // namespace

size_t phase_correlate_workspace_bytes(const DeviceState* ds, int rows, int cols) {
  return layout(ds, rows, cols).total;
}

int phase_correlate_device(DeviceState* ds, const float* a, const float* b, const float* win, int rows, int cols,
                           void* workspace, size_t workspace_bytes, glsx_poc_result* dev_result,
                           hipStream_t stream) {
  if (!a || !b || !workspace || !dev_result) return fail(GLSX_ERR_ARG, "phaseCorrelate: null pointer");
  if (int rc = validate_fft_shape(rows, cols)) return rc;
  if (rows < 2 || cols < 2) return fail(GLSX_ERR_UNSUPPORTED, "phaseCorrelate: images must be at least 2x2");
  const Layout l = layout(ds, rows, cols);
  if (workspace_bytes < l.total) return fail(GLSX_ERR_ARG, "phaseCorrelate: workspace too small");
  char* ws = (char*)workspace;
  float2* spec1 = (float2*)(ws + l.spec1);
  float2* spec2 = (float2*)(ws + l.spec2);
  float* surface = (float*)(ws + l.surface);
  int2* partials = (int2*)(ws + l.partials);

  PassDesc r;
  r.L = cols;
  r.col = false;
  r.n_batch = rows;
  r.es = 1;
  r.bs = cols;
  r.out_kind = OUT_C32;
  r.win = win;
  r.in_kind = win ? IN_R32W : IN_R32;
  PassDesc c;
  c.L = rows;
  c.col = true;
  c.n_batch = cols;
  c.es = cols;
  c.bs = 1;
  c.in_kind = IN_C32;
  c.out_kind = OUT_C32;

  // forward spectra
  r.src = a;
  r.dst = spec1;
  if (int rc = launch_pass(ds, r, stream)) return rc;
  c.src = spec1;
  c.dst = spec1;
  if (int rc = launch_pass(ds, c, stream)) return rc;
  r.src = b;
  r.dst = spec2;
  if (int rc = launch_pass(ds, r, stream)) return rc;
  c.src = spec2;
  c.dst = spec2;
  if (int rc = launch_pass(ds, c, stream)) return rc;
  // inverse of the phase-only cross power, SHIFT, unscaled
  PassDesc ri = r;
  ri.in_kind = IN_POC;
  ri.win = nullptr;
  ri.src = spec1;
  ri.src2 = spec2;
  ri.dst = spec1;
  ri.conj_in = 1;
  ri.shift = 1;
  if (int rc = launch_pass(ds, ri, stream)) return rc;
  PassDesc ci = c;
  ci.src = spec1;
  ci.dst = surface;
  ci.out_kind = OUT_R32;
  ci.conj_out = 1;
  ci.shift = 1;
  ci.argmax = partials;
  ci.out_cols = cols;
  if (int rc = launch_pass(ds, ci, stream)) return rc;
  PocFinalArgs f{partials, (int)l.n_partials, surface, rows, cols, dev_result};
  hipLaunchKernelGGL(poc_finalize_kernel, dim3(1), dim3(256), 0, stream, f);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

}  // namespace glsx
