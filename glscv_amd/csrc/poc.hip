// poc.hip -- gls::phaseCorrelate on the device.
//
// glsCV/glsPhaseCorrelate.cpp:142-231 runs: window multiply x2, fft x2,
// mulSpectrumsPhaseOnly, fft(INVERSE|SHIFT), split, minMaxLoc (2 passes + a
// download), then downloads the whole surface for a CPU 5x5 centroid.  Here
// both images are real, so one complex FFT carries both spectra
// (Z = FFT2(a + i b)), and the chain is four FFT passes plus one tiny kernel;
// only the 40-byte result leaves the device:
//   1. row pass  (img1 * window) + i (img2 * window) -> Z   (merge + multiply fused)
//   2. col pass  Z in place
//   3. row pass  over row pairs (r, -r): split A and B off Z, phase-only
//                cross power A*conj(B)/|A*conj(B)|, conj -> Z in place
//                (mulSpectrumsPhaseOnly fused into the inverse FFT's load)
//   4. col pass  Z -> surface (real part, fft-shifted, unscaled) + argmax
//                partials per workgroup              (split + minMaxLoc pass 0)
//   5. finalize: first maximum over the partials, 5x5 FP64 weighted centroid
//      with border clamping, response / (rows*cols), center - centroid.
// Bytes per pixel: 16 + 16 + 16 + 12 = 60 (two separate spectra: 92).
#include "fft_pass.hpp"
#include "glsx_internal.hpp"
#include "poc_final.hpp"

namespace glsx {

namespace {

struct Layout {
  size_t spec = 0, surface = 0, partials = 0, total = 0;
  long long n_partials = 0;
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

Layout layout(const DeviceState* ds, int rows, int cols) {
  Layout l;
  const size_t n = (size_t)rows * cols;
  PassDesc c;  // the inverse column pass: the real surface with argmax partials
  c.L = rows;
  c.col = true;
  c.n_batch = cols;
  c.es = cols;
  c.bs = 1;
  c.out_kind = OUT_R32;
  l.n_partials = pass_grid(ds, c);
  l.spec = 0;
  l.surface = align_up(l.spec + n * sizeof(float2));
  l.partials = align_up(l.surface + n * sizeof(float));
  l.total = align_up(l.partials + (size_t)l.n_partials * sizeof(int2));
  return l;
}

__global__ void __launch_bounds__(256) poc_finalize_kernel(PocFinalArgs f) {
  __shared__ float sv[16];
  __shared__ int si[16];
  poc_finalize_block(f, sv, si);
}

// Step 3 as one launch: workgroup k runs the tile of rows k and M - k.
template <int L>
__global__ void __launch_bounds__(pocz_threads(L) > 64 ? pocz_threads(L) : 64) pocz_row_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  __shared__ float am_v[16];
  __shared__ int am_i[16];
  run_tile_pocz<L>(a, blockIdx.x, lds, am_v, am_i, threadIdx.x);
}

template <int L>
int launch_pocz_t(const DeviceState* ds, const PassArgs& a, hipStream_t stream) {
  auto kern = pocz_row_kernel<L>;
  constexpr int threads = pocz_threads(L) > 64 ? pocz_threads(L) : 64;
  constexpr size_t lds = pocz_lds_bytes(L);
  if (lds + 1024 > ds->lds_per_block) return fail(GLSX_ERR_UNSUPPORTED, "phaseCorrelate: row too long");
  static LdsLimit lim;
  if (int rc = lim.raise((const void*)kern, ds->lds_per_block - 1024)) return rc;
  const unsigned grid = (unsigned)(a.n_batch / 2 + 1);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, stream, a);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

int launch_pocz(const DeviceState* ds, int L, const PassArgs& a, hipStream_t s) {
  switch (L) {
    case 2: return launch_pocz_t<2>(ds, a, s);
    case 4: return launch_pocz_t<4>(ds, a, s);
    case 8: return launch_pocz_t<8>(ds, a, s);
    case 16: return launch_pocz_t<16>(ds, a, s);
    case 32: return launch_pocz_t<32>(ds, a, s);
    case 64: return launch_pocz_t<64>(ds, a, s);
    case 128: return launch_pocz_t<128>(ds, a, s);
    case 256: return launch_pocz_t<256>(ds, a, s);
    case 512: return launch_pocz_t<512>(ds, a, s);
    case 1024: return launch_pocz_t<1024>(ds, a, s);
    case 2048: return launch_pocz_t<2048>(ds, a, s);
    case 4096: return launch_pocz_t<4096>(ds, a, s);
    case 8192: return launch_pocz_t<8192>(ds, a, s);
    default: return fail(GLSX_ERR_UNSUPPORTED, "phaseCorrelate: cols must be a power of two in [2, 8192]");
  }
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
  if (cols > 8192) return fail(GLSX_ERR_UNSUPPORTED, "phaseCorrelate: cols must be at most 8192");
  const Layout l = layout(ds, rows, cols);
  if (workspace_bytes < l.total) return fail(GLSX_ERR_ARG, "phaseCorrelate: workspace too small");
  char* ws = (char*)workspace;
  float2* spec = (float2*)(ws + l.spec);
  float* surface = (float*)(ws + l.surface);
  int2* partials = (int2*)(ws + l.partials);

  // 1-2: Z = FFT2(a*w + i b*w)
  PassDesc r;
  r.L = cols;
  r.col = false;
  r.n_batch = rows;
  r.es = 1;
  r.bs = cols;
  r.in_kind = IN_R32PAIR;
  r.out_kind = OUT_C32;
  r.src = a;
  r.src2 = b;
  r.win = win;
  r.dst = spec;
  if (int rc = launch_pass(ds, r, stream)) return rc;
  PassDesc c;
  c.L = rows;
  c.col = true;
  c.n_batch = cols;
  c.es = cols;
  c.bs = 1;
  c.in_kind = IN_C32;
  c.out_kind = OUT_C32;
  c.src = spec;
  c.dst = spec;
  if (int rc = launch_pass(ds, c, stream)) return rc;
  // 3: inverse row FFT of the phase-only cross power, SHIFT, unscaled
  PassArgs ri{};
  ri.src = spec;
  ri.dst = spec;
  ri.tw = ds->twiddles;
  ri.n_batch = rows;
  ri.es = 1;
  ri.bs = cols;
  ri.in_kind = IN_POCZ;
  ri.out_kind = OUT_C32;
  ri.w = 2;
  ri.log_w = 1;
  ri.conj_in = 1;
  ri.shift = 1;
  ri.scale = 1.0f;
  if (int rc = launch_pocz(ds, cols, ri, stream)) return rc;
  // 4: inverse column FFT into the real surface, SHIFT, argmax partials
  PassDesc ci = c;
  ci.dst = surface;
  ci.out_kind = OUT_R32;
  ci.conj_out = 1;
  ci.shift = 1;
  ci.argmax = partials;
  ci.out_cols = cols;
  // the workspace holds l.n_partials partials, counted without the pointers
  if (pass_grid(ds, ci) != l.n_partials)
    return fail(GLSX_ERR_STATE, "phaseCorrelate: partial count differs from the workspace layout");
  if (int rc = launch_pass(ds, ci, stream)) return rc;
  PocFinalArgs f{partials, (int)l.n_partials, surface, rows, cols, dev_result};
  hipLaunchKernelGGL(poc_finalize_kernel, dim3(1), dim3(256), 0, stream, f);
  GLSX_HIP(hipGetLastError());
  return GLSX_OK;
}

}  // namespace glsx
