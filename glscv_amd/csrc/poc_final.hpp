// poc_final.hpp -- the last step of gls::phaseCorrelate on the device:
// first maximum over the per-tile argmax partials (glsCV/glsMinMaxLoc.cpp
// pass 1, :99-112), then the 5x5 weighted centroid with border clamping in
// FP64 (glsCV/glsPhaseCorrelate.cpp:60-137), the response normalisation
// (:224-225) and center - centroid (:228-230).
#pragma once
#include "fft_pass.hpp"
#include "glsx_internal.hpp"

namespace glsx {

struct PocFinalArgs {
  const int2* partials;
  int n_partials;
  const float* surface;
  int rows, cols;
  glsx_poc_result* out;
};

// Run by a whole workgroup (blockDim a multiple of 64, at most 1024);
// sv/si hold one entry per wave.
__device__ __forceinline__ void poc_finalize_block(const PocFinalArgs& f, float* sv, int* si) {
  float best = -__FLT_MAX__;
  int best_i = 0x7fffffff;
  for (int i = threadIdx.x; i < f.n_partials; i += blockDim.x) {
    const int2 p = f.partials[i];
    const float v = __int_as_float(p.x);
    if (am_better(v, p.y, best, best_i)) { best = v; best_i = p.y; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const float ov = __shfl_xor(best, off);
    const int oi = __shfl_xor(best_i, off);
    if (am_better(ov, oi, best, best_i)) { best = ov; best_i = oi; }
  }
  if ((threadIdx.x & 63) == 0) { sv[threadIdx.x >> 6] = best; si[threadIdx.x >> 6] = best_i; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      if (am_better(sv[w], si[w], best, best_i)) { best = sv[w]; best_i = si[w]; }
    const int rows = f.rows, cols = f.cols;
    // nothing above -FLT_MAX (e.g. all NaN) reports location (0,0), as the
    // reference's strict '>' scan from -FLT_MAX does
    int px = 0, py = 0;
    if (best_i != 0x7fffffff && best > -__FLT_MAX__) { px = best_i % cols; py = best_i / cols; }
    int minr = py - 2, maxr = py + 2, minc = px - 2, maxc = px + 2;
    if (minr < 0) minr = 0;
    if (minc < 0) minc = 0;
    if (maxr > rows - 1) maxr = rows - 1;
    if (maxc > cols - 1) maxc = cols - 1;
    double sx = 0.0, sy = 0.0, sum = 0.0;
    for (int y = minr; y <= maxr; ++y) {
      for (int x = minc; x <= maxc; ++x) {
        const double v = (double)f.surface[(size_t)y * cols + x];
        sx += (double)x * v;
        sy += (double)y * v;
        sum += v;
      }
    }
    const double response = sum;
    sum += 2.220446049250313e-16;  // DBL_EPSILON
    glsx_poc_result r;
    r.peak_x = px;
    r.peak_y = py;
    r.shift_x = (double)cols / 2.0 - sx / sum;
    r.shift_y = (double)rows / 2.0 - sy / sum;
    r.response = response / ((double)rows * (double)cols);
    *f.out = r;
  }
}

}  // namespace glsx
