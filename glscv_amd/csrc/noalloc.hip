// noalloc.hip -- the allocation-free phase correlation: the chain of
// glsx_phase_correlate_dev (poc.hip) on a static device workspace, for callers
// that cannot allocate (code under HIP graph capture, serving loops that must
// not call hipMalloc).  It has no counterpart in the reference.
#include <string>

#include "glsx_internal.hpp"

namespace glsx {
// The workspace of the allocation-free phase correlation: a static device
// array (each device holds its own copy for the life of the process).  Host
// code never passes `scratch` itself to a kernel -- in host code the name is a
// stand-in, not a device address -- but the address hipGetSymbolAddress
// returns for the current device, kept in DeviceState::static_ws.  Nothing in
// it has to start at zero: every chain writes its spectrum, surface and
// partials before reading them.
static __device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];

int init_static_workspace(DeviceState* ds) {
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(GLSX_ERR_HIP, std::string("static workspace lookup: ") + hipGetErrorString(e));
  ds->static_ws = ptr;
  return GLSX_OK;
}

}  // namespace glsx

using namespace glsx;

#define GLSX_API extern "C" __attribute__((visibility("default")))

// The chain of glsx_phase_correlate_dev on the static workspace: no device or
// host allocation, no synchronisation; the result goes to device memory.
GLSX_API int glsx_phase_correlate_dev_noalloc(const float* src1, const float* src2, const float* window, int rows,
                                              int cols, glsx_poc_result* dev_out, void* stream) {
  if (!src1 || !src2 || !dev_out) return fail(GLSX_ERR_ARG, "phaseCorrelate (no-alloc): null argument");
  int st = 0;
  // the device must have been initialised (glsx_init): initialising here would allocate
  DeviceState* ds = device_state_if_ready(&st);
  if (!ds) return st;
  if (int rc = validate_fft_shape(rows, cols)) return rc;
  if (!ds->static_ws) return fail(GLSX_ERR_STATE, "phaseCorrelate (no-alloc): static workspace not set up");
  const size_t need = phase_correlate_workspace_bytes(ds, rows, cols);
  if (need > kStaticWorkspaceBytes)
    return fail(GLSX_ERR_UNSUPPORTED, "phaseCorrelate (no-alloc): " + std::to_string(rows) + "x" +
                                          std::to_string(cols) + " needs " + std::to_string(need) +
                                          " workspace bytes, more than the static " +
                                          std::to_string(kStaticWorkspaceBytes) + " (GLSX_NOALLOC_WORKSPACE_BYTES)");
  return phase_correlate_device(ds, src1, src2, window, rows, cols, ds->static_ws, kStaticWorkspaceBytes, dev_out,
                                (hipStream_t)stream);
}
