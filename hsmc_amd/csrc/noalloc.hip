// noalloc.hip -- the allocation-free variant of the scaled-overlap scan.
//
// hsmc_gpu_any_overlap_scaled builds its neighbour grid in the context's
// scratch buffers, which grow with hipMalloc on demand.  Callers that may not
// allocate (code under stream capture, tight serving loops) use
// hsmc_gpu_any_overlap_scaled_noalloc instead: the same kernels, with the
// grid, the scale table and the result mask carved from `scratch`, a
// file-scope device array of kScratchBytes that every GPU the library runs
// on holds until the process ends.
//
// The array's name in host code is a host-side stand-in, never a device
// address: its device address on each GPU is looked up once per device with
// hipGetSymbolAddress (at context creation, with that device current) and
// kept in `scratch_ptr`, and only that pointer reaches kernels.  Calls on one
// GPU share the array, which keeps the last call's contents; every region
// that must start at zero is cleared by a kernel in each call, and calls
// that overlap in time on one GPU race.
#include <mutex>

#include "ctx.h"

namespace {

constexpr size_t kScratchBytes = 16u << 20;  // 16 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex scratch_mu;
unsigned char *scratch_ptr[kMaxDevices] = {};

// Device address of the array on the current device.  Defined here, outside
// namespace hsmc, so that `scratch` names this array and not the context's
// scratch-buffer function hsmc::scratch.
hipError_t scratch_symbol_address(void **p) {
  return hipGetSymbolAddress(p, HIP_SYMBOL(scratch));
}

}  // namespace

namespace hsmc {

// Device address of `scratch` on `device`, which must be the current device.
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return HSMC_ERR_NOT_SUPPORTED;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[device]) return HSMC_OK;
  void *p = nullptr;
  if (scratch_symbol_address(&p) != hipSuccess || !p) return HSMC_ERR_HIP;
  scratch_ptr[device] = (unsigned char *)p;
  return HSMC_OK;
}

}  // namespace hsmc

extern "C" {

int hsmc_gpu_any_overlap_scaled_noalloc(hsmc_gpu_ctx *ctx, int replica, const double *sf, int n_sf,
                                        uint8_t *no_overlap) {
  if (!ctx || replica < 0 || replica >= ctx->R || n_sf < 1 || n_sf > 64 || !sf || !no_overlap)
    return HSMC_ERR_BAD_ARG;
  unsigned char *ws = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) ws = scratch_ptr[ctx->device];
  }
  if (!ws) {
    ctx->last_error = "the allocation-free scratch array was not resolved on this device";
    return HSMC_ERR_NOT_SUPPORTED;
  }
  const size_t need = hsmc::overlap_scan_workspace_bytes(ctx, replica, sf, n_sf);
  if (need == 0) return HSMC_ERR_BAD_ARG;
  if (need > kScratchBytes) {
    ctx->last_error = "the scan of this replica needs " + std::to_string(need) +
                      " bytes, above the 16 MiB allocation-free scratch";
    return HSMC_ERR_NOT_SUPPORTED;
  }
  return hsmc::overlap_scan_in_workspace(ctx, replica, sf, n_sf, ws, kScratchBytes, no_overlap);
}

}  // extern "C"
