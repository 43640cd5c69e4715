// Weighted control update that allocates nothing: the partial rows of launch_weighted_update live in a file-scope
// device array instead of the per-device workspace that mppi_weighted_update grows with hipMalloc. For callers that
// cannot allocate (code under graph capture, tight serving loops).
//
// Every GPU holds its own copy of `scratch` until the process ends. Host code never passes the array's name to a
// kernel: the name is a host-side stand-in, not a device address. The device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and kept; all launches use it. Every call on one GPU shares that
// GPU's array and leaves its contents behind; the partial rows are written in full before they are read within a
// call, so nothing needs clearing, but calls that overlap on one GPU (different streams) race on the array.
#include <mutex>
#include <string>

#include "host.hpp"
#include "kernels.hpp"

namespace mppi_amd {

constexpr int MPPI_NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) float scratch[MPPI_NOALLOC_SCRATCH_BYTES / sizeof(float)];

// Device address of `scratch` on the current device, looked up on the device's first call.
static int scratch_address(void** out)
{
  static std::mutex mu;
  static void* addr[MPPI_NOALLOC_MAX_DEVICES] = {};
  int dev = 0;
  MPPI_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MPPI_NOALLOC_MAX_DEVICES) return fail(MPPI_E_INVALID, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (p == nullptr) return fail(MPPI_E_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    addr[dev] = p;
  }
  *out = addr[dev];
  return MPPI_OK;
}

}  // namespace mppi_amd

using namespace mppi_amd;

extern "C" int mppi_weighted_update_noalloc(const float* weights_d, const mppi_weight_stats* stats_d,
                                            const float* samples_d, int32_t K, int32_t T, int32_t C,
                                            float* mean_out_d, void* stream)
{
  if (!weights_d || !stats_d || !samples_d || !mean_out_d) return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(validate_sizes(K, T));
  if (C < 1 || C > MPPI_MAX_CONTROL) return fail(MPPI_E_INVALID, "control_dim out of range");
  const size_t need = weighted_update_workspace_bytes(K, T, C);
  if (need > MPPI_NOALLOC_SCRATCH_BYTES)
    return fail(MPPI_E_SIZE, "weighted_update_noalloc: ceil(K / 64) * T * C * 4 = " + std::to_string(need) +
                                 " bytes exceeds the " + std::to_string(MPPI_NOALLOC_SCRATCH_BYTES) +
                                 "-byte scratch array");
  void* ws = nullptr;
  MPPI_CHECK(scratch_address(&ws));
  return launch_weighted_update(weights_d, stats_d, samples_d, K, T, C, mean_out_d, ws, (hipStream_t)stream);
}
