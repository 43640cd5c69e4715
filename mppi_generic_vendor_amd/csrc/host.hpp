// Host-side helpers shared by the controller and the batch objects: parameter validation, device/pinned
// allocation and the computeControl post-processing (smoothing, state propagation, clamping, free energy).
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace mppi_amd {

int model_dims(int model, int* S, int* C, int* O);
int validate_dynamics(const mppi_dynamics_params& d);
int validate_cost(const mppi_cost_params& c, const mppi_dynamics_params& d, bool need_map_ptr);
int validate_goal_traj(const mppi_cost_params& c);
// gamma / r of the Tsallis weights: off (either 0) unless `required`, else gamma > 0 and r != 1.
int validate_tsallis(float gamma, float r, bool required);
int validate_sampler(const mppi_sampler_params& s, const mppi_dynamics_params& d);
int validate_sizes(int K, int T);

// Savitzky-Golay smoothing keeping u[0] (controllers/controller.cuh:573-607).
void smooth_control(std::vector<float>& u, int T, int C);
// computeOutputTrajectoryHelper (controllers/controller.cuh:664-684): states [T][S] of the controls from x0.
void propagate(const mppi_dynamics_params& p, const float* x0, const float* u, int T, float dt, float* states);
// enforceConstraints on every step (controllers/MPPI/mppi_controller.cu:224-227).
void clamp_controls_seq(const mppi_dynamics_params& p, float* u, int T);
// Free-energy statistics (core/mppi_common.cu:1101-1117) from the FP64 weight sums; fe->previous_baseline is input.
void free_energy_from_stats(const mppi_weight_stats& st, double lambda, int K, mppi_free_energy* fe);

// Waits for the completion signal `want` in the mapped word `flag` (HostSignal): polls it, since waking from a
// blocking wait costs about as much as a short control call, and after HOST_SPIN_MS blocks on `ev` (recorded after
// the call's last kernel), which also reports a failed call.
constexpr int HOST_SPIN_MS = 20;
int wait_host_signal(const uint32_t* flag, uint32_t want, hipEvent_t ev);

template <class Tp>
int dev_alloc(Tp** p, size_t bytes)
{
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) return fail(MPPI_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return MPPI_OK;
}

template <class Tp>
int host_alloc(Tp** p, size_t bytes)
{
  hipError_t e = hipHostMalloc((void**)p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return fail(MPPI_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  std::memset(*p, 0, bytes);
  return MPPI_OK;
}

// Pinned host memory that kernels read and write directly: mapped, fine-grained (uncached on the device side, so a
// kernel's stores are in host memory once it has completed and the host's stores are seen by the next launch).
// *dev receives the device-side address of the same bytes.
template <class Tp>
int host_alloc_mapped(Tp** p, Tp** dev, size_t bytes)
{
  hipError_t e = hipHostMalloc((void**)p, bytes, hipHostMallocMapped | hipHostMallocCoherent);
  if (e != hipSuccess) return fail(MPPI_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  std::memset(*p, 0, bytes);
  e = hipHostGetDevicePointer((void**)dev, *p, 0);
  if (e != hipSuccess) return fail(MPPI_E_HIP, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
  return MPPI_OK;
}

}  // namespace mppi_amd
