// Kernels for user-written dynamics and cost plugins, instantiated in the user's own hipcc translation unit.
//
// The reference lets users write CRTP Dynamics / Cost classes whose __device__ methods its rollout kernel calls
// (dynamics/dynamics.cuh:366-432, cost_functions/cost.cu:39-53). Device code cannot cross a C ABI, so here the
// kernels are templates over the plugin types: the user's TU includes this header (through
// include/mppi_amd/device_plugin.hpp), which instantiates
//   plugin_iteration_kernel<DYN, COST, GEN, WRITE> -- one fused MPPI iteration of the controller: Gaussian samples
//       drawn into the LDS control tile (or read from the sample buffer, e.g. after the library's colored sampler),
//       the user's constraint, the engine's pipelined rollout (rollout_tile_t) calling the user's dynamics and cost,
//       and the tile's partials for the library's combine kernel -- the same pieces as the built-in fused kernel;
//   plugin_rollout_kernel<DYN, COST> -- the kernel-level rollout (mppi_rollout's counterpart): costs and clamped
//       samples of a sample buffer.
// DYN and COST are the user's final classes (trivially copyable, passed by value as kernel arguments):
//   DYN:  STATE_DIM, CONTROL_DIM, OUTPUT_DIM; computeStateDeriv(x, u, x_dot); updateState(x, x_next, x_dot, dt);
//         stateToOutput(x, y); enforceControlConstraint(i, u_i)  (all const, __host__ __device__)
//   COST: computeRunningCost(y, u, t); terminalCost(y)  (const, __host__ __device__)
// The host launchers return MPPI_* status codes; the library's combine kernel is launched by the controller.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <set>
#include <utility>

#include "iteration_core.hpp"

namespace mppi_amd {
namespace plugin {

template <class DYN>
struct PluginModel {
  static constexpr int S = DYN::STATE_DIM, C = DYN::CONTROL_DIM, O = DYN::OUTPUT_DIM;
  __device__ static void deriv(const DYN& d, const float* x, const float* u, float* xd) { d.computeStateDeriv(x, u, xd); }
  __device__ static void update(const DYN& d, const float* x, float* xn, const float* xd, float dt)
  {
    d.updateState(x, xn, xd, dt);
  }
  __device__ static void output(const DYN& d, const float* x, float* y) { d.stateToOutput(x, y); }
};

template <class COST>
struct PluginCost {
  __device__ static float running(const COST& c, const float* y, const float* u, int t)
  {
    return c.computeRunningCost(y, u, t);
  }
  __device__ static float terminal(const COST& c, const float* y) { return c.terminalCost(y); }
};

template <class DYN>
struct PluginConstrain {
  const DYN* d;
  __device__ float operator()(int col, float v) const { return d->enforceControlConstraint(col % DYN::CONTROL_DIM, v); }
};

template <class DYN>
constexpr void check_dims()
{
  static_assert(DYN::STATE_DIM >= 1 && DYN::STATE_DIM <= MPPI_MAX_STATE, "plugin STATE_DIM out of range");
  static_assert(DYN::CONTROL_DIM >= 1 && DYN::CONTROL_DIM <= MPPI_MAX_CONTROL, "plugin CONTROL_DIM out of range");
  static_assert(DYN::OUTPUT_DIM >= 1 && DYN::OUTPUT_DIM <= 64, "plugin OUTPUT_DIM out of range");
}

// One fused iteration over `rows` samples per workgroup (16, 32 or 64); 1024 threads: wavefront 0 runs the state
// chain, wavefronts 1-15 the costs.
template <class DYN, class COST, bool GEN, bool WRITE>
__global__ void __launch_bounds__(RC_THREADS) plugin_iteration_kernel(
    DYN dyn, COST cost, mppi_sampler_params sp, int K, int T, float dt, float lambda, float alpha,
    const float* __restrict__ x0, const float* __restrict__ mean, float* __restrict__ samples,
    const IterArgs* __restrict__ itr, int pstart, float* __restrict__ costs_out, Partial* __restrict__ parts,
    float* __restrict__ sums, int chunk, int rows)
{
  constexpr int C = DYN::CONTROL_DIM, S = DYN::STATE_DIM;
  constexpr int NT = RC_THREADS;
  extern __shared__ float lds[];
  const int TC = T * C, TC4 = (TC + 3) & ~3, ROW = TC + 1;
  float* tile = lds;
  float* mu = tile + ((rows * ROW + 3) & ~3);
  float* wl = mu + TC4;              // [64] local weights; x0 until the rollout has read it
  float* scratch = wl + RC_LANES;    // pipelined-rollout buffers
  const int tid = threadIdx.x;
  const int g = blockIdx.x;
  const int k0 = g * rows;
  const int nrows = min(rows, K - k0);
  for (int j = tid; j < TC; j += NT) mu[j] = mean[j];
  if (tid < S) wl[tid] = x0[tid];
  const PluginConstrain<DYN> f{&dyn};
  if constexpr (GEN) {
    __syncthreads();
    gaussian_tile_fill<C, NT, false>(sp, *itr, k0, nrows, TC, ROW, pstart, mu, nullptr, TC4, tile, f);
  } else {
    tile_load<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid, f);
  }
  __syncthreads();
  float* cost_l;
  float* raw_l;
  rollout_tile_t<PluginModel<DYN>, PluginCost<COST>>(dyn, cost, sp, T, dt, lambda, alpha, wl, tile, ROW, mu, pstart,
                                                     k0, scratch, chunk, cost_l, raw_l, nullptr, rows);
  tile_partials<NT, WRITE, false>(cost_l, raw_l, wl, tile, ROW, TC, k0, nrows, g, lambda, costs_out, parts, sums,
                                  samples);
}

// Kernel-level rollout of a [K][T][C] sample buffer (rolloutKernel, core/mppi_common.cu:35-153): constrain in place,
// roll out, costs[K] = (sum_t running(y_{t+1}, u_t, t) + LR(u_t) + terminal(y_T)) / T.
template <class DYN, class COST>
__global__ void __launch_bounds__(RC_THREADS) plugin_rollout_kernel(DYN dyn, COST cost, mppi_sampler_params sp, int K,
                                                                    int T, float dt, float lambda, float alpha,
                                                                    const float* __restrict__ x0,
                                                                    const float* __restrict__ mean,
                                                                    float* __restrict__ samples,
                                                                    float* __restrict__ costs, int pstart, int chunk)
{
  constexpr int C = DYN::CONTROL_DIM, S = DYN::STATE_DIM;
  constexpr int NT = RC_THREADS;
  extern __shared__ float lds[];
  const int TC = T * C, TC4 = (TC + 3) & ~3, ROW = TC + 1;
  float* tile = lds;
  float* mu = tile + ((RC_LANES * ROW + 3) & ~3);
  float* x0_l = mu + TC4;
  float* scratch = x0_l + RC_LANES;
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * RC_LANES;
  const int nrows = min(RC_LANES, K - k0);
  for (int j = tid; j < TC; j += NT) mu[j] = mean[j];
  if (tid < S) x0_l[tid] = x0[tid];
  tile_load<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid, PluginConstrain<DYN>{&dyn});
  __syncthreads();
  float* cost_l;
  float* raw_l;
  rollout_tile_t<PluginModel<DYN>, PluginCost<COST>>(dyn, cost, sp, T, dt, lambda, alpha, x0_l, tile, ROW, mu, pstart,
                                                     k0, scratch, chunk, cost_l, raw_l, nullptr, nrows);
  if (tid < nrows) costs[k0 + tid] = cost_l[tid];
  tile_store<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid);
}

// ------------------------------------------------------------------------------------------------ host side
// Compute units of the current device (0 when the query fails).
inline int device_cus()
{
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return n;
}

// The dynamic-LDS limit of `fn` raised on the current device, once per (device, kernel).
inline int allow_lds(const void* fn)
{
  static std::mutex mu;
  static std::set<std::pair<int, const void*>> done;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return MPPI_E_HIP;
  std::lock_guard<std::mutex> lock(mu);
  if (done.count({dev, fn}) != 0) return MPPI_OK;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_LIMIT) != hipSuccess) return MPPI_E_HIP;
  done.insert({dev, fn});
  return MPPI_OK;
}

// LDS floats for `rows` samples per workgroup: tile, mean, 64 slots, then the rollout buffers of the largest chunk
// that fits; 0 (and *chunk 0) when the horizon does not fit.
template <int C, int O>
inline size_t lds_floats(int T, int rows, int* chunk)
{
  const size_t TC4 = (size_t)((T * C + 3) & ~3);
  const size_t base = (((size_t)rows * (T * C + 1) + 3) & ~(size_t)3) + TC4 + RC_LANES;
  const size_t limit = LDS_LIMIT / sizeof(float);
  *chunk = base < limit ? rollout_chunk<O>(limit - base) : 0;
  return *chunk == 0 ? 0 : base + rollout_scratch_floats<O>(*chunk);
}

inline int sampler_ok(const mppi_sampler_params& sp, int C, int K, int T)
{
  if (K < 1 || T < 1 || T > MPPI_MAX_TIMESTEPS) return MPPI_E_INVALID;
  if (sp.control_dim != C) return MPPI_E_INVALID;
  return MPPI_OK;
}

// Kernel-level rollout of device buffers (the plugin counterpart of mppi_rollout).
template <class DYN, class COST>
int launch_rollout(const DYN& dyn, const COST& cost, const mppi_sampler_params& sp, int K, int T, float dt,
                   float lambda, float alpha, const float* x0_d, const float* mean_d, float* samples_d, float* costs_d,
                   hipStream_t s)
{
  check_dims<DYN>();
  constexpr int C = DYN::CONTROL_DIM, O = DYN::OUTPUT_DIM;
  if (!x0_d || !mean_d || !samples_d || !costs_d) return MPPI_E_INVALID;
  if (const int st = sampler_ok(sp, C, K, T); st != MPPI_OK) return st;
  int chunk = 0;
  const size_t lf = lds_floats<C, O>(T, RC_LANES, &chunk);
  if (lf == 0) return MPPI_E_SIZE;
  const void* fn = (const void*)plugin_rollout_kernel<DYN, COST>;
  if (const int st = allow_lds(fn); st != MPPI_OK) return st;
  int pstart = 0;
  if (mppi_pure_noise_start(K, sp.pure_noise_trajectories_percentage, &pstart) != MPPI_OK) return MPPI_E_INVALID;
  plugin_rollout_kernel<DYN, COST><<<(K + RC_LANES - 1) / RC_LANES, RC_THREADS, sizeof(float) * lf, s>>>(
      dyn, cost, sp, K, T, dt, lambda, alpha, x0_d, mean_d, samples_d, costs_d, pstart, chunk);
  return hipGetLastError() == hipSuccess ? MPPI_OK : MPPI_E_HIP;
}

// One fused iteration for the controller (the mppi_plugin::enqueue_iteration callback): partials and weighted sums
// for the library's combine kernel; *num_partials receives their count.
template <class DYN, class COST>
int enqueue_iteration(const DYN& dyn, const COST& cost, const mppi_plugin_iteration& it, int32_t* num_partials)
{
  check_dims<DYN>();
  constexpr int C = DYN::CONTROL_DIM, O = DYN::OUTPUT_DIM;
  const int K = it.num_rollouts, T = it.num_timesteps;
  if (!it.x0_d || !it.mean_d || !it.iter_d || !it.partials_d || !it.sums_d || !num_partials) return MPPI_E_INVALID;
  if (const int st = sampler_ok(it.sampler, C, K, T); st != MPPI_OK) return st;
  if (!it.generate && !it.samples_d) return MPPI_E_INVALID;
  const int rows = fused_rows_for(K, device_cus());
  const int G = (K + rows - 1) / rows;
  if (G > it.max_partials) return MPPI_E_SIZE;
  int chunk = 0;
  const size_t lf = lds_floats<C, O>(T, rows, &chunk);
  if (lf == 0) return MPPI_E_SIZE;
  int pstart = 0;
  if (mppi_pure_noise_start(K, it.sampler.pure_noise_trajectories_percentage, &pstart) != MPPI_OK)
    return MPPI_E_INVALID;
  const hipStream_t s = (hipStream_t)it.stream;
  const IterArgs* itr = static_cast<const IterArgs*>(it.iter_d);
  Partial* parts = static_cast<Partial*>(it.partials_d);
#define MPPI_PLUGIN_LAUNCH(GEN, WRITE)                                                                              \
  do {                                                                                                              \
    const void* fn = (const void*)plugin_iteration_kernel<DYN, COST, GEN, WRITE>;                                   \
    if (const int st = allow_lds(fn); st != MPPI_OK) return st;                                                     \
    plugin_iteration_kernel<DYN, COST, GEN, WRITE><<<G, RC_THREADS, sizeof(float) * lf, s>>>(                       \
        dyn, cost, it.sampler, K, T, it.dt, it.lambda, it.alpha, it.x0_d, it.mean_d, it.samples_d, itr, pstart,    \
        it.costs_d, parts, it.sums_d, chunk, rows);                                                                 \
  } while (0)
  if (it.generate) {
    if (it.write_samples) MPPI_PLUGIN_LAUNCH(true, true);
    else MPPI_PLUGIN_LAUNCH(true, false);
  } else {
    if (it.write_samples) MPPI_PLUGIN_LAUNCH(false, true);
    else MPPI_PLUGIN_LAUNCH(false, false);
  }
#undef MPPI_PLUGIN_LAUNCH
  if (hipGetLastError() != hipSuccess) return MPPI_E_HIP;
  *num_partials = G;
  return MPPI_OK;
}

}  // namespace plugin
}  // namespace mppi_amd
