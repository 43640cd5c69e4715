// Fused iteration (sampling + rollout + local weights + partial update, then a log-sum-exp combine).
#pragma once

#include "colored_core.hpp"
#include "kernels.hpp"
#include "models.hpp"

namespace mppi_amd {

// Per-workgroup partial of the weighted update, relative to the workgroup's own baseline.
struct Partial {
  float baseline;
  int argmin;
  double eta;
  double s2;
  double wc;  // sum_k w_k cost_k with the workgroup's weights (finite costs)
};

size_t fused_iteration_workspace_bytes(int K, int T, int C);

// Optional copies of an iteration's results that the combine kernel also writes, typically to mapped host memory
// (device-side addresses): the new mean [T C], the statistics, and cost0_src[0] to cost0. Null members are skipped.
// `signal` (optional) is given once all of these are visible to the host. With a signal and `control`, the combine's
// last workgroup also writes the new mean Savitzky-Golay smoothed and constrained by `limits` to `control` [T C]
// (host) and `nominal` [T C] (device).
struct ResultCopies {
  float* mean = nullptr;
  mppi_weight_stats* stats = nullptr;
  const float* cost0_src = nullptr;
  float* cost0 = nullptr;
  HostSignal signal;
  float* control = nullptr;
  float* nominal = nullptr;
  ControlLimits limits{};
  int T = 0, C = 0;
};

// One MPPI iteration in two launches. With `generate` the Gaussian control samples are drawn inside the rollout
// kernel (rng offset, stride and iteration from `it`); otherwise they are read from `samples` (e.g. written by the
// colored sampler). `write_samples` stores the clamped samples back to `samples`; `costs_out` (optional) receives
// the per-sample costs. `advance` (optional) gets its rng offset incremented by `advance_step` in the combine kernel;
// `between` (optional) is recorded on the stream between the rollout and the combine launches; `partials_out`
// (optional) receives the number of per-workgroup partials the combine reduces; `copies`
// (optional) names further destinations of the results. With `colored` (generate false) the samples are colored
// noise of that plan: drawn inside the rollout kernel where its workgroup shape allows (the same values as the colored
// sampler's), else by the colored sampler into `samples` before it.
int launch_fused_iteration(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                           const mppi_sampler_params& sp, int K, int T, float dt, float lambda, float alpha,
                           const float* x0, const float* mean, float* samples, bool generate, bool write_samples,
                           IterRef it, float* costs_out, void* ws, mppi_weight_stats* stats, float* mean_out,
                           IterArgs* advance, hipStream_t s, const float* prop_dmu = nullptr,
                           const float* prop_sig = nullptr, const ResultCopies* copies = nullptr,
                           const ColoredPlan* colored = nullptr, int advance_step = 1,
                           hipEvent_t between = nullptr, int* partials_out = nullptr);

// The combine launch alone, over G per-workgroup partials parts[G] and weighted sums sums[G][TC] produced by any
// rollout kernel (the built-in fused kernel or a user plugin's): global first-strict baseline, FP64 normalizer, new
// mean sums / eta into mean_out, statistics, and the optional advance / copies as above. G <= MAX_PARTIALS.
constexpr int MAX_PARTIALS = 4096;

// The fused two-system iteration (Tube-MPPI / RMPPI, Gaussian samples of `sp` for both systems, both drawn from the
// same normals around mean[d][T][C]): rollout of both systems, then per system d the Partials parts[d gstride + g] and
// partial rows psums[(d gstride + g) TC ..] of the G workgroups (G <= gstride, returned), the costs [2][K] and the
// applied controls [2][K][T][C]. launch_combine with systems = 2 and pstride = gstride forms both new means.
int launch_two_system_fused(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                            const mppi_sampler_params& sp, const mppi_two_system_params& tp, int K, int T, float dt,
                            float lambda, float alpha, const float* x0, const float* mean, float* samples,
                            float* costs, IterRef itr, Partial* parts, float* psums, int gstride, int* G,
                            hipStream_t s);
// With systems = 2 one launch combines two independent sets: set y's partials at parts[y pstride ..], its rows at
// sums[y pstride TC ..], its statistics at stats[y], its mean at mean_out[y TC ..] (copies.mean / copies.stats
// likewise); `advance` and cost0 belong to set 0, and the smoothed-control copy is not available.
int launch_combine(const Partial* parts, const float* sums, int G, int K, int TC, float lambda,
                   mppi_weight_stats* stats, float* mean_out, IterArgs* advance, int advance_step,
                   const ResultCopies* copies, hipStream_t s, int systems = 1, int pstride = 0);

}  // namespace mppi_amd
