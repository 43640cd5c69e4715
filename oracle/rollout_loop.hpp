/*
 * rollout_loop.hpp -- the oracle's rollout loop over any model and cost. TEST INFRASTRUCTURE ONLY.
 *
 * Restates launchCPURolloutKernel (tests/include/kernel_tests/core/rollout_kernel_test.cu:504-541) with the
 * likelihood-ratio argument order of sampling_distributions/gaussian/gaussian.cuh:135-137 and the clamped control
 * written back as in core/mppi_common.cu:117-124: for each sample, for t < T, constrain u_t (written back), step the
 * dynamics, add the running cost of the post-step output y_{t+1} at index t plus the likelihood-ratio control cost
 * (sampling_distributions/gaussian/gaussian.cu:650-675, the mean the samples were drawn around, 0 for pure-noise
 * samples); cost = (sum + terminal(y_T)) / T.
 *
 * oracle.cpp instantiates it for the engine's built-in models; a test's own plugin model instantiates it with a host
 * restatement of that model (the oracle side of a user-plugin parity test).
 *
 * Sys provides, for the arithmetic type R:
 *   int state_dim() const, int control_dim() const, int output_dim() const
 *   void enforce(R* u) const                                   -- Dynamics::enforceConstraints
 *   void step(const R* x, R* xn, R* xd, const R* u, R* y, R dt) const   -- Dynamics::step
 *   R running(const R* y, const R* u, int t) const             -- Cost::computeRunningCost
 *   R terminal(const R* y) const                               -- Cost::terminalCost
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../include/mppi_amd.h"

namespace oracle {

/* sign() of utils/math_utils.h:744-747 (sign(0) = 1) */
template <class R>
inline R sgn(R v)
{
  return v >= (R)0 ? (R)1 : (R)-1;
}

/* Dynamics::enforceConstraints (dynamics/dynamics.cu:107-126): deadband, then clamp to the control range. */
template <class R>
inline void enforce_constraints(const mppi_dynamics_params& p, R* u)
{
  for (int i = 0; i < p.control_dim; ++i) {
    if (std::fabs(u[i]) < (R)p.control_deadband[i]) {
      u[i] = (R)p.zero_control[i];
    } else {
      u[i] += (R)p.control_deadband[i] * -sgn(u[i]);
    }
    u[i] = std::min(std::max((R)p.control_min[i], u[i]), (R)p.control_max[i]);
  }
}

/* first k with (float)k >= (1 - p) K in FP32 (sampling_distributions/gaussian/gaussian.cu:96) */
inline int pure_noise_start(int K, float pct)
{
  const float thr = (1.0f - pct) * (float)K;
  int k = 0;
  while (k < K && !((float)k >= thr)) ++k;
  return k;
}

/* Likelihood-ratio control cost of step t, gaussian.cu:650-675, mean = the mean the samples were drawn around; with
 * time_specific_std_dev the step's standard deviations (:665-668; s.std_dev_t is a host pointer here). */
template <class R>
inline R lr_cost(const mppi_sampler_params& s, int C, const R* u, const float* mean_t, bool pure, R lambda, R alpha,
                 int t)
{
  R cost = 0;
  for (int i = 0; i < C; ++i) {
    const R m = pure ? (R)0 : (R)mean_t[i];
    const R sd = s.std_dev_t ? s.std_dev_t[(size_t)t * C + i] : s.std_dev[i];
    cost += (R)s.control_cost_coeff[i] * m * (m - (R)2 * u[i]) / (sd * sd);
  }
  return (R)0.5 * lambda * ((R)1 - alpha) * cost;
}

constexpr int LOOP_MAX_DIM = 64;

/* Records nothing (the plain loop). */
struct NoTrace {
  template <class R>
  void step(int, int, const R*, int, R) const {}
  template <class R>
  void terminal(int, R) const {}
};

template <class R, class Sys, class Trace = NoTrace>
void rollout_loop(const Sys& sys, const mppi_sampler_params& smp, int K, int T, R dt, R lambda, R alpha,
                  const float* x0, const float* mean, float* samples, R* costs, int k_begin, int k_end,
                  const Trace& trace = Trace())
{
  const int S = sys.state_dim(), C = sys.control_dim();
  const int pstart = pure_noise_start(K, smp.pure_noise_trajectories_percentage);
  for (int k = k_begin; k < k_end; ++k) {
    R x[LOOP_MAX_DIM] = {0}, xn[LOOP_MAX_DIM] = {0}, xd[LOOP_MAX_DIM] = {0}, y[LOOP_MAX_DIM] = {0};
    R u[LOOP_MAX_DIM];
    for (int i = 0; i < S; ++i) x[i] = x0[i];
    R run = 0;
    for (int t = 0; t < T; ++t) {
      float* ut = samples + ((size_t)k * T + t) * C;
      for (int i = 0; i < C; ++i) u[i] = ut[i];
      sys.enforce(u);
      for (int i = 0; i < C; ++i) ut[i] = (float)u[i];
      sys.step(x, xn, xd, u, y, dt);
      const R c = sys.running(y, u, t) + lr_cost<R>(smp, C, u, mean + (size_t)t * C, k >= pstart, lambda, alpha, t);
      trace.step(k, t, y, sys.output_dim(), c / (R)T);
      run += c;
      for (int i = 0; i < S; ++i) x[i] = xn[i];
    }
    trace.terminal(k, sys.terminal(y) / (R)T);
    costs[k] = run / (R)T + sys.terminal(y) / (R)T;
  }
}

/* The loop split over `nthreads` std::thread workers by sample range (the multi-core CPU baseline). */
template <class R, class Sys>
void rollout_loop_threads(const Sys& sys, const mppi_sampler_params& smp, int K, int T, R dt, R lambda, R alpha,
                          const float* x0, const float* mean, float* samples, R* costs, int nthreads)
{
  if (nthreads <= 1) {
    rollout_loop<R>(sys, smp, K, T, dt, lambda, alpha, x0, mean, samples, costs, 0, K);
    return;
  }
  std::vector<std::thread> th;
  const int chunk = (K + nthreads - 1) / nthreads;
  for (int w = 0; w < nthreads; ++w) {
    const int b = w * chunk, e = std::min(K, b + chunk);
    if (b >= e) break;
    th.emplace_back([&, b, e]() { rollout_loop<R>(sys, smp, K, T, dt, lambda, alpha, x0, mean, samples, costs, b, e); });
  }
  for (auto& t : th) t.join();
}

} /* namespace oracle */
