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

/* computeFeedbackCost (sampling_distributions/gaussian/gaussian.cu:632-647): 0.5 lambda (1 - alpha) sum_c coeff_c
 * fb_c^2 / sigma_c^2, the distribution's std_dev (or the step's row of std_dev_t). */
template <class R>
inline R feedback_cost(const mppi_sampler_params& s, int C, const R* fb, R lambda, R alpha, int t)
{
  R cost = 0;
  for (int i = 0; i < C; ++i) {
    const R sd = s.std_dev_t ? s.std_dev_t[(size_t)t * C + i] : s.std_dev[i];
    cost += (R)s.control_cost_coeff[i] * fb[i] * fb[i] / (sd * sd);
  }
  return (R)0.5 * lambda * ((R)1 - alpha) * cost;
}

/* Two-system rollout: systems d = 0, 1 with initial states x0[d][S], means mean[d][T][C] and samples
 * samples[d][K][T][C] (written back with the applied controls), costs[d][K].
 *   nom < 0 (Tube-MPPI, controllers/Tube-MPPI/tube_mppi_controller.cu:158-236): each system is rolled out alone as
 *     rollout_loop does, with its own sampler smp[d] in the likelihood-ratio cost.
 *   nom = 0 or 1 (RMPPI, launchCPURMPPIRolloutKernel, tests/include/kernel_tests/core/rmppi_kernel_test.cu:7-70, with
 *     the likelihood-ratio arguments of the GPU kernel, core/rmppi_kernels.cu:775-813, since the CPU test's
 *     computeLikelihoodRatioCost(u, t, idx, lambda, alpha) call binds them to (sample_index, t, distribution_idx,
 *     lambda) of gaussian.cu:650-653): real r = 1 - nom, fb_t = K_t (x_r - x_n) with K_t[j][i] = gains[(t S + i) C +
 *     j] (DDPFeedback::k_, feedback_controllers/DDP/ddp.cuh:144-151; gains NULL = no feedback), u_r += fb_t, both
 *     constrained and stepped, then
 *       cost_r  = (sum_t [running(y_r) + lr_r(u_r)] + terminal(y_r)) / T
 *       track_r = (sum_t [running(y_r) + fb_cost(fb_t)] + terminal(y_r)) / T
 *       state_n = (sum_t running(y_n) + terminal(y_n)) / T,   lr_n = sum_t lr_n(u_n) / T
 *       cost_n  = 0.5 state_n + 0.5 max(min(track_r, vft), state_n) + lr_n.
 * parts (optional, [K][4]) receives (cost_r, track_r, state_n, lr_n) of RMPPI. */
template <class R, class Sys>
void two_system_loop(const Sys& sys, const mppi_sampler_params* smp, int nom, R vft, const float* gains, int K, int T,
                     R dt, R lambda, R alpha, const float* x0, const float* mean, float* samples, R* costs, R* parts,
                     int k_begin, int k_end)
{
  const int S = sys.state_dim(), C = sys.control_dim();
  const int pstart = pure_noise_start(K, smp[0].pure_noise_trajectories_percentage);
  const bool rmppi = nom >= 0;
  const int n = rmppi ? nom : 1, r = 1 - n;
  for (int k = k_begin; k < k_end; ++k) {
    R x[2][LOOP_MAX_DIM] = {{0}}, xn[2][LOOP_MAX_DIM] = {{0}}, xd[LOOP_MAX_DIM] = {0}, y[2][LOOP_MAX_DIM] = {{0}};
    R u[2][LOOP_MAX_DIM], fb[LOOP_MAX_DIM];
    for (int d = 0; d < 2; ++d)
      for (int i = 0; i < S; ++i) x[d][i] = x0[d * S + i];
    R acc[4] = {0, 0, 0, 0};
    const bool pure = k >= pstart;
    for (int t = 0; t < T; ++t) {
      float* ut[2];
      for (int d = 0; d < 2; ++d) {
        ut[d] = samples + (((size_t)d * K + k) * T + t) * C;
        for (int i = 0; i < C; ++i) u[d][i] = ut[d][i];
      }
      for (int j = 0; j < C; ++j) fb[j] = 0;
      if (rmppi && gains) {
        for (int i = 0; i < S; ++i)
          for (int j = 0; j < C; ++j) fb[j] += (R)gains[((size_t)t * S + i) * C + j] * (x[r][i] - x[n][i]);
        for (int j = 0; j < C; ++j) u[r][j] += fb[j];
      }
      for (int d = 0; d < 2; ++d) {
        sys.enforce(u[d]);
        for (int i = 0; i < C; ++i) ut[d][i] = (float)u[d][i];
        sys.step(x[d], xn[d], xd, u[d], y[d], dt);
      }
      const float* mt[2] = {mean + (size_t)t * C, mean + ((size_t)T + t) * C};
      if (!rmppi) {
        for (int d = 0; d < 2; ++d)
          acc[d] += sys.running(y[d], u[d], t) + lr_cost<R>(smp[d], C, u[d], mt[d], pure, lambda, alpha, t);
      } else {
        const R rc_r = sys.running(y[r], u[r], t);
        acc[0] += rc_r + lr_cost<R>(smp[r], C, u[r], mt[r], pure, lambda, alpha, t);
        acc[1] += rc_r + feedback_cost<R>(smp[r], C, fb, lambda, alpha, t);
        acc[2] += sys.running(y[n], u[n], t);
        acc[3] += lr_cost<R>(smp[n], C, u[n], mt[n], pure, lambda, alpha, t);
      }
      for (int d = 0; d < 2; ++d)
        for (int i = 0; i < S; ++i) x[d][i] = xn[d][i];
    }
    if (!rmppi) {
      for (int d = 0; d < 2; ++d) costs[(size_t)d * K + k] = acc[d] / (R)T + sys.terminal(y[d]) / (R)T;
    } else {
      const R cost_r = (acc[0] + sys.terminal(y[r])) / (R)T;
      const R track_r = (acc[1] + sys.terminal(y[r])) / (R)T;
      const R state_n = (acc[2] + sys.terminal(y[n])) / (R)T;
      const R lr_n = acc[3] / (R)T;
      costs[(size_t)r * K + k] = cost_r;
      costs[(size_t)n * K + k] = (R)0.5 * state_n + (R)0.5 * std::max(std::min(track_r, vft), state_n) + lr_n;
      if (parts) {
        parts[(size_t)k * 4 + 0] = cost_r;
        parts[(size_t)k * 4 + 1] = track_r;
        parts[(size_t)k * 4 + 2] = state_n;
        parts[(size_t)k * 4 + 3] = lr_n;
      }
    }
  }
}

/* RMPPI candidate evaluation (initEvalKernel, core/rmppi_kernels.cu:12-358; the CPU form of
 * tests/include/kernel_tests/core/rmppi_kernel_test.cu:73-127 with the GPU kernel's likelihood-ratio arguments, the
 * mean at t and the sample index i M + j): candidate i rolls out samples j < M from states[i] with the controls
 * samples[j][min(t + strides[i], T - 1)], constrained; costs[i M + j] = sum_t (running + lr) / T + terminal / T. */
template <class R, class Sys>
void candidate_eval_loop(const Sys& sys, const mppi_sampler_params& smp, int N, int M, int K, int T, R dt, R lambda,
                         R alpha, const float* states, const int* strides, const float* mean, const float* samples,
                         R* costs)
{
  const int S = sys.state_dim(), C = sys.control_dim();
  const int pstart = pure_noise_start(K, smp.pure_noise_trajectories_percentage);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < M; ++j) {
      R x[LOOP_MAX_DIM] = {0}, xn[LOOP_MAX_DIM] = {0}, xd[LOOP_MAX_DIM] = {0}, y[LOOP_MAX_DIM] = {0};
      R u[LOOP_MAX_DIM];
      for (int s = 0; s < S; ++s) x[s] = states[i * S + s];
      R run = 0;
      for (int t = 0; t < T; ++t) {
        const float* ut = samples + ((size_t)j * T + std::min(t + strides[i], T - 1)) * C;
        for (int c = 0; c < C; ++c) u[c] = ut[c];
        sys.enforce(u);
        sys.step(x, xn, xd, u, y, dt);
        run += sys.running(y, u, t) +
               lr_cost<R>(smp, C, u, mean + (size_t)t * C, i * M + j >= pstart, lambda, alpha, t);
        for (int s = 0; s < S; ++s) x[s] = xn[s];
      }
      costs[(size_t)i * M + j] = run / (R)T + sys.terminal(y) / (R)T;
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
