// Internal launch interface of the engine's HIP kernels (used by the C-ABI and the controller object).
#pragma once

#include "common.hpp"

namespace mppi_amd {

// Per-call arguments that change between MPC solves. The controller keeps them in device memory so that a captured
// hipGraph can be replayed with new values; the kernel-level C-ABI passes them by value.
struct IterArgs {
  uint64_t rng_offset;
  int32_t stride;
  int32_t iteration;
};

struct IterRef {
  const IterArgs* ptr;  // device pointer or nullptr
  IterArgs val;         // used when ptr == nullptr
};

int launch_generate_normals(uint64_t seed, uint64_t offset, float* out, int64_t n, hipStream_t s);
// With rows >= 0 only the first `rows` samples of the K-sample draw are written (the same values: the pure-noise
// share still follows K).
int launch_sample_gaussian(const mppi_sampler_params& sp, int K, int T, int C, const float* mean, float* samples,
                           IterRef it, const float* eps_in, hipStream_t s, int rows = -1);
int launch_sample_colored(const mppi_sampler_params& sp, int K, int T, int C, const float* mean, float* samples,
                          IterRef it, const float* spectra_in, hipStream_t s);
int launch_rollout(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp,
                   int K, int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                   float* samples, float* costs, hipStream_t s);
int launch_compute_weights(float* costs, int K, float lambda, mppi_weight_stats* stats, hipStream_t s);
// The weight transform into w_out (may be costs): normExp, or Tsallis when gamma != 0 and r != 0; the statistics also
// to stats_copy and the rng offset of *advance advanced by advance_step when those are given.
int launch_weights(const float* costs, float* w_out, int K, float lambda, float gamma, float r,
                   mppi_weight_stats* stats, mppi_weight_stats* stats_copy, IterArgs* advance, int advance_step,
                   hipStream_t s);
// Two-system rollout (Tube / RMPPI) over x0[2][S], mean[2][T][C], samples[2][K][T][C] (in place), costs[2][K].
int launch_two_system(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp0,
                      const mppi_sampler_params& sp1, const mppi_two_system_params& tp, int K, int T, float dt,
                      float lambda, float alpha, const float* x0, const float* mean, float* samples, float* costs,
                      hipStream_t s);
// RMPPI candidate evaluation: N candidates x the first M of K samples [K][T][C], each candidate from states[i][S] with
// its controls shifted by strides[i]; costs [N][M].
int launch_candidate_eval(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                          const mppi_sampler_params& sp, int N, int M, int K, int T, float dt, float lambda,
                          float alpha, const float* states, const int* strides, const float* mean,
                          const float* samples, float* costs, hipStream_t s);
// `ws` provides weighted_update_workspace_bytes(K, T, C) bytes of device memory.
size_t weighted_update_workspace_bytes(int K, int T, int C);
int launch_weighted_update(const float* w, const mppi_weight_stats* stats, const float* samples, int K, int T, int C,
                           float* out, void* ws, hipStream_t s);

// Re-rolls of n selected samples idx[n] of a [K][T][C] buffer of constrained samples (one lane each), recording
// controls [n][T][C], post-step outputs [n][T][O], per-step costs [n][T] and terminal costs [n], each / T.
int launch_vis_rollouts(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp,
                        int K, int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                        const float* samples, const int* idx, int n, float* controls, float* outputs,
                        float* step_costs, float* terminal, hipStream_t s);

// The small per-call exchanges of a control call in one launch: up to four spans of 32-bit words, each between
// device memory and mapped fine-grained host memory (device-side alias), read and written by the kernel itself.
struct WordSpan {
  const void* src;
  void* dst;
  int words;
};
constexpr int MAX_WORD_SPANS = 4;
// `sig` (optional): the host completion signal, given after the copies.
int launch_copy_words(const WordSpan* spans, int count, hipStream_t s, const HostSignal* sig = nullptr);

// 32-bit words of n objects of type Tp.
template <class Tp>
constexpr int words_of(int n)
{
  static_assert(sizeof(Tp) % 4 == 0, "copied in 32-bit words");
  return n * (int)(sizeof(Tp) / 4);
}

}  // namespace mppi_amd
