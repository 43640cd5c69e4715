// Re-rolls of selected samples for the sampled-trajectory export (the visualisation rollouts of the reference,
// core/mppi_common.cu:371-691, behind copySampledControlFromDevice / copyTopControlFromDevice /
// calculateSampledStateTrajectories, controllers/controller.cu:57-181,287-354): one lane per selected sample reads its
// stored (constrained) controls from the sample buffer -- one gather launch, no per-sample copies -- and rolls it out
// from x0 around the mean the samples were drawn around, recording for every step t its control, its post-step output
// y_{t+1} and its cost (running(y_{t+1}) + likelihood-ratio(u_t)) / T, and the terminal cost / T; the per-step costs
// and the terminal cost add up to the sample's rollout cost.
#include "kernels.hpp"
#include "models.hpp"
#include "nn_bf16.hpp"

namespace mppi_amd {

template <class M, int COST>
__global__ void __launch_bounds__(64) vis_rollout_kernel(mppi_dynamics_params dyn, mppi_cost_params cost,
                                                         mppi_sampler_params sp, int T, float dt, float lambda,
                                                         float alpha, const float* __restrict__ x0,
                                                         const float* __restrict__ mean,
                                                         const float* __restrict__ samples,
                                                         const int* __restrict__ idx, int n, int pstart,
                                                         float* __restrict__ controls, float* __restrict__ outputs,
                                                         float* __restrict__ step_costs, float* __restrict__ terminal)
{
  using CF = Cost<COST, M::O>;
  constexpr int S = M::S, C = M::C, O = M::O;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int k = idx[j];
  const bool pure = k >= pstart;
  float x[S], xd[S], xn[S], u[C];
#pragma unroll
  for (int i = 0; i < S; ++i) x[i] = x0[i];
  const float* row = samples + (size_t)k * T * C;
  const float invT = 1.0f / (float)T;
  for (int t = 0; t < T; ++t) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      u[c] = row[t * C + c];
      controls[((size_t)j * T + t) * C + c] = u[c];
    }
    M::deriv(dyn, x, u, xd);
    M::update(dyn, x, xn, xd, dt);
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = xn[i];
#pragma unroll
    for (int i = 0; i < O; ++i) outputs[((size_t)j * T + t) * O + i] = x[i];
    step_costs[(size_t)j * T + t] =
        (CF::running(cost, x, t) + lr_cost<C>(sp, u, mean + t * C, pure, lambda, alpha, t)) * invT;
  }
  terminal[j] = CF::terminal(cost, x) * invT;
}

template <class M, int COST>
static int launch_vis_t(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp,
                        int K, int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                        const float* samples, const int* idx, int n, float* controls, float* outputs,
                        float* step_costs, float* terminal, hipStream_t s)
{
  const int pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  vis_rollout_kernel<M, COST><<<(n + 63) / 64, 64, 0, s>>>(dyn, cost, sp, T, dt, lambda, alpha, x0, mean, samples,
                                                           idx, n, pstart, controls, outputs, step_costs, terminal);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

int launch_vis_rollouts(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp,
                        int K, int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                        const float* samples, const int* idx, int n, float* controls, float* outputs,
                        float* step_costs, float* terminal, hipStream_t s)
{
  if (n < 1) return MPPI_OK;
#define MPPI_VIS_ARGS dyn, cost, sp, K, T, dt, lambda, alpha, x0, mean, samples, idx, n, controls, outputs, step_costs, \
                      terminal, s
  if (dyn.model == MPPI_MODEL_CARTPOLE && cost.cost == MPPI_COST_CARTPOLE_QUADRATIC)
    return launch_vis_t<Model<MPPI_MODEL_CARTPOLE>, MPPI_COST_CARTPOLE_QUADRATIC>(MPPI_VIS_ARGS);
  if (cost.cost == MPPI_COST_QUADRATIC_MAP) {
    switch (dyn.model) {
      case MPPI_MODEL_DOUBLE_INTEGRATOR:
        return launch_vis_t<Model<MPPI_MODEL_DOUBLE_INTEGRATOR>, MPPI_COST_QUADRATIC_MAP>(MPPI_VIS_ARGS);
      case MPPI_MODEL_RACER_DUBINS:
        return launch_vis_t<Model<MPPI_MODEL_RACER_DUBINS>, MPPI_COST_QUADRATIC_MAP>(MPPI_VIS_ARGS);
      case MPPI_MODEL_RACER_DUBINS_NN:
        if (dyn.nn_precision == MPPI_NN_BF16)
          return launch_vis_t<BicycleNNBf16Lane, MPPI_COST_QUADRATIC_MAP>(MPPI_VIS_ARGS);
        return launch_vis_t<Model<MPPI_MODEL_RACER_DUBINS_NN>, MPPI_COST_QUADRATIC_MAP>(MPPI_VIS_ARGS);
      default:
        break;
    }
  }
#undef MPPI_VIS_ARGS
  return fail(MPPI_E_UNSUPPORTED, "sampled trajectories: model/cost combination not compiled");
}

}  // namespace mppi_amd
