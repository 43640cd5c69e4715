// The unicycle of unicycle_model.hpp as a user plugin of the engine (test program): the Dynamics / Cost classes a
// user writes, the kernel-level plugin rollout, a C-ABI plugin controller, and the C++ VanillaMPPIController over the
// plugin types -- exported for the Python tests.
#define UNI_HD __host__ __device__
#include "unicycle_model.hpp"

#include <hip/hip_runtime.h>

#include <memory>

#include "../../include/mppi_amd/device_plugin.hpp"

class Unicycle : public mppi_amd::plugin::Dynamics<Unicycle, UnicycleParams, 3, 2, 4> {
 public:
  __host__ __device__ void computeStateDeriv(const float* x, const float* u, float* xd) const
  {
    unicycle_deriv<float>(params_, x, u, xd);
  }
  __host__ __device__ void updateState(const float* x, float* xn, const float* xd, float dt) const
  {
    unicycle_update<float>(x, xn, xd, dt);
  }
  __host__ __device__ void stateToOutput(const float* x, float* y) const { unicycle_output<float>(x, y); }
};

class UnicycleCost : public mppi_amd::plugin::Cost<UnicycleCost, UnicycleCostParams, 4> {
 public:
  __host__ __device__ float computeStateCost(const float* y, int t) const
  {
    return unicycle_state_cost<float>(params_, y, t);
  }
  __host__ __device__ float computeControlCost(const float* u, int t) const
  {
    return unicycle_control_cost<float>(params_, u, t);
  }
  __host__ __device__ float terminalCost(const float* y) const { return unicycle_terminal_cost<float>(params_, y); }
};

static void configure(const UnicycleSetup& s, Unicycle& d, UnicycleCost& c)
{
  d.setParams(s.dyn);
  for (int i = 0; i < 2; ++i) {
    d.control_rngs_[i] = {s.control_min[i], s.control_max[i]};
    d.control_deadband_[i] = s.control_deadband[i];
    d.zero_control_[i] = s.zero_control[i];
  }
  c.setParams(s.cost);
}

// the objects and callback tables of the controllers created through the C-ABI (kept for the process's lifetime)
struct PluginInstance {
  Unicycle dyn;
  UnicycleCost cost;
  std::unique_ptr<mppi_amd::plugin::Binding<Unicycle, UnicycleCost>> binding;
};

extern "C" {

// Kernel-level plugin rollout of device buffers; synchronises.
int unicycle_rollout_gpu(const UnicycleSetup* s, const mppi_sampler_params* sp, int K, int T, float dt, float lambda,
                         float alpha, const float* x0_d, const float* mean_d, float* samples_d, float* costs_d)
{
  Unicycle d;
  UnicycleCost c;
  configure(*s, d, c);
  const int st = mppi_amd::plugin::launch_rollout(d, c, *sp, K, T, dt, lambda, alpha, x0_d, mean_d, samples_d,
                                                  costs_d, nullptr);
  if (st != MPPI_OK) return st;
  return hipDeviceSynchronize() == hipSuccess ? MPPI_OK : MPPI_E_HIP;
}

// A C-ABI controller over the plugin (desc->dynamics: control limits and zero control are taken from the setup).
int unicycle_controller_create(const mppi_controller_desc* desc, const UnicycleSetup* s, mppi_controller** out)
{
  auto* inst = new PluginInstance();
  configure(*s, inst->dyn, inst->cost);
  inst->binding = mppi_amd::plugin::bind_plugin(&inst->dyn, &inst->cost);
  const mppi_plugin table = inst->binding->table();
  mppi_controller_desc d = *desc;
  d.dynamics = inst->dyn.desc();
  return mppi_controller_create_plugin(&d, &table, out);
}

// The plugin's host step and constraint (what the controller's state propagation calls).
void unicycle_host_step(const UnicycleSetup* s, const float* x, const float* u, float dt, float* xn)
{
  Unicycle d;
  UnicycleCost c;
  configure(*s, d, c);
  float uc[2] = {u[0], u[1]}, xd[3], y[4];
  d.enforceConstraints(uc);
  d.step(x, xn, xd, uc, y, dt);
}

// Closed loop through the C++ API: VanillaMPPIController<Unicycle, UnicycleCost, ...> (the reference's template,
// compiled here with the plugin types), driving the unicycle from x0 for `steps` control steps with the first
// control applied each step (dt, K = 2048, T = 60). Writes the final state and the last baseline cost.
int unicycle_cpp_closed_loop(const UnicycleSetup* s, const float* x0, int steps, float dt, float* x_final,
                             float* baseline)
{
  using DYN = Unicycle;
  using COST = UnicycleCost;
  using FB = mppi_amd::NoFeedback<DYN, 60>;
  using SAMPLER = mppi_amd::sampling_distributions::GaussianDistributionC<2>;
  using CTRL = mppi_amd::VanillaMPPIController<DYN, COST, FB, 60, 2048, SAMPLER>;
  try {
    DYN model;
    COST cost;
    configure(*s, model, cost);
    FB fb(&model, dt);
    SAMPLER::SAMPLING_PARAMS_T sp;
    sp.std_dev[0] = 0.6f;
    sp.std_dev[1] = 0.8f;
    SAMPLER sampler(sp);
    CTRL controller(&model, &cost, &fb, &sampler, dt, 1, /*lambda=*/0.05f, /*alpha=*/0.0f);
    DYN::state_array x;
    for (int i = 0; i < 3; ++i) x[i] = x0[i];
    for (int k = 0; k < steps; ++k) {
      controller.computeControl(x, 1);
      const auto u = controller.getControlSeq();
      float uc[2] = {u(0, 0), u(1, 0)}, xn[3], xd[3], y[4];
      model.enforceConstraints(uc);
      model.step(x.data(), xn, xd, uc, y, dt);
      for (int i = 0; i < 3; ++i) x[i] = xn[i];
      controller.slideControlSequence(1);
    }
    // the state trajectory of the last call, propagated on request through the plugin's host step
    const auto traj = controller.getTargetStateSeq();
    if (!(traj(0, 0) == traj(0, 0))) return MPPI_E_INVALID;
    for (int i = 0; i < 3; ++i) x_final[i] = x[i];
    *baseline = controller.getBaselineCost();
  } catch (const std::exception&) {
    return MPPI_E_INVALID;
  }
  return MPPI_OK;
}

}  // extern "C"
