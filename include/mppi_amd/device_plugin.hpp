// device_plugin.hpp -- user-written device plugins for the MI355X MPPI engine (compile with hipcc for gfx950).
//
// The reference's users derive their own models and costs from the CRTP bases Dynamics<CLASS_T, PARAMS_T>
// (dynamics/dynamics.cuh:68-522) and Cost<CLASS_T, PARAMS_T, DYN_PARAMS_T> (cost_functions/cost.cuh:34-234) and write
// __host__ __device__ methods that the rollout kernel calls. The same is done here:
//
//   struct MyParams { float k; };
//   class MyDynamics : public mppi_amd::plugin::Dynamics<MyDynamics, MyParams, /*S*/ 3, /*C*/ 2, /*O*/ 4> {
//    public:
//     __host__ __device__ void computeStateDeriv(const float* x, const float* u, float* x_dot) const;
//     __host__ __device__ void stateToOutput(const float* x, float* y) const;        // optional (default: copy)
//     __host__ __device__ void updateState(const float* x, float* xn, const float* xd, float dt) const;  // optional
//   };
//   class MyCost : public mppi_amd::plugin::Cost<MyCost, MyCostParams, /*O*/ 4> {
//    public:
//     __host__ __device__ float computeStateCost(const float* y, int t) const;
//     __host__ __device__ float terminalCost(const float* y) const;                     // optional (default: 0)
//   };
//
// The classes are passed to the kernels by value: no virtual functions, no pointers to host memory. Every method
// a derived class defines hides the base's default (they are called on the final type). The engine's kernels
// (mppi_generic_vendor_amd/csrc/plugin_kernels.hpp) are templates over them and get instantiated in the user's
// translation unit: the controller (mppi_amd::VanillaMPPIController<MyDynamics, MyCost, ...> of mppi.hpp, or
// mppi_controller_create_plugin with bind_plugin()) then runs its sampling, graphs, weights, combine and host
// post-processing around them; mppi_amd::plugin::launch_rollout is the kernel-level rollout.
#pragma once

#include <cfloat>
#include <memory>

#include "../../mppi_generic_vendor_amd/csrc/plugin_kernels.hpp"
#include "mppi.hpp"

namespace mppi_amd {
namespace plugin {

// Dynamics<CLASS_T, PARAMS_T> with the dimensions as template arguments (the reference takes them from PARAMS_T's
// enums, dynamics/dynamics.cuh:75-77). control_rngs_ / control_deadband_ / zero_control_ are the reference's public
// members (dynamics/dynamics.cuh:130,508-509).
template <class CLASS_T, class PARAMS_T, int S_DIM, int C_DIM, int O_DIM = S_DIM>
class Dynamics {
 public:
  static constexpr int STATE_DIM = S_DIM, CONTROL_DIM = C_DIM, OUTPUT_DIM = O_DIM;
  static constexpr bool DEVICE_PLUGIN = true;
  using DYN_PARAMS_T = PARAMS_T;
  using state_array = Mat<float, S_DIM, 1>;
  using control_array = Mat<float, C_DIM, 1>;
  using output_array = Mat<float, O_DIM, 1>;

  Dynamics()
  {
    for (int i = 0; i < C_DIM; ++i) {
      control_rngs_[i] = {-FLT_MAX, FLT_MAX};
      control_deadband_[i] = 0.0f;
      zero_control_[i] = 0.0f;
    }
  }

  PARAMS_T params_{};
  float2 control_rngs_[C_DIM];
  float control_deadband_[C_DIM];
  float zero_control_[C_DIM];

  void setParams(const PARAMS_T& p) { params_ = p; }
  PARAMS_T getParams() const { return params_; }
  void setControlRanges(const std::array<float2, C_DIM>& r)
  {
    for (int i = 0; i < C_DIM; ++i) control_rngs_[i] = r[i];
  }
  void setControlDeadbands(const std::array<float, C_DIM>& d)
  {
    for (int i = 0; i < C_DIM; ++i) control_deadband_[i] = d[i];
  }
  void GPUSetup() {}
  void freeCudaMem() {}
  void bindToStream(void*) {}

  // Euler step (dynamics/dynamics.cu:128-139 / dynamics.cuh:270-276)
  __host__ __device__ void updateState(const float* x, float* xn, const float* xd, float dt) const
  {
#pragma unroll
    for (int i = 0; i < S_DIM; ++i) xn[i] = x[i] + xd[i] * dt;
  }
  // the first min(S, O) states (Dynamics::stateToOutput default)
  __host__ __device__ void stateToOutput(const float* x, float* y) const
  {
#pragma unroll
    for (int i = 0; i < O_DIM; ++i) y[i] = i < S_DIM ? x[i] : 0.0f;
  }
  // enforceConstraints for control i (dynamics/dynamics.cu:107-126): deadband then clamp; sign(0) = 1
  __host__ __device__ float enforceControlConstraint(int i, float v) const
  {
    if (fabsf(v) < control_deadband_[i]) v = zero_control_[i];
    else v += control_deadband_[i] * -(v >= 0.0f ? 1.0f : -1.0f);
    return fminf(fmaxf(control_rngs_[i].x, v), control_rngs_[i].y);
  }
  __host__ __device__ void enforceConstraints(float* u) const
  {
#pragma unroll
    for (int i = 0; i < C_DIM; ++i) u[i] = derived().enforceControlConstraint(i, u[i]);
  }
  // Dynamics::enforceLeash (dynamics/dynamics.cuh:445-463), host side: the nominal value per state when it lies within
  // the leash of the true one, else the true value moved by the leash towards it
  void enforceLeash(const state_array& xt, const state_array& xn, const state_array& lv, state_array& out) const
  {
    for (int i = 0; i < S_DIM; ++i) {
      const float diff = xn[i] - xt[i];
      out[i] = lv[i] < fabsf(diff) ? xt[i] + fminf(fmaxf(diff, -lv[i]), lv[i]) : xn[i];
    }
  }
  // Dynamics::step (dynamics/dynamics.cuh:278-285): derivative, update, output
  __host__ __device__ void step(const float* x, float* xn, float* xd, const float* u, float* y, float dt) const
  {
    derived().computeStateDeriv(x, u, xd);
    derived().updateState(x, xn, xd, dt);
    derived().stateToOutput(xn, y);
  }

  // the POD block the controller keeps for the host side: dimensions, control limits, zero control
  mppi_dynamics_params desc() const
  {
    mppi_dynamics_params d;
    std::memset(&d, 0, sizeof(d));
    d.state_dim = S_DIM;
    d.control_dim = C_DIM;
    d.output_dim = O_DIM;
    for (int i = 0; i < C_DIM; ++i) {
      d.control_min[i] = control_rngs_[i].x;
      d.control_max[i] = control_rngs_[i].y;
      d.control_deadband[i] = control_deadband_[i];
      d.zero_control[i] = zero_control_[i];
    }
    return d;
  }

 protected:
  __host__ __device__ const CLASS_T& derived() const { return *static_cast<const CLASS_T*>(this); }
};

// Cost<CLASS_T, PARAMS_T, DYN_PARAMS_T>: computeRunningCost = computeStateCost + computeControlCost
// (cost_functions/cost.cu:39-53); the control cost is 0 by default, as in the reference (cost_functions/cost.cuh:
// 205-208), the sampler's likelihood-ratio term being the live control cost; the terminal cost is 0 by default.
template <class CLASS_T, class PARAMS_T, int O_DIM>
class Cost {
 public:
  static constexpr int OUTPUT_DIM = O_DIM;
  static constexpr bool DEVICE_PLUGIN = true;
  using COST_PARAMS_T = PARAMS_T;
  PARAMS_T params_{};
  void setParams(const PARAMS_T& p) { params_ = p; }
  PARAMS_T getParams() const { return params_; }
  void GPUSetup() {}
  void freeCudaMem() {}
  void bindToStream(void*) {}

  __host__ __device__ float computeControlCost(const float*, int) const { return 0.0f; }
  __host__ __device__ float terminalCost(const float*) const { return 0.0f; }
  __host__ __device__ float computeRunningCost(const float* y, const float* u, int t) const
  {
    return derived().computeStateCost(y, t) + derived().computeControlCost(u, t);
  }

 protected:
  __host__ __device__ const CLASS_T& derived() const { return *static_cast<const CLASS_T*>(this); }
};

// The callback table of mppi_controller_create_plugin for (DYN, COST): the binding holds pointers to the caller's
// objects, which must outlive the controller; parameter changes take effect when the controller is created again
// (the kernel arguments of captured graphs are fixed at capture).
template <class DYN, class COST>
struct Binding {
  const DYN* dyn;
  const COST* cost;

  static int enqueue(void* user, const mppi_plugin_iteration* it, int32_t* num_partials)
  {
    const Binding* b = static_cast<const Binding*>(user);
    return enqueue_iteration<DYN, COST>(*b->dyn, *b->cost, *it, num_partials);
  }
  static int host_step(void* user, const float* x, const float* u, float dt, float* x_next)
  {
    const Binding* b = static_cast<const Binding*>(user);
    float xd[DYN::STATE_DIM], y[DYN::OUTPUT_DIM];
    b->dyn->step(x, x_next, xd, u, y, dt);
    return MPPI_OK;
  }
  static int host_enforce(void* user, float* u)
  {
    static_cast<const Binding*>(user)->dyn->enforceConstraints(u);
    return MPPI_OK;
  }
  mppi_plugin table()
  {
    mppi_plugin p;
    std::memset(&p, 0, sizeof(p));
    p.state_dim = DYN::STATE_DIM;
    p.control_dim = DYN::CONTROL_DIM;
    p.output_dim = DYN::OUTPUT_DIM;
    p.user = this;
    p.enqueue_iteration = &Binding::enqueue;
    p.host_step = &Binding::host_step;
    p.host_enforce_constraints = &Binding::host_enforce;
    return p;
  }
};

template <class DYN, class COST>
std::unique_ptr<Binding<DYN, COST>> bind_plugin(const DYN* dyn, const COST* cost)
{
  return std::unique_ptr<Binding<DYN, COST>>(new Binding<DYN, COST>{dyn, cost});
}

}  // namespace plugin

// VanillaMPPIController's creation hook for plugin types (declared in mppi.hpp).
template <class DYN, class COST>
struct PluginControllerFactory {
  static mppi_controller* create(mppi_controller_desc d, const DYN* dyn, const COST* cost,
                                 std::shared_ptr<void>& keep)
  {
    auto binding = std::shared_ptr<plugin::Binding<DYN, COST>>(new plugin::Binding<DYN, COST>{dyn, cost});
    const mppi_plugin table = binding->table();
    d.dynamics = dyn->desc();
    mppi_controller* c = nullptr;
    check(mppi_controller_create_plugin(&d, &table, &c), "mppi_controller_create_plugin");
    keep = binding;
    return c;
  }
};

}  // namespace mppi_amd
