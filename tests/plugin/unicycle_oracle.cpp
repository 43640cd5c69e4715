// CPU side of the unicycle plugin test (test program, compiled with g++ -ffp-contract=off): the oracle's rollout loop
// (oracle/rollout_loop.hpp, the restated reference loop that checks the built-in models) over the unicycle's host
// formulas, in FP32 and FP64, and its state propagation.
#include <cstring>

#include "../../oracle/rollout_loop.hpp"
#include "unicycle_model.hpp"

namespace {

mppi_dynamics_params constraints_of(const UnicycleSetup& s)
{
  mppi_dynamics_params p;
  std::memset(&p, 0, sizeof(p));
  p.state_dim = 3;
  p.control_dim = 2;
  p.output_dim = 4;
  for (int i = 0; i < 2; ++i) {
    p.control_min[i] = s.control_min[i];
    p.control_max[i] = s.control_max[i];
    p.control_deadband[i] = s.control_deadband[i];
    p.zero_control[i] = s.zero_control[i];
  }
  return p;
}

template <class R>
struct UnicycleSystem {
  const UnicycleSetup& s;
  mppi_dynamics_params lim;
  int state_dim() const { return 3; }
  int control_dim() const { return 2; }
  int output_dim() const { return 4; }
  void enforce(R* u) const { oracle::enforce_constraints<R>(lim, u); }
  void step(const R* x, R* xn, R* xd, const R* u, R* y, R dt) const
  {
    unicycle_deriv<R>(s.dyn, x, u, xd);
    unicycle_update<R>(x, xn, xd, dt);
    unicycle_output<R>(xn, y);
  }
  R running(const R* y, const R* u, int t) const
  {
    return unicycle_state_cost<R>(s.cost, y, t) + unicycle_control_cost<R>(s.cost, u, t);
  }
  R terminal(const R* y) const { return unicycle_terminal_cost<R>(s.cost, y); }
};

}  // namespace

extern "C" {

void unicycle_rollout_oracle_f32(const UnicycleSetup* s, const mppi_sampler_params* sp, int K, int T, float dt,
                                 float lambda, float alpha, const float* x0, const float* mean, float* samples,
                                 float* costs, int nthreads)
{
  const UnicycleSystem<float> sys{*s, constraints_of(*s)};
  oracle::rollout_loop_threads<float>(sys, *sp, K, T, dt, lambda, alpha, x0, mean, samples, costs, nthreads);
}

void unicycle_rollout_oracle_f64(const UnicycleSetup* s, const mppi_sampler_params* sp, int K, int T, double dt,
                                 double lambda, double alpha, const float* x0, const float* mean, float* samples,
                                 double* costs, int nthreads)
{
  const UnicycleSystem<double> sys{*s, constraints_of(*s)};
  oracle::rollout_loop_threads<double>(sys, *sp, K, T, dt, lambda, alpha, x0, mean, samples, costs, nthreads);
}

// states [T][3]: x_0 = x0, x_{t+1} = step(x_t, constrained u_t)
void unicycle_state_trajectory_f32(const UnicycleSetup* s, const float* x0, const float* u, int T, float dt,
                                   float* states)
{
  const UnicycleSystem<float> sys{*s, constraints_of(*s)};
  for (int i = 0; i < 3; ++i) states[i] = x0[i];
  for (int t = 0; t + 1 < T; ++t) {
    float uc[2] = {u[2 * t], u[2 * t + 1]}, xd[3], y[4];
    sys.enforce(uc);
    sys.step(states + 3 * t, states + 3 * (t + 1), xd, uc, y, dt);
  }
}

}  // extern "C"
