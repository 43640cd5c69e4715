// Goal-trajectory tracking through the C++ API: a double integrator follows a point moving on the unit circle
// (1 rad/s), given to the controller as a QuadraticCostTrajectory goal table (cost_functions/quadratic_cost/
// quadratic_cost.cuh:12-63) whose current_time the loop advances every control step (setCurrentTime), with per-step
// standard deviations growing along the horizon (GaussianTimeVaryingStdDevParams, sampling_distributions/gaussian/
// gaussian.cuh:66-80). The same loop with current_time held at 0 is run for contrast: the controller then chases the
// first second of the circle forever. Exit status 0 when the advancing run tracks the circle (mean position error
// over the second half below 0.1) and the frozen one does not (above 0.3).
#include <cmath>
#include <cstdio>

#include "mppi_amd/mppi.hpp"

using namespace mppi_amd;
constexpr int T_STEPS = 50;
constexpr int HORIZON = 1200;  // 24 s of goals at dt = 0.02
using DYN = DoubleIntegratorDynamics;
using COST = QuadraticCostTrajectory<DYN, HORIZON>;
using SAMPLER = sampling_distributions::GaussianTimeVaryingDistributionC<DYN::CONTROL_DIM, T_STEPS>;
using FB = NoFeedback<DYN, T_STEPS>;
using CTRL = VanillaMPPIController<DYN, COST, FB, T_STEPS, 2048, SAMPLER>;

static double run(bool advance, const float dt, int steps)
{
  DYN model;
  COST cost;
  COST::COST_PARAMS_T cp;
  for (int k = 0; k < HORIZON; ++k) {
    const float t = k * dt;
    cp.s_goal[k * 4 + 0] = std::cos(t);
    cp.s_goal[k * 4 + 1] = std::sin(t);
    cp.s_goal[k * 4 + 2] = -std::sin(t);
    cp.s_goal[k * 4 + 3] = std::cos(t);
  }
  cp.s_coeffs[0] = cp.s_coeffs[1] = 10.0f;
  cp.s_coeffs[2] = cp.s_coeffs[3] = 1.0f;
  SAMPLER::SAMPLING_PARAMS_T sp;
  for (int t = 0; t < T_STEPS; ++t)
    for (int c = 0; c < DYN::CONTROL_DIM; ++c) sp.std_dev[t * DYN::CONTROL_DIM + c] = 0.5f + t / (T_STEPS - 1.0f);
  SAMPLER sampler(sp);
  FB fb(&model, dt);
  CTRL controller(&model, &cost, &fb, &sampler, dt, 1, 0.1f, 0.0f);
  auto params = controller.getParams();
  params.seed_ = 3;
  controller.setParams(params);

  DYN::state_array x = DYN::state_array::Zero(), xn, xd;
  DYN::output_array y;
  x[0] = 1.0f;
  double err_sum = 0.0;
  int err_n = 0;
  for (int k = 0; k < steps; ++k) {
    // the output after the rollout's step t is the state at k + t + 1: charge it against goal row k + t + 1
    cp.setCurrentTime(advance ? k + 1 : 0);
    cost.setParams(cp);
    controller.computeControl(x, 1);
    DYN::control_array u = controller.getControlSeq().col(0);
    model.step(x, xn, xd, u, y, k * dt, dt);
    x = xn;
    controller.slideControlSequence(1);
    if (k >= steps / 2) {
      err_sum += std::hypot(x[0] - cp.s_goal[(k + 1) * 4 + 0], x[1] - cp.s_goal[(k + 1) * 4 + 1]);
      ++err_n;
    }
  }
  return err_sum / err_n;
}

int main()
{
  const float dt = 0.02f;
  const int steps = 600;
  const double tracked = run(true, dt, steps);
  const double frozen = run(false, dt, steps);
  std::printf("[tracking] mean position error over the second half: %.4f with current_time advancing, %.4f with it "
              "held at 0\n", tracked, frozen);
  const bool ok = tracked < 0.1 && frozen > 0.3;
  std::printf("[tracking] %s\n", ok ? "all checks passed" : "some checks failed");
  return ok ? 0 : 1;
}
