// Tube-MPPI and RMPPI through the C++ API on a double integrator pushed by a constant disturbance (a velocity bias the
// controllers' model does not know) towards a goal at (2, 2):
//   - TubeMPPIController: the actual system is controlled with getActualControlSeq's first control; the loop checks
//     that the state approaches the goal and that the nominal system was replaced by the actual one at least once;
//   - RobustMPPIController with PD feedback gains (GainFeedback), in the reference's order: updateImportanceSampling-
//     Control (the candidate evaluation that picks the nominal state and slides the nominal control), computeControl,
//     then the real system applies the real control plus getFeedbackControl(x, x_nominal, 0); the loop checks
//     getFeedbackControl against the PD formula, that the real system approaches the goal, that it stays within a tube
//     around the nominal one, and that the evaluation chose a candidate other than the first at some step.
// Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>

#include "mppi_amd/mppi.hpp"

using namespace mppi_amd;
constexpr int T_STEPS = 50;
using DYN = DoubleIntegratorDynamics;
using COST = QuadraticCost<DYN>;
using SAMPLER = sampling_distributions::GaussianDistributionC<DYN::CONTROL_DIM>;
using FB = GainFeedback<DYN, T_STEPS>;
using TUBE = TubeMPPIController<DYN, COST, FB, T_STEPS, 1024, SAMPLER>;
using RMPPI = RobustMPPIController<DYN, COST, FB, T_STEPS, 1024, SAMPLER>;

static int failures = 0;
static void check_that(bool ok, const char* what)
{
  std::printf("[tube_loop] %s: %s\n", what, ok ? "ok" : "FAILED");
  if (!ok) failures++;
}

static void setup_cost(COST& cost)
{
  COST::COST_PARAMS_T p;
  p.s_goal[0] = p.s_goal[1] = 2.0f;
  p.s_coeffs[0] = p.s_coeffs[1] = 5.0f;
  p.s_coeffs[2] = p.s_coeffs[3] = 0.5f;
  cost.setParams(p);
}

// the true plant: the model's step plus a constant velocity disturbance
static DYN::state_array plant_step(DYN& model, const DYN::state_array& x, const DYN::control_array& u, float dt)
{
  DYN::state_array xn, xd;
  DYN::output_array y;
  DYN::state_array xc = x;
  model.step(xc, xn, xd, u, y, 0, dt);
  xn[0] += 0.1f * dt;
  xn[1] -= 0.05f * dt;
  return xn;
}

int main()
{
  const float dt = 0.02f;
  {
    DYN model;
    COST cost;
    setup_cost(cost);
    SAMPLER sampler;
    FB fb(&model, dt);
    TUBE ctrl(&model, &cost, &fb, &sampler, dt, 1, 1.0f, 0.0f);
    ctrl.setNominalThreshold(2.0f);
    DYN::state_array x = DYN::state_array::Zero();
    int replaced = 0;
    for (int k = 0; k < 200; ++k) {
      ctrl.computeControl(x, 1);
      replaced += ctrl.getFreeEnergyStatistics().nominal_state_used == 0;
      const auto u = ctrl.getActualControlSeq();
      DYN::control_array u0;
      for (int c = 0; c < 2; ++c) u0[c] = u(c, 0);
      x = plant_step(model, x, u0, dt);
      ctrl.slideControlSequence(1);
    }
    const float dist = std::hypot(x[0] - 2.0f, x[1] - 2.0f);
    std::printf("[tube_loop] Tube: final distance %.3f, nominal replaced in %d of 200 calls, baselines %.3f / %.3f\n",
                dist, replaced, ctrl.getBaselineCost(0), ctrl.getBaselineCost(1));
    check_that(dist < 0.35f * std::hypot(2.0f, 2.0f), "Tube approaches the goal");
    check_that(replaced > 0, "Tube replaced the nominal system at least once");
  }
  {
    DYN model;
    COST cost;
    setup_cost(cost);
    SAMPLER sampler;
    FB fb(&model, dt);
    const float kp = 4.0f, kd = 2.0f;
    std::vector<float> gains((size_t)T_STEPS * DYN::STATE_DIM * DYN::CONTROL_DIM, 0.0f);
    for (int t = 0; t < T_STEPS; ++t) {
      float* g = gains.data() + (size_t)t * DYN::STATE_DIM * DYN::CONTROL_DIM;
      g[0 * 2 + 0] = -kp;  // (state px) -> control ax
      g[2 * 2 + 0] = -kd;  // (state vx) -> control ax
      g[1 * 2 + 1] = -kp;  // (state py) -> control ay
      g[3 * 2 + 1] = -kd;  // (state vy) -> control ay
    }
    fb.setGains(gains);
    RMPPI ctrl(&model, &cost, &fb, &sampler, dt, 1, 1.0f, 0.0f);
    ctrl.setValueFunctionThreshold(50.0f);
    DYN::state_array x = DYN::state_array::Zero(), xn;
    float max_gap = 0.0f;
    bool formula = true;
    int moved = 0;
    for (int k = 0; k < 200; ++k) {
      ctrl.updateImportanceSamplingControl(x, 1);
      moved += ctrl.getBestIndex() != 0;
      xn = ctrl.getNominalState();
      ctrl.computeControl(x, 1);
      const auto u = ctrl.getControlSeq();
      const auto fbu = ctrl.getFeedbackControl(x, xn, 0);
      const float want0 = -kp * (x[0] - xn[0]) - kd * (x[2] - xn[2]);
      const float want1 = -kp * (x[1] - xn[1]) - kd * (x[3] - xn[3]);
      formula = formula && std::fabs(fbu[0] - want0) <= 1e-5f * (1 + std::fabs(want0)) &&
                std::fabs(fbu[1] - want1) <= 1e-5f * (1 + std::fabs(want1));
      DYN::control_array ua;
      for (int c = 0; c < 2; ++c) ua[c] = u(c, 0) + fbu[c];
      x = plant_step(model, x, ua, dt);
      max_gap = std::fmax(max_gap, std::hypot(x[0] - xn[0], x[1] - xn[1]));
    }
    const float dist = std::hypot(x[0] - 2.0f, x[1] - 2.0f);
    std::printf("[tube_loop] RMPPI: final distance %.3f, largest real-nominal gap %.3f, candidate > 0 in %d calls\n",
                dist, max_gap, moved);
    check_that(moved > 0, "RMPPI candidate evaluation moved the nominal state along the line");
    check_that(formula, "getFeedbackControl is the PD law K_t (x - x_nominal)");
    check_that(dist < 0.35f * std::hypot(2.0f, 2.0f), "RMPPI real system approaches the goal");
    check_that(max_gap < 0.5f, "RMPPI real system stays in a tube around the nominal one");
  }
  std::printf("[tube_loop] %s\n", failures == 0 ? "all checks passed" : "some checks failed");
  return failures == 0 ? 0 : 1;
}
