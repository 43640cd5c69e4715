// ColoredMPPIController through the C++ API (controllers/ColoredMPPI/colored_mppi_controller.cuh:39-200):
//   part 1: with gamma = 0 (normExp weights), no leash and unbounded controls, the colored controller is the vanilla
//     iteration: five open-loop calls of both controllers, same sampler and seed, give the same control sequences
//     bit for bit; with Tsallis weights (gamma 1, r 2) the sequences differ.
//   part 2: closed loop: the double integrator driven to the goal (2, 2) under power-law noise with the Tsallis
//     weights and the state leash on (0.5 on every state), 300 steps of 0.02 s.
// Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "mppi_amd/mppi.hpp"

using namespace mppi_amd;
constexpr int T_STEPS = 50;
using DYN = DoubleIntegratorDynamics;
using COST = QuadraticMapCost<DYN>;
using SAMPLER = sampling_distributions::ColoredNoiseDistributionC<DYN::CONTROL_DIM>;
using FB = NoFeedback<DYN, T_STEPS>;
using COLORED = ColoredMPPIController<DYN, COST, FB, T_STEPS, 1024, SAMPLER>;
using VANILLA = VanillaMPPIController<DYN, COST, FB, T_STEPS, 1024, SAMPLER>;

static COST::COST_PARAMS_T cost_params()
{
  COST::COST_PARAMS_T p;
  const float q[4] = {2.0f, 2.0f, 0.5f, 0.5f};
  for (int i = 0; i < 4; ++i) p.s_coeffs[i] = q[i];
  p.s_goal[0] = p.s_goal[1] = 2.0f;
  return p;
}

static SAMPLER::SAMPLING_PARAMS_T sampler_params()
{
  SAMPLER::SAMPLING_PARAMS_T sp;
  sp.exponents[0] = 1.0f;
  sp.exponents[1] = 2.0f;
  return sp;
}

template <class CTRL>
static void seed(CTRL& c, unsigned long long s)
{
  auto p = c.getParams();
  p.seed_ = s;
  c.setParams(p);
}

int main()
{
  int failures = 0;
  auto check = [&](bool ok, const char* what) {
    std::printf("[colored] %s: %s\n", what, ok ? "ok" : "FAILED");
    if (!ok) failures++;
  };
  const float dt = 0.02f;
  DYN model;
  COST cost;
  cost.setParams(cost_params());
  SAMPLER sampler(sampler_params());
  FB fb(&model, dt);

  // ------------------------------------------------------------------------------------------------ part 1
  {
    COLORED colored(&model, &cost, &fb, &sampler, dt, 1, 1.0f, 0.0f);
    VANILLA vanilla(&model, &cost, &fb, &sampler, dt, 1, 1.0f, 0.0f);
    COLORED tsallis(&model, &cost, &fb, &sampler, dt, 1, 1.0f, 0.0f);
    tsallis.setGamma(1.0f);
    tsallis.setRExp(2.0f);
    seed(colored, 11);
    seed(vanilla, 11);
    seed(tsallis, 11);
    DYN::state_array x = DYN::state_array::Zero();
    bool same = true, differs = false;
    for (int k = 0; k < 5; ++k) {
      colored.computeControl(x, 1);
      vanilla.computeControl(x, 1);
      tsallis.computeControl(x, 1);
      const auto a = colored.getControlSeq(), b = vanilla.getControlSeq(), t = tsallis.getControlSeq();
      same = same && std::memcmp(a.data(), b.data(), sizeof(float) * DYN::CONTROL_DIM * T_STEPS) == 0;
      differs = differs || std::memcmp(a.data(), t.data(), sizeof(float) * DYN::CONTROL_DIM * T_STEPS) != 0;
    }
    std::printf("[colored] %s vs %s: baselines %.6f / %.6f, Tsallis baseline %.6f normalizer %.3f\n",
                colored.getControllerName().c_str(), vanilla.getControllerName().c_str(), colored.getBaselineCost(),
                vanilla.getBaselineCost(), tsallis.getBaselineCost(), tsallis.getNormalizerCost());
    check(same, "gamma 0, no leash: the same control sequences as the vanilla controller");
    check(differs, "Tsallis weights change the sequences");
  }

  // ------------------------------------------------------------------------------------------------ part 2
  {
    COLORED ctrl(&model, &cost, &fb, &sampler, dt, 1, 1.0f, 0.0f);
    ctrl.setGamma(1.0f);
    ctrl.setRExp(2.0f);
    ctrl.setLeashActive(true);
    for (int i = 0; i < DYN::STATE_DIM; ++i) ctrl.setStateLeashLength(0.5f, i);
    seed(ctrl, 5);
    DYN::state_array x = DYN::state_array::Zero(), xn, xd;
    DYN::output_array y;
    for (int k = 0; k < 300; ++k) {
      ctrl.computeControl(x, 1);
      DYN::control_array u = ctrl.getControlSeq().col(0);
      model.step(x, xn, xd, u, y, k * dt, dt);
      x = xn;
      ctrl.slideControlSequence(1);
    }
    const float dist = std::hypot(x[0] - 2.0f, x[1] - 2.0f);
    std::printf("[colored] Tsallis + leash closed loop: state %.3f %.3f %.3f %.3f, %.4f from the goal\n", x[0], x[1],
                x[2], x[3], dist);
    check(dist < 0.1f, "reached the goal");
  }
  std::printf("[colored] %s\n", failures == 0 ? "all checks passed" : "some checks failed");
  return failures == 0 ? 0 : 1;
}
