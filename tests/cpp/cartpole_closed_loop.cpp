// Closed-loop cart-pole swing-up through the C++ plugin API, written as the reference's examples/cartpole_example.cu
// and tests/controllers/vanilla_mppi_test.cu:79-136 are: plugins, VanillaMPPIController, computeControl,
// getControlSeq, slideControlSequence. Prints the final baseline cost and state; exit status 0 when the swing-up
// reaches a baseline cost below 1.0.
#include <cstdio>

#include "mppi_amd/mppi.hpp"

using namespace mppi_amd;
using SAMPLER_T = sampling_distributions::GaussianDistribution<CartpoleDynamics::DYN_PARAMS_T>;

int main()
{
  auto model = new CartpoleDynamics(1.0, 1.0, 1.0);
  auto cost = new CartpoleQuadraticCost;
  CartpoleQuadraticCostParams new_params;
  new_params.cart_position_coeff = 100;
  new_params.pole_angle_coeff = 200;
  new_params.cart_velocity_coeff = 10;
  new_params.pole_angular_velocity_coeff = 20;
  new_params.terminal_cost_coeff = 0;
  new_params.desired_terminal_state[0] = -20;
  new_params.desired_terminal_state[1] = 0;
  new_params.desired_terminal_state[2] = M_PI;
  new_params.desired_terminal_state[3] = 0;
  cost->setParams(new_params);

  auto sampler_params = SAMPLER_T::SAMPLING_PARAMS_T();
  sampler_params.std_dev[0] = 5.0;
  sampler_params.control_cost_coeff[0] = 1.0;
  auto sampler = new SAMPLER_T(sampler_params);
  const float dt = 0.01f;
  const int num_timesteps = 100;
  auto fb = new NoFeedback<CartpoleDynamics, num_timesteps>(model, dt);
  auto controller = new VanillaMPPIController<CartpoleDynamics, CartpoleQuadraticCost,
                                              NoFeedback<CartpoleDynamics, num_timesteps>, num_timesteps, 2048>(
      model, cost, fb, sampler, dt, 1, 0.25f, 0.01f);
  auto params = controller->getParams();
  params.slide_control_scale_[0] = 1.0f;
  params.seed_ = 42;
  controller->setParams(params);

  CartpoleDynamics::state_array x = CartpoleDynamics::state_array::Zero();
  CartpoleDynamics::state_array xdot, xn;
  CartpoleDynamics::output_array y;
  for (int i = 0; i < 1000; ++i) {
    controller->computeControl(x, 1);
    CartpoleDynamics::control_array u = controller->getControlSeq().col(0);
    model->step(x, xn, xdot, u, y, i * dt, dt);
    x = xn;
    controller->slideControlSequence(1);
  }
  const float b = controller->getBaselineCost();
  std::printf("baseline %.6f state %.4f %.4f %.4f %.4f\n", b, x[0], x[1], x[2], x[3]);
  delete controller;
  delete fb;
  delete sampler;
  delete cost;
  delete model;
  return b < 1.0f ? 0 : 1;
}
