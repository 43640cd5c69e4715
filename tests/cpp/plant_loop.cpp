// The plant loop (include/mppi_amd/plant.hpp) around the cart-pole controller, the way a robot integration uses
// BasePlant (core/base_plant.hpp): a simulated robot feeds states through updateState and applies what pubControl
// publishes; the controller runs runControlIteration / runControlLoop.
//
//   part 1 (one thread, deterministic): the robot is simulated at twice the controller's rate, so every other state
//     arrives between two solution steps and pubControl must carry the solution's control linearly interpolated at
//     that time (checked against the interpolation of getControlTraj, constrained); the pole swings up to the goal
//     cart position 0, then a cost-parameter change queued with setCostParams moves the goal to 1.0 m, which the
//     controller picks up in its next iteration without being recreated.
//   part 2 (two threads): runControlLoop at 50 Hz against a robot thread publishing states at 100 Hz of robot time;
//     a dynamics-parameter change (pole mass) queued from the robot thread is applied by the loop.
// Prints the host-inclusive optimisation rate through the plant; exit status 0 when every check holds.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <memory>
#include <thread>

#include "mppi_amd/plant.hpp"

using namespace mppi_amd;
constexpr int T_STEPS = 100;
using FB = NoFeedback<CartpoleDynamics, T_STEPS>;
using SAMPLER_T = sampling_distributions::GaussianDistribution<CartpoleDynamics::DYN_PARAMS_T>;
using CTRL = VanillaMPPIController<CartpoleDynamics, CartpoleQuadraticCost, FB, T_STEPS, 2048, SAMPLER_T>;

class SimPlant : public BasePlant<CTRL> {
 public:
  using BasePlant<CTRL>::BasePlant;
  std::atomic<int> published{0};
  c_array last_u = c_array::Zero();
  std::mutex u_guard;
  double max_interp_err = 0.0;
  bool check_interp = false;
  double sim_time = 0.0;

  void pubControl(const c_array& u) override
  {
    if (check_interp) {
      // the solution's control at the time since the solution's state, constrained (controllers/controller.cuh:
      // 347-394), recomputed from the published trajectory
      const double rel = getStateTime() - last_used_state_update_time_;
      const c_traj ct = getControlTraj();
      const double dt = controller_->getDt();
      const int lo = std::min((int)(rel / dt), T_STEPS - 1), hi = std::min(lo + 1, T_STEPS - 1);
      const double a = (rel - lo * dt) / dt;
      float want = (float)((1.0 - a) * ct(0, lo) + a * ct(0, hi));
      CartpoleDynamics::state_array z = CartpoleDynamics::state_array::Zero();
      c_array w;
      w[0] = want;
      controller_->model_->enforceConstraints(z, w);
      max_interp_err = std::max(max_interp_err, (double)std::fabs(w[0] - u[0]));
    }
    std::lock_guard<std::mutex> g(u_guard);
    last_u = u;
    published++;
  }
  void pubNominalState(const s_array&) override {}
  void pubFreeEnergyStatistics(MPPIFreeEnergyStatistics&) override {}
  int checkStatus() override { return 0; }
  double getCurrentTime() override { return sim_time; }
  double getPoseTime() override { return getStateTime(); }
  c_array control()
  {
    std::lock_guard<std::mutex> g(u_guard);
    return last_u;
  }
};

static CartpoleQuadraticCostParams cost_params(float goal_x)
{
  CartpoleQuadraticCostParams p;
  // the reference's swing-up test weights (tests/controllers/vanilla_mppi_test.cu:79-136)
  p.cart_position_coeff = 100;
  p.pole_angle_coeff = 200;
  p.cart_velocity_coeff = 10;
  p.pole_angular_velocity_coeff = 20;
  p.terminal_cost_coeff = 0;
  p.desired_terminal_state[0] = goal_x;
  p.desired_terminal_state[1] = 0;
  p.desired_terminal_state[2] = M_PI;
  p.desired_terminal_state[3] = 0;
  return p;
}

int main()
{
  int failures = 0;
  auto check = [&](bool ok, const char* what) {
    std::printf("[plant] %s: %s\n", what, ok ? "ok" : "FAILED");
    if (!ok) failures++;
  };
  const float dt = 0.01f;
  auto model = std::make_shared<CartpoleDynamics>(1.0f, 1.0f, 1.0f);
  auto cost = std::make_shared<CartpoleQuadraticCost>();
  cost->setParams(cost_params(0.0f));
  SAMPLER_T::SAMPLING_PARAMS_T sp;
  sp.std_dev[0] = 5.0f;
  sp.control_cost_coeff[0] = 1.0f;
  auto sampler = std::make_shared<SAMPLER_T>(sp);
  auto fb = std::make_shared<FB>(model.get(), dt);

  // ------------------------------------------------------------------------------------------------ part 1
  {
    auto controller = std::make_shared<CTRL>(model.get(), cost.get(), fb.get(), sampler.get(), dt, 1, 0.25f, 0.01f);
    auto params = controller->getParams();
    params.slide_control_scale_[0] = 1.0f;
    params.seed_ = 7;
    controller->setParams(params);
    SimPlant plant(controller, 100, 1);
    plant.check_interp = true;
    std::atomic<bool> alive{true};
    CartpoleDynamics::state_array x = CartpoleDynamics::state_array::Zero(), xn, xd;
    CartpoleDynamics::output_array y;
    const double h = dt / 2;  // the robot runs at twice the controller's rate
    double t = 0.0;
    for (int k = 0; k < 4400; ++k) {
      plant.sim_time = t;
      plant.updateState(x, t);
      if (k % 2 == 0) plant.runControlIteration(&alive);
      if (k == 2400) plant.setCostParams(cost_params(1.0f));
      CartpoleDynamics::control_array u = plant.control();
      model->step(x, xn, xd, u, y, (float)t, (float)h);
      x = xn;
      t += h;
      if (k == 2399)
        std::printf("[plant] goal 0: state %.3f %.3f %.3f %.3f after %d solutions\n", x[0], x[1], x[2], x[3],
                    plant.getNumIterations());
    }
    std::printf("[plant] goal 1: state %.3f %.3f %.3f %.3f; %d solutions, %d published controls\n", x[0], x[1],
                x[2], x[3], plant.getNumIterations(), plant.published.load());
    std::printf("[plant] interpolated control max |error| %.3g; optimisation %.4f ms avg (%.0f Hz host-inclusive), "
                "iteration %.4f ms avg\n", plant.max_interp_err, plant.getAvgOptimizationTime(),
                1000.0 / plant.getAvgOptimizationTime(), plant.getAvgLoopTime());
    check(plant.max_interp_err < 1e-5, "published controls are the interpolated solution");
    check(std::fabs(x[0] - 1.0f) < 0.25f, "cart moved to the new goal after the cost-parameter swap");
    const float ang = std::fmod(std::fabs(x[2]), 2.0f * (float)M_PI);
    check(std::fabs(ang - (float)M_PI) < 0.3f, "pole upright");
    check(controller->cost_->getParams().desired_terminal_state[0] == 1.0f, "cost parameters applied");
  }

  // ------------------------------------------------------------------------------------------------ part 2
  {
    model->setParams(CartpoleDynamicsParams(1.0f, 1.0f, 1.0f));
    cost->setParams(cost_params(0.0f));
    auto controller = std::make_shared<CTRL>(model.get(), cost.get(), fb.get(), sampler.get(), dt, 1, 0.25f, 0.01f);
    auto params = controller->getParams();
    params.slide_control_scale_[0] = 1.0f;
    controller->setParams(params);
    SimPlant plant(controller, 50, 1);
    std::atomic<bool> alive{true};
    std::atomic<double> robot_time{0.0};
    std::thread robot([&]() {
      // the robot's own model: the controller's is updated by the loop thread
      CartpoleDynamics robot_model(1.0f, 1.0f, 1.0f);
      CartpoleDynamics::state_array x = CartpoleDynamics::state_array::Zero(), xn, xd;
      CartpoleDynamics::output_array y;
      const double h = 0.01;
      for (int k = 0; k < 300; ++k) {
        plant.sim_time = k * h;
        plant.updateState(x, k * h);
        if (k == 150) {
          plant.setDynamicsParams(CartpoleDynamicsParams(1.0f, 0.8f, 1.0f));
          robot_model.setParams(CartpoleDynamicsParams(1.0f, 0.8f, 1.0f));
        }
        CartpoleDynamics::control_array u = plant.control();
        robot_model.step(x, xn, xd, u, y, (float)(k * h), (float)h);
        x = xn;
        robot_time = k * h;
        std::this_thread::sleep_for(std::chrono::milliseconds(2));  // robot time runs 5x faster than wall time
      }
      alive = false;
    });
    std::thread loop([&]() { plant.runControlLoop(&alive); });
    robot.join();
    loop.join();
    std::printf("[plant] threaded loop: %d solutions over %.2f s of robot time, optimisation %.4f ms avg "
                "(%.0f Hz), waiting %.4f ms avg\n", plant.getNumIterations(), robot_time.load(),
                plant.getAvgOptimizationTime(), 1000.0 / plant.getAvgOptimizationTime(), plant.getAvgSleepTime());
    check(plant.getNumIterations() >= 60, "the loop kept up with the robot (about one solution per 0.02 s)");
    check(std::fabs(controller->model_->getParams().pole_mass - 0.8f) < 1e-6f,
          "dynamics parameters queued from the robot thread were applied");
  }
  std::printf("[plant] %s\n", failures == 0 ? "all checks passed" : "some checks failed");
  return failures == 0 ? 0 : 1;
}
