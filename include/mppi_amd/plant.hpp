// plant.hpp -- the control loop around an MPPI controller (header-only, over mppi.hpp).
//
// BasePlant<CONTROLLER_T> keeps the reference's plant interface (core/base_plant.hpp:25-633): a robot-side
// integration implements pubControl / pubNominalState / pubFreeEnergyStatistics / checkStatus / getCurrentTime /
// getPoseTime, feeds states with updateState(state, time), and runs runControlLoop (or runControlIteration) on a thread
// of its own. Between solutions updateState publishes the solution's control interpolated at the state's time
// (getCurrentControl, controllers/controller.cuh:347-404); each iteration picks up parameter changes queued from any
// thread (setDynamicsParams / setCostParams / setControllerParams / setSamplerParams), slides the previous solution
// by the robot time elapsed, and solves from the latest state. The parameter changes reach the running device
// controller through mppi_controller_set_params (no reallocation, graphs recaptured).
//
// Left out (robot-specific or other algorithms): the logger, the state buffer of buffered plants, the feedback
// controller (NoFeedback is the only one), visualisation publishing.
#pragma once

#include <atomic>
#include <chrono>
#include <cmath>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "mppi.hpp"

namespace mppi_amd {

template <class CONTROLLER_T>
class BasePlant {
 public:
  using c_array = typename CONTROLLER_T::control_array;
  using c_traj = typename CONTROLLER_T::control_trajectory;
  using s_array = typename CONTROLLER_T::state_array;
  using s_traj = typename CONTROLLER_T::state_trajectory;
  using o_traj = typename CONTROLLER_T::output_trajectory;
  using DYN_T = typename CONTROLLER_T::TEMPLATED_DYNAMICS;
  using DYN_PARAMS_T = typename DYN_T::DYN_PARAMS_T;
  using COST_T = typename CONTROLLER_T::TEMPLATED_COSTS;
  using COST_PARAMS_T = typename COST_T::COST_PARAMS_T;
  using CONTROLLER_PARAMS_T = typename CONTROLLER_T::TEMPLATED_PARAMS;
  using SAMPLER_PARAMS_T = typename CONTROLLER_T::TEMPLATED_SAMPLING_PARAMS;
  struct FB_STATE_T {};  // NoFeedback carries no state

  std::shared_ptr<CONTROLLER_T> controller_;

  // core/base_plant.hpp:126-141
  BasePlant(std::shared_ptr<CONTROLLER_T> controller, int hz, int optimization_stride)
    : controller_(std::move(controller)), hz_(hz), optimization_stride_(optimization_stride)
  {
    dynamics_params_ = controller_->model_->getParams();
    cost_params_ = controller_->cost_->getParams();
    controller_params_ = controller_->getParams();
    sampler_params_ = controller_->sampler_->getParams();
    control_traj_ = c_traj::Zero();
    state_traj_ = s_traj::Zero();
  }
  virtual ~BasePlant() = default;

  // ---------------------------------------------------------------------- the robot side (core/base_plant.hpp:146-183)
  virtual void pubControl(const c_array& u) = 0;
  virtual void pubNominalState(const s_array& s) = 0;
  virtual void pubFreeEnergyStatistics(MPPIFreeEnergyStatistics& fe_stats) = 0;
  // 0: nominal, 1: something is wrong but no action needed, 2: stop
  virtual int checkStatus() = 0;
  virtual double getCurrentTime() = 0;
  virtual double getPoseTime() = 0;
  virtual double getStateTime() { return state_time_; }

  // --------------------------------------------------------------------------------------------------- accessors
  s_traj getStateTraj()
  {
    std::lock_guard<std::mutex> lck(access_guard_);
    return state_traj_;
  }
  c_traj getControlTraj()
  {
    std::lock_guard<std::mutex> lck(access_guard_);
    return control_traj_;
  }
  virtual s_array getState()
  {
    std::lock_guard<std::mutex> lck(access_guard_);
    return state_;
  }
  virtual void setState(const s_array& state) { state_ = state; }
  virtual void setControl(const c_array& u) { u_ = u; }
  virtual void setDebugMode(bool mode) { debug_mode_ = mode; }
  void resetStateTime() { last_used_state_update_time_ = -1; }
  double getAvgOptimizationTime() const { return avg_optimize_time_ms_; }
  double getAvgLoopTime() const { return avg_loop_time_ms_; }
  double getAvgSleepTime() const { return avg_sleep_time_ms_; }
  double getLastOptimizationTime() const { return optimization_duration_; }
  int getNumIterations() const { return num_iter_; }
  int getStatus() const { return status_; }
  int getTargetOptimizationStride() const { return optimization_stride_; }
  int getLastOptimizationStride() const { return last_optimization_stride_; }
  void setTargetOptimizationStride(int v) { optimization_stride_ = v; }
  int getHz() const { return hz_; }
  void setHz(int hz) { hz_ = hz; }

  // core/base_plant.hpp:253-265
  virtual void setSolution(const s_traj& state_seq, const c_traj& control_seq, const o_traj& output_seq,
                           const FB_STATE_T& fb_state, double timestamp)
  {
    last_used_state_update_time_ = timestamp;
    std::lock_guard<std::mutex> guard(access_guard_);
    state_traj_ = state_seq;
    output_traj_ = output_seq;
    control_traj_ = control_seq;
    feedback_state_ = fb_state;
    num_iter_++;
  }

  // A new state from the estimator (core/base_plant.hpp:272-304): stores it and, once a solution exists and the
  // state's time falls inside it, publishes the solution's control interpolated at the time since that solution's
  // state (and the interpolated nominal state in debug mode).
  virtual void updateState(const s_array& state, double time)
  {
    const double last = last_used_state_update_time_;
    const double since = time - last;
    {
      std::lock_guard<std::mutex> lck(access_guard_);
      state_ = state;
    }
    state_time_ = time;
    has_received_state_ = true;
    if (num_iter_ == 0) return;
    const bool within = time >= last && time < last + controller_->getDt() * controller_->getNumTimesteps();
    if (since > 0 && within) {
      s_traj st;
      c_traj ct;
      FB_STATE_T fb;
      {
        std::lock_guard<std::mutex> lck(access_guard_);
        st = state_traj_;
        ct = control_traj_;
        fb = feedback_state_;
      }
      const s_array target = controller_->interpolateState(st, since);
      pubControl(controller_->getCurrentControl(state, since, target, ct, fb));
      if (debug_mode_) pubNominalState(target);
    }
  }

  // ------------------------------------------------------------- parameter hot-swap (core/base_plant.hpp:335-456)
  virtual bool hasNewDynamicsParams() { return has_new_dynamics_params_; }
  virtual bool hasNewCostParams() { return has_new_cost_params_; }
  virtual bool hasNewControllerParams() { return has_new_controller_params_; }
  virtual bool hasNewSamplerParams() { return has_new_sampler_params_; }
  virtual DYN_PARAMS_T getNewDynamicsParams(bool set_flag = false)
  {
    has_new_dynamics_params_ = set_flag;
    return dynamics_params_;
  }
  virtual COST_PARAMS_T getNewCostParams(bool set_flag = false)
  {
    has_new_cost_params_ = set_flag;
    return cost_params_;
  }
  virtual CONTROLLER_PARAMS_T getNewControllerParams(bool set_flag = false)
  {
    has_new_controller_params_ = set_flag;
    return controller_params_;
  }
  virtual SAMPLER_PARAMS_T getNewSamplerParams(bool set_flag = false)
  {
    has_new_sampler_params_ = set_flag;
    return sampler_params_;
  }
  virtual void setDynamicsParams(const DYN_PARAMS_T& p)
  {
    std::lock_guard<std::mutex> guard(dynamics_params_guard_);
    dynamics_params_ = p;
    has_new_dynamics_params_ = true;
  }
  virtual void setCostParams(const COST_PARAMS_T& p)
  {
    std::lock_guard<std::mutex> guard(cost_params_guard_);
    cost_params_ = p;
    has_new_cost_params_ = true;
  }
  virtual void setControllerParams(const CONTROLLER_PARAMS_T& p)
  {
    std::lock_guard<std::mutex> guard(controller_params_guard_);
    controller_params_ = p;
    has_new_controller_params_ = true;
  }
  virtual void setSamplerParams(const SAMPLER_PARAMS_T& p)
  {
    std::lock_guard<std::mutex> guard(sampler_params_guard_);
    sampler_params_ = p;
    has_new_sampler_params_ = true;
  }
  // Applies the queued parameter changes to the controller's plugins; the controller swaps them into its device side
  // on its next computeControl. Returns whether anything changed.
  virtual bool updateParameters()
  {
    bool changed = false;
    if (hasNewCostParams()) {
      std::lock_guard<std::mutex> guard(cost_params_guard_);
      changed = true;
      controller_->cost_->setParams(getNewCostParams());
    }
    if (hasNewDynamicsParams()) {
      std::lock_guard<std::mutex> guard(dynamics_params_guard_);
      changed = true;
      controller_->model_->setParams(getNewDynamicsParams());
    }
    if (hasNewControllerParams()) {
      std::lock_guard<std::mutex> guard(controller_params_guard_);
      changed = true;
      controller_->setParams(getNewControllerParams());
    }
    if (hasNewSamplerParams()) {
      std::lock_guard<std::mutex> guard(sampler_params_guard_);
      changed = true;
      controller_->setSamplingParams(getNewSamplerParams());
    }
    return changed;
  }

  // ------------------------------------------------------------------------- the loop (core/base_plant.hpp:466-633)
  // One optimisation: waits for a state newer than the last solution's, applies parameter changes, slides the
  // previous solution by the robot time elapsed (at least the target stride), solves from the latest state and
  // publishes the solution; keeps running averages of the optimisation and loop times (wall clock).
  void runControlIteration(std::atomic<bool>* is_alive)
  {
    const auto loop_start = std::chrono::steady_clock::now();
    if (!is_alive->load()) return;
    double last_state_time = getStateTime();
    const double last_used = last_used_state_update_time_;
    bool skip_first = num_iter_ == 0 && has_received_state_;
    while (last_used == last_state_time && !skip_first && is_alive->load()) {
      std::this_thread::sleep_for(std::chrono::microseconds(50));
      last_state_time = getStateTime();
      skip_first = num_iter_ == 0 && has_received_state_;
    }
    if (!is_alive->load()) return;
    updateParameters();
    s_array state;
    {
      std::lock_guard<std::mutex> lck(access_guard_);
      state = state_;
      last_state_time = state_time_;
    }
    for (int i = 0; i < s_array::rows(); ++i)
      if (!std::isfinite(state[i])) return;
    const int status = checkStatus();
    double dt = last_state_time - last_used;
    if (num_iter_ == 0) {
      dt = 0;
      last_optimization_stride_ = 0;
    } else {
      last_optimization_stride_ =
          std::max((int)std::round(dt / controller_->getDt()), optimization_stride_);
    }
    if (last_optimization_stride_ > 0 && last_optimization_stride_ < controller_->getNumTimesteps()) {
      controller_->updateImportanceSamplingControl(state, last_optimization_stride_);
      controller_->slideControlSequence(last_optimization_stride_);
    }
    const auto opt_start = std::chrono::steady_clock::now();
    controller_->computeControl(state, last_optimization_stride_);
    MPPIFreeEnergyStatistics fe = controller_->getFreeEnergyStatistics();
    const c_traj control_traj = controller_->getControlSeq();
    const s_traj state_traj = controller_->getTargetStateSeq();
    const o_traj output_traj = controller_->getTargetOutputSeq();
    for (int i = 0; i < c_traj::rows() * controller_->getNumTimesteps(); ++i)
      if (!std::isfinite(control_traj.data()[i])) throw std::runtime_error("control trajectory inside plant has a NaN");
    for (int i = 0; i < s_traj::rows() * controller_->getNumTimesteps(); ++i)
      if (!std::isfinite(state_traj.data()[i])) throw std::runtime_error("state trajectory inside plant has a NaN");
    optimization_duration_ = ms_since(opt_start);
    setSolution(state_traj, control_traj, output_traj, FB_STATE_T{}, last_state_time);
    status_ = status;
    pubFreeEnergyStatistics(fe);
    const double keep = (num_iter_ - 1.0) / num_iter_;
    avg_optimize_time_ms_ = keep * avg_optimize_time_ms_ + optimization_duration_ / num_iter_;
    optimize_loop_duration_ = ms_since(loop_start);
    avg_loop_time_ms_ = keep * avg_loop_time_ms_ + optimize_loop_duration_ / num_iter_;
  }

  // Runs iterations until *is_alive is cleared, each followed by a wait until the robot time has advanced by
  // optimization_stride / hz past the state the last solution used (parameter changes are applied while waiting).
  void runControlLoop(std::atomic<bool>* is_alive)
  {
    controller_->resetControls();
    while (is_alive->load()) {
      runControlIteration(is_alive);
      const double wait_until = last_used_state_update_time_ + (1.0 / hz_) * optimization_stride_;
      const auto sleep_start = std::chrono::steady_clock::now();
      while (is_alive->load() && wait_until > getStateTime()) {
        updateParameters();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
      }
      sleep_duration_ = ms_since(sleep_start);
      if (num_iter_ > 0) {
        const double keep = (num_iter_ - 1.0) / num_iter_;
        avg_sleep_time_ms_ = keep * avg_sleep_time_ms_ + sleep_duration_ / num_iter_;
      }
    }
  }

 protected:
  static double ms_since(std::chrono::steady_clock::time_point t0)
  {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  std::mutex access_guard_;
  int hz_ = 10;
  bool debug_mode_ = false;
  DYN_PARAMS_T dynamics_params_;
  std::mutex dynamics_params_guard_;
  COST_PARAMS_T cost_params_;
  std::mutex cost_params_guard_;
  CONTROLLER_PARAMS_T controller_params_;
  std::mutex controller_params_guard_;
  SAMPLER_PARAMS_T sampler_params_;
  std::mutex sampler_params_guard_;
  std::atomic<bool> has_new_dynamics_params_{false};
  std::atomic<bool> has_new_cost_params_{false};
  std::atomic<bool> has_new_controller_params_{false};
  std::atomic<bool> has_new_sampler_params_{false};
  std::atomic<bool> has_received_state_{false};
  s_array state_ = s_array::Zero();
  c_array u_ = c_array::Zero();
  s_traj state_traj_;
  c_traj control_traj_;
  o_traj output_traj_;
  FB_STATE_T feedback_state_;
  int optimization_stride_ = 1;
  int last_optimization_stride_ = 0;
  std::atomic<double> last_used_state_update_time_{0.0};
  std::atomic<double> state_time_{0.0};
  double optimize_loop_duration_ = 0, optimization_duration_ = 0, sleep_duration_ = 0;
  double avg_loop_time_ms_ = 0, avg_optimize_time_ms_ = 0, avg_sleep_time_ms_ = 0;
  std::atomic<int> num_iter_{0};
  int status_ = 1;
};

}  // namespace mppi_amd
