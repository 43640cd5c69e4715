// mppi.hpp -- C++ template plugin API of the MI355X MPPI engine, header-only, over the C-ABI of mppi_amd.h.
//
// It keeps the reference's type and method names so that user code written against MPPI-Generic ports by
// changing includes and namespaces:
//   VanillaMPPIController<DYN_T, COST_T, FB_T, MAX_TIMESTEPS, NUM_ROLLOUTS, SAMPLING_T, PARAMS_T>
//       (controllers/MPPI/mppi_controller.cuh:14-41, controllers/controller.cuh:70-1058)
//   CartpoleDynamics, DoubleIntegratorDynamics, RacerDubins  (dynamics/*: Dynamics<CLASS_T, PARAMS_T>)
//   CartpoleQuadraticCost, QuadraticMapCost<DYN_T>          (cost_functions/*: Cost<CLASS_T, PARAMS_T, DYN_PARAMS_T>)
//   sampling_distributions::GaussianDistribution, ColoredNoiseDistribution  (sampling_distributions/*)
// Eigen is replaced by the fixed-size column-major Mat below (col(), data(), operator()), so `control_trajectory`
// holds the same bytes as the reference's Eigen::Matrix<float, C, T>.
//
// Device work never happens in these classes: the plugins are parameter holders whose `desc()` is a POD block of
// mppi_amd.h, and the kernel library instantiates the registered (model x cost x sampler) combinations. Errors from
// the C-ABI are thrown as std::runtime_error (the reference exits the process, utils/gpu_err_chk.cuh:32-40).
#pragma once

#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../mppi_amd.h"

namespace mppi_amd {

inline void check(int status, const char* what)
{
  if (status != MPPI_OK) throw std::runtime_error(std::string(what) + ": " + mppi_last_error());
}

struct float2 {
  float x, y;
};

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

// ------------------------------------------------------------------------------------------------ fixed matrices
template <class T, int R, int C>
class Mat {
 public:
  static constexpr int RowsAtCompileTime = R, ColsAtCompileTime = C;

  class ColRef {
   public:
    explicit ColRef(T* p) : p_(p) {}
    T& operator[](int i) { return p_[i]; }
    T& operator()(int i) { return p_[i]; }
    ColRef& operator=(const Mat<T, R, 1>& v)
    {
      std::memcpy(p_, v.data(), sizeof(T) * R);
      return *this;
    }
    operator Mat<T, R, 1>() const
    {
      Mat<T, R, 1> v;
      std::memcpy(v.data(), p_, sizeof(T) * R);
      return v;
    }
    T* data() { return p_; }

   private:
    T* p_;
  };

  Mat() { setZero(); }
  static Mat Zero() { return Mat(); }
  static Mat Ones()
  {
    Mat m;
    for (auto& v : m.d_) v = T(1);
    return m;
  }
  static Mat Constant(T v)
  {
    Mat m;
    for (auto& x : m.d_) x = v;
    return m;
  }
  void setZero()
  {
    for (auto& v : d_) v = T(0);
  }
  T& operator()(int r, int c) { return d_[c * R + r]; }
  const T& operator()(int r, int c) const { return d_[c * R + r]; }
  T& operator()(int i) { return d_[i]; }
  const T& operator()(int i) const { return d_[i]; }
  T& operator[](int i) { return d_[i]; }
  const T& operator[](int i) const { return d_[i]; }
  ColRef col(int c) { return ColRef(d_ + c * R); }
  Mat<T, R, 1> col(int c) const
  {
    Mat<T, R, 1> v;
    std::memcpy(v.data(), d_ + c * R, sizeof(T) * R);
    return v;
  }
  T* data() { return d_; }
  const T* data() const { return d_; }
  static constexpr int rows() { return R; }
  static constexpr int cols() { return C; }

 private:
  T d_[R * C];
};

// =================================================================================================== dynamics
// Dynamics<CLASS_T, PARAMS_T> (dynamics/dynamics.cuh:68-522): host methods run the engine's model code through
// mppi_host_step / mppi_host_enforce_constraints; the device copy is made by the controller.
template <int MODEL, int S, int C, int O>
class DynamicsBase {
 public:
  static const int STATE_DIM = S, CONTROL_DIM = C, OUTPUT_DIM = O;
  using state_array = Mat<float, S, 1>;
  using control_array = Mat<float, C, 1>;
  using output_array = Mat<float, O, 1>;

  DynamicsBase()
  {
    check(mppi_dynamics_defaults(MODEL, &desc_), "mppi_dynamics_defaults");
    for (int i = 0; i < C; ++i) {
      control_rngs_[i] = {desc_.control_min[i], desc_.control_max[i]};
      control_deadband_[i] = 0.0f;
    }
  }

  // Dynamics::control_rngs_ / control_deadband_ / zero_control_ are public in the reference (examples write them)
  float2 control_rngs_[C];
  float control_deadband_[C];
  control_array zero_control_ = control_array::Zero();

  void setControlRanges(const std::array<float2, C>& r)
  {
    for (int i = 0; i < C; ++i) control_rngs_[i] = r[i];
  }
  void setControlDeadbands(const std::array<float, C>& d)
  {
    for (int i = 0; i < C; ++i) control_deadband_[i] = d[i];
  }
  std::array<float2, C> getControlRanges() const
  {
    std::array<float2, C> r;
    for (int i = 0; i < C; ++i) r[i] = control_rngs_[i];
    return r;
  }
  state_array getZeroState() const { return state_array::Zero(); }

  void enforceConstraints(state_array&, control_array& u) const
  {
    const mppi_dynamics_params d = desc();
    check(mppi_host_enforce_constraints(&d, u.data()), "enforceConstraints");
  }
  void computeStateDeriv(const state_array& x, const control_array& u, state_array& xd) const
  {
    const mppi_dynamics_params d = host_desc();
    check(mppi_host_step(&d, x.data(), u.data(), 0.0f, nullptr, xd.data(), nullptr), "computeStateDeriv");
  }
  void updateState(const state_array& x, state_array& xn, const state_array& xd, float dt) const
  {
    for (int i = 0; i < S; ++i) xn[i] = x[i] + xd[i] * dt;
  }
  // Dynamics::enforceLeash (dynamics/dynamics.cuh:445-463; the bicycle's own, racer_dubins.cu:176-238)
  void enforceLeash(const state_array& state_true, const state_array& state_nominal, const state_array& leash_values,
                    state_array& state_output) const
  {
    const mppi_dynamics_params d = desc();
    check(mppi_host_enforce_leash(&d, state_true.data(), state_nominal.data(), leash_values.data(),
                                  state_output.data()),
          "enforceLeash");
  }
  void step(const state_array& x, state_array& xn, state_array& xd, const control_array& u, output_array& y,
            float /*t*/, float dt) const
  {
    const mppi_dynamics_params d = host_desc();
    check(mppi_host_step(&d, x.data(), u.data(), dt, xn.data(), xd.data(), y.data()), "step");
  }
  void initializeDynamics(const state_array&, const control_array&, output_array&, float, float) {}
  void GPUSetup() {}
  void freeCudaMem() {}
  void bindToStream(void*) {}

  // The POD block the kernel library consumes.
  mppi_dynamics_params desc() const
  {
    mppi_dynamics_params d = desc_;
    for (int i = 0; i < C; ++i) {
      d.control_min[i] = control_rngs_[i].x;
      d.control_max[i] = control_rngs_[i].y;
      d.control_deadband[i] = control_deadband_[i];
      d.zero_control[i] = zero_control_[i];
    }
    return d;
  }
  virtual const float* hostParameters() const { return nullptr; }
  virtual ~DynamicsBase() = default;

 protected:
  mppi_dynamics_params host_desc() const
  {
    mppi_dynamics_params d = desc();
    d.nn_theta = hostParameters();
    return d;
  }
  mppi_dynamics_params desc_{};
};

struct CartpoleDynamicsParams {
  float cart_mass = 1.0f;
  float pole_mass = 1.0f;
  float pole_length = 1.0f;
  CartpoleDynamicsParams() = default;
  CartpoleDynamicsParams(float m_c, float m_p, float l) : cart_mass(m_c), pole_mass(m_p), pole_length(l) {}
};

// dynamics/cartpole/cartpole_dynamics.cuh:40-101
class CartpoleDynamics : public DynamicsBase<MPPI_MODEL_CARTPOLE, 4, 1, 4> {
 public:
  using DYN_PARAMS_T = CartpoleDynamicsParams;
  explicit CartpoleDynamics(float cart_mass = 1.0f, float pole_mass = 1.0f, float pole_length = 1.0f,
                            void* /*stream*/ = nullptr)
  {
    setParams(CartpoleDynamicsParams(cart_mass, pole_mass, pole_length));
  }
  void setParams(const CartpoleDynamicsParams& p)
  {
    params_ = p;
    desc_.cart_mass = p.cart_mass;
    desc_.pole_mass = p.pole_mass;
    desc_.pole_length = p.pole_length;
  }
  CartpoleDynamicsParams getParams() const { return params_; }
  float getCartMass() const { return params_.cart_mass; }
  float getPoleMass() const { return params_.pole_mass; }
  float getPoleLength() const { return params_.pole_length; }
  float getGravity() const { return desc_.gravity; }

 private:
  CartpoleDynamicsParams params_;
};

struct DoubleIntegratorParams {};

// dynamics/double_integrator/di_dynamics.cuh
class DoubleIntegratorDynamics : public DynamicsBase<MPPI_MODEL_DOUBLE_INTEGRATOR, 4, 2, 4> {
 public:
  using DYN_PARAMS_T = DoubleIntegratorParams;
  explicit DoubleIntegratorDynamics(float /*system_noise*/ = 1.0f, void* /*stream*/ = nullptr) {}
};

// RacerDubinsParams (dynamics/racer_dubins/racer_dubins.cuh:81-103)
struct RacerDubinsParams {
  float c_t[3] = {1.3f, 2.6f, 3.9f};
  float c_b[3] = {2.5f, 3.5f, 4.5f};
  float c_v[3] = {3.7f, 4.7f, 5.7f};
  float c_0 = 4.9f;
  float steering_constant = 0.6f;
  float steer_command_angle_scale = 5.0f;
  float steer_angle_scale = -9.1f;
  float max_steer_angle = 0.5f;
  float max_steer_rate = 5.0f;
  float brake_delay_constant = 6.6f;
  float max_brake_rate_neg = 0.9f;
  float max_brake_rate_pos = 0.33f;
  float wheel_base = 0.3f;
  int gear_sign = 1;
};

template <int MODEL>
class RacerDubinsImpl : public DynamicsBase<MODEL, 7, 2, 7> {
 public:
  using DYN_PARAMS_T = RacerDubinsParams;
  RacerDubinsImpl() { setParams(RacerDubinsParams()); }
  void setParams(const RacerDubinsParams& p)
  {
    params_ = p;
    auto& d = this->desc_;
    d.c_t = p.c_t[0];
    d.c_b = p.c_b[0];
    d.c_v = p.c_v[0];
    d.c_0 = p.c_0;
    d.steering_constant = p.steering_constant;
    d.steer_command_angle_scale = p.steer_command_angle_scale;
    d.steer_angle_scale = p.steer_angle_scale;
    d.max_steer_angle = p.max_steer_angle;
    d.max_steer_rate = p.max_steer_rate;
    d.brake_delay_constant = p.brake_delay_constant;
    d.max_brake_rate_neg = p.max_brake_rate_neg;
    d.max_brake_rate_pos = p.max_brake_rate_pos;
    d.wheel_base = p.wheel_base;
    d.gear_sign = p.gear_sign;
  }
  RacerDubinsParams getParams() const { return params_; }

 protected:
  RacerDubinsParams params_;
};

// Kinematic bicycle (dynamics/racer_dubins/racer_dubins.cu).
class RacerDubins : public RacerDubinsImpl<MPPI_MODEL_RACER_DUBINS> {};

// Kinematic bicycle + 9-64-64-4 tanh residual MLP on (VEL_X, YAW, STEER_ANGLE, BRAKE_STATE) derivatives, the
// NeuralNetModel pattern of dynamics/autorally/ar_nn_model.cu with FNNHelper's parameter layout.
class RacerDubinsNN : public RacerDubinsImpl<MPPI_MODEL_RACER_DUBINS_NN> {
 public:
  RacerDubinsNN() : theta_(desc_.nn_num_params, 0.0f) {}
  // FNNHelper::updateModel(layers, theta) (utils/nn_helpers/fnn_helper.cu:207-246)
  void updateModel(const std::vector<int>& layers, const std::vector<float>& theta)
  {
    if (layers != std::vector<int>{9, 64, 64, 4})
      throw std::runtime_error("RacerDubinsNN: only the 9-64-64-4 network is compiled into the kernel library");
    if ((int)theta.size() != desc_.nn_num_params) throw std::runtime_error("RacerDubinsNN: parameter count mismatch");
    theta_ = theta;
  }
  // FNNHelper::loadParams(prefix, npz) (utils/nn_helpers/fnn_helper.cu:44-127) from an .npz file
  void loadParams(const std::string& model_path, const std::string& prefix = "")
  {
    int32_t nl = 0, np = 0;
    check(mppi_fnn_load_npz(model_path.c_str(), prefix.c_str(), 0, &nl, nullptr, nullptr, 0, &np), "loadParams");
    std::vector<int32_t> layers(nl);
    std::vector<float> theta(np);
    check(mppi_fnn_load_npz(model_path.c_str(), prefix.c_str(), nl, nullptr, layers.data(), theta.data(), np, nullptr),
          "loadParams");
    updateModel(std::vector<int>(layers.begin(), layers.end()), theta);
  }
  void setPrecision(int precision) { desc_.nn_precision = precision; }
  const std::vector<float>& getTheta() const { return theta_; }
  const float* hostParameters() const override { return theta_.data(); }

 private:
  std::vector<float> theta_;
};

// ======================================================================================================= costs
// CartpoleQuadraticCostParams (cost_functions/cartpole/cartpole_quadratic_cost.cuh:10-23)
struct CartpoleQuadraticCostParams {
  float cart_position_coeff = 1000;
  float cart_velocity_coeff = 100;
  float pole_angle_coeff = 2000;
  float pole_angular_velocity_coeff = 100;
  float terminal_cost_coeff = 0;
  float desired_terminal_state[4] = {0, 0, (float)M_PI, 0};
  float control_cost_coeff[1] = {10.0f};  // inert on the device path (cost_functions/cost.cuh:205-208)
};

class CartpoleQuadraticCost {
 public:
  using COST_PARAMS_T = CartpoleQuadraticCostParams;
  explicit CartpoleQuadraticCost(void* /*stream*/ = nullptr) {}
  void setParams(const CartpoleQuadraticCostParams& p) { params_ = p; }
  CartpoleQuadraticCostParams getParams() const { return params_; }
  void GPUSetup() {}
  void freeCudaMem() {}
  void bindToStream(void*) {}
  template <class DYN>
  mppi_cost_params desc(const DYN& dyn) const
  {
    const mppi_dynamics_params d = dyn.desc();
    mppi_cost_params c;
    check(mppi_cost_defaults(MPPI_COST_CARTPOLE_QUADRATIC, &d, &c), "mppi_cost_defaults");
    c.coeff[0] = params_.cart_position_coeff;
    c.coeff[1] = params_.cart_velocity_coeff;
    c.coeff[2] = params_.pole_angle_coeff;
    c.coeff[3] = params_.pole_angular_velocity_coeff;
    for (int i = 0; i < 4; ++i) c.goal[i] = params_.desired_terminal_state[i];
    c.terminal_cost_coeff = params_.terminal_cost_coeff;
    return c;
  }
  const float* costMapHost() const { return nullptr; }
  unsigned long long costMapVersion() const { return 0; }
  const float* goalTrajectoryHost() const { return nullptr; }
  int goalTrajectoryLength() const { return 0; }

 private:
  CartpoleQuadraticCostParams params_;
};

// Quadratic goal term (QuadraticCostTrajectoryParams, cost_functions/quadratic_cost/quadratic_cost.cuh:13-63 with a
// single goal) + bilinear cost map + speed term.
template <class DYN_T>
struct QuadraticMapCostParams {
  float s_coeffs[DYN_T::OUTPUT_DIM];
  float s_goal[DYN_T::OUTPUT_DIM];
  float terminal_cost_coeff = 0.0f;
  float map_coeff = 1.0f;
  float map_origin[2] = {0.0f, 0.0f};
  float map_resolution = 1.0f;
  int map_x_index = 0, map_y_index = 1;
  float speed_coeff = 0.0f, speed_target = 0.0f;
  int speed_index = 0;
  bool map_rotated = false;
  float map_rotation[4] = {1.0f, 0.0f, 0.0f, 1.0f};  // row-major 2 x 2 map-frame rotation
  QuadraticMapCostParams()
  {
    for (int i = 0; i < DYN_T::OUTPUT_DIM; ++i) {
      s_coeffs[i] = 1.0f;
      s_goal[i] = 0.0f;
    }
  }
};

template <class DYN_T>
class QuadraticMapCost {
 public:
  using COST_PARAMS_T = QuadraticMapCostParams<DYN_T>;
  explicit QuadraticMapCost(void* /*stream*/ = nullptr) {}
  void setParams(const COST_PARAMS_T& p) { params_ = p; }
  COST_PARAMS_T getParams() const { return params_; }
  // A row-major [height][width] FP32 map (the TwoDTextureHelper::updateTexture role, texture_helper.cu:147-222).
  void setCostMap(const std::vector<float>& map, int width, int height)
  {
    if ((int)map.size() != width * height || width < 2 || height < 2)
      throw std::runtime_error("setCostMap: bad map size");
    map_ = map;
    width_ = width;
    height_ = height;
    ++map_version_;
  }
  // TwoDTextureHelper's update interface (utils/texture_helpers/two_d_texture_helper.cu:21-60, texture_helper.cu:
  // 279-316): new values of the current extent (row-major [height][width] or column-major [width][height]) or with a
  // new extent, the map frame's origin, resolution and (planar, row-major 2 x 2) rotation. They take effect at the
  // controller's next computeControl; values of an unchanged extent are uploaded in place.
  void updateTexture(const std::vector<float>& values, bool column_major = false)
  {
    updateTexture(values, width_, height_, column_major);
  }
  void updateTexture(const std::vector<float>& values, int width, int height, bool column_major = false)
  {
    if (!column_major) {
      setCostMap(values, width, height);
      return;
    }
    if ((int)values.size() != width * height) throw std::runtime_error("updateTexture: bad map size");
    std::vector<float> rm(values.size());
    for (int j = 0; j < width; ++j)
      for (int i = 0; i < height; ++i) rm[(size_t)i * width + j] = values[(size_t)j * height + i];
    setCostMap(rm, width, height);
  }
  void updateOrigin(float x, float y)
  {
    params_.map_origin[0] = x;
    params_.map_origin[1] = y;
  }
  void updateResolution(float resolution) { params_.map_resolution = resolution; }
  void updateRotation(const std::array<float, 4>& rotation)
  {
    params_.map_rotated = true;
    for (int i = 0; i < 4; ++i) params_.map_rotation[i] = rotation[i];
  }
  // ARStandardCost::loadTrackData (cost_functions/autorally/ar_standard_cost.cu:85-142): the map, origin and
  // resolution of an .npz track map (xBounds, yBounds, pixelsPerMeter, channel)
  void loadTrackData(const std::string& map_path, const std::string& channel = "channel0")
  {
    int32_t w = 0, h = 0;
    float ox = 0.0f, oy = 0.0f, res = 1.0f;
    check(mppi_track_map_load_npz(map_path.c_str(), channel.c_str(), nullptr, 0, &w, &h, &ox, &oy, &res),
          "loadTrackData");
    std::vector<float> v((size_t)w * h);
    check(mppi_track_map_load_npz(map_path.c_str(), channel.c_str(), v.data(), (int64_t)v.size(), nullptr, nullptr,
                                  nullptr, nullptr, nullptr),
          "loadTrackData");
    setCostMap(v, w, h);
    updateOrigin(ox, oy);
    updateResolution(res);
  }
  int getWidth() const { return width_; }
  int getHeight() const { return height_; }
  // changes with every setCostMap: the controller uploads the new map on its next computeControl
  unsigned long long costMapVersion() const { return map_version_; }
  void GPUSetup() {}
  void freeCudaMem() {}
  void bindToStream(void*) {}
  template <class DYN>
  mppi_cost_params desc(const DYN& dyn) const
  {
    const mppi_dynamics_params d = dyn.desc();
    mppi_cost_params c;
    check(mppi_cost_defaults(MPPI_COST_QUADRATIC_MAP, &d, &c), "mppi_cost_defaults");
    for (int i = 0; i < DYN_T::OUTPUT_DIM; ++i) {
      c.coeff[i] = params_.s_coeffs[i];
      c.goal[i] = params_.s_goal[i];
    }
    c.terminal_cost_coeff = params_.terminal_cost_coeff;
    c.map_width = width_;
    c.map_height = height_;
    c.map_origin_x = params_.map_origin[0];
    c.map_origin_y = params_.map_origin[1];
    c.map_resolution = params_.map_resolution;
    c.map_coeff = map_.empty() ? 0.0f : params_.map_coeff;
    c.map_x_index = params_.map_x_index;
    c.map_y_index = params_.map_y_index;
    c.speed_coeff = params_.speed_coeff;
    c.speed_target = params_.speed_target;
    c.speed_index = params_.speed_index;
    c.map_rotated = params_.map_rotated ? 1 : 0;
    for (int i = 0; i < 4; ++i) c.map_rotation[i] = params_.map_rotation[i];
    return c;
  }
  const float* costMapHost() const { return map_.empty() ? nullptr : map_.data(); }
  const float* goalTrajectoryHost() const { return nullptr; }
  int goalTrajectoryLength() const { return 0; }

 private:
  COST_PARAMS_T params_;
  std::vector<float> map_;
  int width_ = 0, height_ = 0;
  unsigned long long map_version_ = 0;
};

// QuadraticCostTrajectoryParams (cost_functions/quadratic_cost/quadratic_cost.cuh:12-63): a goal per step for the
// first SIM_TIME_HORIZON steps from current_time on; step t of a rollout is charged against row
// getIndex(t) = min(current_time + t, SIM_TIME_HORIZON - 1).
template <class DYN_T, int SIM_TIME_HORIZON = 1>
struct QuadraticCostTrajectoryParams {
  static const int TIME_HORIZON = SIM_TIME_HORIZON;
  float s_goal[DYN_T::OUTPUT_DIM * SIM_TIME_HORIZON] = {0};
  float s_coeffs[DYN_T::OUTPUT_DIM];
  int current_time = 0;
  QuadraticCostTrajectoryParams()
  {
    for (int i = 0; i < DYN_T::OUTPUT_DIM; ++i) s_coeffs[i] = 1.0f;
  }
  int getIndex(int t) const
  {
    int index = current_time + t;
    if (index >= TIME_HORIZON) index = TIME_HORIZON - 1;
    return index * DYN_T::OUTPUT_DIM;
  }
  Mat<float, DYN_T::OUTPUT_DIM, 1> getDesiredState(int t) const
  {
    Mat<float, DYN_T::OUTPUT_DIM, 1> g;
    for (int i = 0; i < DYN_T::OUTPUT_DIM; ++i) g[i] = s_goal[getIndex(t) + i];
    return g;
  }
  void setCurrentTime(int new_time) { current_time = new_time; }
};

// QuadraticCostTrajectory / QuadraticCost (quadratic_cost.cuh:111-127): the quadratic goal term alone, terminal cost
// 0 (quadratic_cost.cu:28-32). Runs as the built-in quadratic cost with the map and speed terms off and, for a
// horizon above 1, the goal table uploaded with the parameters.
template <class DYN_T, int SIM_TIME_HORIZON>
class QuadraticCostTrajectory {
 public:
  using COST_PARAMS_T = QuadraticCostTrajectoryParams<DYN_T, SIM_TIME_HORIZON>;
  explicit QuadraticCostTrajectory(void* /*stream*/ = nullptr) {}
  void setParams(const COST_PARAMS_T& p) { params_ = p; }
  COST_PARAMS_T getParams() const { return params_; }
  void GPUSetup() {}
  void freeCudaMem() {}
  void bindToStream(void*) {}
  // computeStateCost (quadratic_cost.cu:10-26): sum_i s_coeffs[i] (y_i - goal_i(t))^2
  float computeStateCost(const Mat<float, DYN_T::OUTPUT_DIM, 1>& y, int t = 0) const
  {
    float c = 0.0f;
    for (int i = 0; i < DYN_T::OUTPUT_DIM; ++i) {
      const float e = y[i] - params_.s_goal[params_.getIndex(t) + i];
      c += e * e * params_.s_coeffs[i];
    }
    return c;
  }
  template <class DYN>
  mppi_cost_params desc(const DYN& dyn) const
  {
    const mppi_dynamics_params d = dyn.desc();
    mppi_cost_params c;
    check(mppi_cost_defaults(MPPI_COST_QUADRATIC_MAP, &d, &c), "mppi_cost_defaults");
    for (int i = 0; i < DYN_T::OUTPUT_DIM; ++i) {
      c.coeff[i] = params_.s_coeffs[i];
      c.goal[i] = params_.s_goal[params_.getIndex(0) + i];
    }
    c.terminal_cost_coeff = 0.0f;
    c.map_coeff = 0.0f;
    c.goal_traj_len = goalTrajectoryLength();
    c.goal_time = SIM_TIME_HORIZON > 1 ? params_.current_time : 0;
    return c;
  }
  const float* costMapHost() const { return nullptr; }
  unsigned long long costMapVersion() const { return 0; }
  const float* goalTrajectoryHost() const { return SIM_TIME_HORIZON > 1 ? params_.s_goal : nullptr; }
  int goalTrajectoryLength() const { return SIM_TIME_HORIZON > 1 ? SIM_TIME_HORIZON : 0; }

 private:
  COST_PARAMS_T params_;
};

template <class DYN_T>
using QuadraticCost = QuadraticCostTrajectory<DYN_T, 1>;

// ==================================================================================================== samplers
namespace sampling_distributions {

// GaussianParamsImpl (sampling_distributions/gaussian/gaussian.cuh:21-42)
template <int C_DIM>
struct GaussianParams {
  static const int CONTROL_DIM = C_DIM;
  float std_dev[C_DIM];
  float control_cost_coeff[C_DIM];
  float pure_noise_trajectories_percentage = 0.01f;
  float std_dev_decay = 1.0f;
  dim3 rewrite_controls_block_dim = dim3(32, 16, 1);  // accepted, unused: the engine picks its own launch shapes
  int sum_strides = 32;                               // accepted, unused
  bool time_specific_std_dev = false;                 // set by GaussianTimeVaryingStdDevParams
  GaussianParams()
  {
    for (int i = 0; i < C_DIM; ++i) {
      std_dev[i] = 1.0f;
      control_cost_coeff[i] = 0.0f;
    }
  }
};

// GaussianTimeVaryingStdDevParams (sampling_distributions/gaussian/gaussian.cuh:66-80): std_dev holds a row of
// C_DIM standard deviations per step for the first MAX_TIMESTEPS steps (time_specific_std_dev set); the controller
// uses the first num_timesteps rows, in sampling and in the likelihood-ratio cost.
template <int C_DIM, int MAX_TIMESTEPS = 1>
struct GaussianTimeVaryingStdDevParams : GaussianParams<C_DIM> {
  float std_dev[C_DIM * MAX_TIMESTEPS];
  GaussianTimeVaryingStdDevParams()
  {
    this->time_specific_std_dev = true;
    for (int i = 0; i < C_DIM * MAX_TIMESTEPS; ++i) std_dev[i] = 1.0f;
  }
};

// ColoredNoiseParamsImpl (sampling_distributions/colored_noise/colored_noise.cuh:44-73)
template <int C_DIM>
struct ColoredNoiseParams : GaussianParams<C_DIM> {
  float exponents[C_DIM];
  float offset_decay_rate = 0.97f;
  float fmin = 0.0f;
  int fft_len_mode = MPPI_FFT_LEN_2T;
  ColoredNoiseParams()
  {
    for (int i = 0; i < C_DIM; ++i) exponents[i] = 0.0f;
  }
};

template <class PARAMS_T, int KIND>
class DistributionBase {
 public:
  using SAMPLING_PARAMS_T = PARAMS_T;
  static const int CONTROL_DIM = PARAMS_T::CONTROL_DIM;
  DistributionBase() = default;
  explicit DistributionBase(const PARAMS_T& p, void* /*stream*/ = nullptr) : params_(p) {}
  void setParams(const PARAMS_T& p) { params_ = p; }
  PARAMS_T getParams() const { return params_; }
  void GPUSetup() {}
  void freeCudaMem() {}
  void bindToStream(void*) {}
  virtual mppi_sampler_params desc() const
  {
    mppi_sampler_params s;
    check(mppi_sampler_defaults(KIND, CONTROL_DIM, &s), "mppi_sampler_defaults");
    for (int i = 0; i < CONTROL_DIM; ++i) {
      s.std_dev[i] = params_.std_dev[i];
      s.control_cost_coeff[i] = params_.control_cost_coeff[i];
    }
    s.pure_noise_trajectories_percentage = params_.pure_noise_trajectories_percentage;
    s.std_dev_decay = params_.std_dev_decay;
    return s;
  }
  // the per-step standard deviations [T][C] (row-major, the layout of the reference's time-varying std_dev), or
  // nullptr: params_.std_dev of the final params type, which for the time-varying params is the [MAX_T][C] table
  const float* stdDevTHost() const { return params_.time_specific_std_dev ? params_.std_dev : nullptr; }
  int stdDevTRows() const { return (int)(sizeof(params_.std_dev) / sizeof(float)) / CONTROL_DIM; }
  virtual ~DistributionBase() = default;

 protected:
  PARAMS_T params_;
};

template <class DYN_PARAMS_T, int C_DIM = 0>
class GaussianDistributionT;

// GaussianDistribution<DYN_PARAMS_T> (sampling_distributions/gaussian/gaussian.cuh:120-260); the control dimension
// is taken from the dynamics model the params type belongs to.
template <int C_DIM>
class GaussianDistributionC : public DistributionBase<GaussianParams<C_DIM>, MPPI_SAMPLER_GAUSSIAN> {
 public:
  using DistributionBase<GaussianParams<C_DIM>, MPPI_SAMPLER_GAUSSIAN>::DistributionBase;
};

// GaussianDistribution over GaussianTimeVaryingStdDevParams (gaussian.cuh:120 with that params template).
template <int C_DIM, int MAX_TIMESTEPS>
class GaussianTimeVaryingDistributionC
    : public DistributionBase<GaussianTimeVaryingStdDevParams<C_DIM, MAX_TIMESTEPS>, MPPI_SAMPLER_GAUSSIAN> {
 public:
  static const int MAX_T = MAX_TIMESTEPS;
  using DistributionBase<GaussianTimeVaryingStdDevParams<C_DIM, MAX_TIMESTEPS>,
                         MPPI_SAMPLER_GAUSSIAN>::DistributionBase;
};

template <int C_DIM>
class ColoredNoiseDistributionC : public DistributionBase<ColoredNoiseParams<C_DIM>, MPPI_SAMPLER_COLORED> {
 public:
  using DistributionBase<ColoredNoiseParams<C_DIM>, MPPI_SAMPLER_COLORED>::DistributionBase;
  mppi_sampler_params desc() const override
  {
    mppi_sampler_params s = DistributionBase<ColoredNoiseParams<C_DIM>, MPPI_SAMPLER_COLORED>::desc();
    for (int i = 0; i < C_DIM; ++i) s.exponents[i] = this->params_.exponents[i];
    s.offset_decay_rate = this->params_.offset_decay_rate;
    s.fmin = this->params_.fmin;
    s.fft_len_mode = this->params_.fft_len_mode;
    return s;
  }
};

template <class DYN_PARAMS_T>
struct control_dim_of;
template <>
struct control_dim_of<CartpoleDynamicsParams> {
  static const int value = 1;
};
template <>
struct control_dim_of<DoubleIntegratorParams> {
  static const int value = 2;
};
template <>
struct control_dim_of<RacerDubinsParams> {
  static const int value = 2;
};

template <class DYN_PARAMS_T>
using GaussianDistribution = GaussianDistributionC<control_dim_of<DYN_PARAMS_T>::value>;
template <class DYN_PARAMS_T>
using ColoredNoiseDistribution = ColoredNoiseDistributionC<control_dim_of<DYN_PARAMS_T>::value>;

}  // namespace sampling_distributions

// ==================================================================================================== feedback
// FB_T is a required template argument of the reference controller (controllers/controller.cuh:85-89); feedback is
// disabled by default there (:1020). NoFeedback satisfies the slot and carries no computation.
template <class DYN_T, int FB_TIMESTEPS>
class NoFeedback {
 public:
  static const int TEMPLATED_FB_TIMESTEPS = FB_TIMESTEPS;
  NoFeedback(DYN_T* = nullptr, float = 0.0f) {}
};

// ================================================================================================== controller
struct freeEnergyEstimate {
  float increase = -1, previousBaseline = -1, freeEnergyMean = -1, freeEnergyVariance = -1,
        freeEnergyModifiedVariance = -1, normalizerPercent = -1;
};

struct MPPIFreeEnergyStatistics {
  int nominal_state_used = 0;
  freeEnergyEstimate nominal_sys;
  freeEnergyEstimate real_sys;
};

// ControllerParams (controllers/controller.cuh:46-68)
template <int S_DIM, int C_DIM, int MAX_TIMESTEPS>
struct ControllerParams {
  float dt_ = 0.02f;
  float lambda_ = 1.0f;
  float alpha_ = 0.0f;
  int num_timesteps_ = MAX_TIMESTEPS;
  int num_iters_ = 1;
  unsigned long long seed_ = 0;
  dim3 dynamics_rollout_dim_;  // accepted, unused
  dim3 cost_rollout_dim_;      // accepted, unused
  Mat<float, C_DIM, MAX_TIMESTEPS> init_control_traj_ = Mat<float, C_DIM, MAX_TIMESTEPS>::Zero();
  Mat<float, C_DIM, 1> slide_control_scale_ = Mat<float, C_DIM, 1>::Zero();
  bool use_graph_ = false;
};

// User-written device plugins (include/mppi_amd/device_plugin.hpp) mark themselves with DEVICE_PLUGIN; a controller
// over them is created through PluginControllerFactory, which that header defines (it needs hipcc).
template <class T, class = void>
struct is_device_plugin : std::false_type {};
template <class T>
struct is_device_plugin<T, std::void_t<decltype(T::DEVICE_PLUGIN)>> : std::bool_constant<T::DEVICE_PLUGIN> {};
template <class DYN, class COST>
struct PluginControllerFactory;

// (The default sampler is GaussianDistribution<DYN_T::DYN_PARAMS_T>, i.e. GaussianDistributionC of the model's control
// dimension, which also covers plugin models.)
template <class DYN_T, class COST_T, class FB_T, int MAX_TIMESTEPS, int NUM_ROLLOUTS,
          class SAMPLING_T = sampling_distributions::GaussianDistributionC<DYN_T::CONTROL_DIM>,
          class PARAMS_T = ControllerParams<DYN_T::STATE_DIM, DYN_T::CONTROL_DIM, MAX_TIMESTEPS>>
class VanillaMPPIController {
 public:
  static const int STATE_DIM = DYN_T::STATE_DIM, CONTROL_DIM = DYN_T::CONTROL_DIM, OUTPUT_DIM = DYN_T::OUTPUT_DIM;
  using state_array = typename DYN_T::state_array;
  using control_array = typename DYN_T::control_array;
  using output_array = Mat<float, OUTPUT_DIM, 1>;
  using control_trajectory = Mat<float, CONTROL_DIM, MAX_TIMESTEPS>;
  using state_trajectory = Mat<float, STATE_DIM, MAX_TIMESTEPS>;
  using output_trajectory = Mat<float, OUTPUT_DIM, MAX_TIMESTEPS>;
  // the template arguments, as the reference's plant reads them (controllers/controller.cuh:70-110)
  using TEMPLATED_DYNAMICS = DYN_T;
  using TEMPLATED_COSTS = COST_T;
  using TEMPLATED_FEEDBACK = FB_T;
  using TEMPLATED_SAMPLING = SAMPLING_T;
  using TEMPLATED_SAMPLING_PARAMS = typename SAMPLING_T::SAMPLING_PARAMS_T;
  using TEMPLATED_PARAMS = PARAMS_T;

  // controllers/MPPI/mppi_controller.cuh:36-41
  VanillaMPPIController(DYN_T* model, COST_T* cost, FB_T* fb_controller, SAMPLING_T* sampler, float dt, int max_iter,
                        float lambda, float alpha, int num_timesteps = MAX_TIMESTEPS,
                        const control_trajectory& init_control_traj = control_trajectory::Zero(),
                        void* /*stream*/ = nullptr)
    : model_(model), cost_(cost), fb_(fb_controller), sampler_(sampler)
  {
    params_.dt_ = dt;
    params_.num_iters_ = max_iter;
    params_.lambda_ = lambda;
    params_.alpha_ = alpha;
    params_.num_timesteps_ = num_timesteps;
    params_.init_control_traj_ = init_control_traj;
    control_ = init_control_traj;
  }
  VanillaMPPIController(DYN_T* model, COST_T* cost, FB_T* fb_controller, SAMPLING_T* sampler, const PARAMS_T& params,
                        void* /*stream*/ = nullptr)
    : model_(model), cost_(cost), fb_(fb_controller), sampler_(sampler), params_(params)
  {
    control_ = params.init_control_traj_;
  }
  virtual ~VanillaMPPIController() { release(); }
  VanillaMPPIController(const VanillaMPPIController&) = delete;
  VanillaMPPIController& operator=(const VanillaMPPIController&) = delete;

  virtual std::string getControllerName() const { return "Vanilla MPPI"; }

  // VanillaMPPIController::computeControl (controllers/MPPI/mppi_controller.cu:152-237)
  virtual void computeControl(const state_array& state, int optimization_stride = 1)
  {
    ensure();
    check(mppi_controller_set_control_seq(ctrl_, control_.data()), "updateImportanceSampler");
    check(mppi_controller_compute_control(ctrl_, state.data(), optimization_stride), "computeControl");
    check(mppi_controller_get_control_seq(ctrl_, control_.data()), "getControlSeq");
    state_fresh_ = false;  // the state trajectory is propagated when getTargetStateSeq asks for it
    mppi_weight_stats st;
    check(mppi_controller_get_weight_stats(ctrl_, &st), "weights");
    baseline_ = st.baseline;
    normalizer_ = (float)st.normalizer;
    mppi_free_energy fe;
    check(mppi_controller_get_free_energy(ctrl_, &fe), "free energy");
    fe_.real_sys.freeEnergyMean = fe.free_energy_mean;
    fe_.real_sys.freeEnergyVariance = fe.free_energy_variance;
    fe_.real_sys.freeEnergyModifiedVariance = fe.free_energy_modified_variance;
    fe_.real_sys.normalizerPercent = fe.normalizer_percent;
    fe_.real_sys.increase = fe.increase;
    fe_.real_sys.previousBaseline = fe.previous_baseline;
  }

  control_trajectory getControlSeq() const { return control_; }
  // computeStateTrajectory of the last call's controls (controllers/MPPI/mppi_controller.cu:222), propagated on the
  // host on the first request after each call
  state_trajectory getTargetStateSeq() const
  {
    if (!state_fresh_ && ctrl_) {
      check(mppi_controller_get_state_seq(ctrl_, state_.data()), "getTargetStateSeq");
      state_fresh_ = true;
    }
    return state_;
  }
  float getBaselineCost(int = 0) const { return baseline_; }
  float getNormalizerCost(int = 0) const { return normalizer_; }
  MPPIFreeEnergyStatistics getFreeEnergyStatistics() const { return fe_; }
  void updateImportanceSampler(const control_trajectory& u) { control_ = u; }
  // Controller::slideControlSequence / slideControlSequenceHelper (controllers/controller.cuh:370-376,609-621)
  virtual void slideControlSequence(int steps)
  {
    const int T = params_.num_timesteps_;
    for (int i = 0; i < T; ++i) {
      const int ind = std::min(i + steps, T - 1);
      for (int c = 0; c < CONTROL_DIM; ++c) {
        float v = control_(c, ind);
        if (i + steps > T - 1)
          v = (v - model_->zero_control_[c]) * params_.slide_control_scale_[c] + model_->zero_control_[c];
        control_(c, i) = v;
      }
    }
  }
  PARAMS_T getParams() const { return params_; }
  // New controller parameters; dt, lambda, alpha, the slide scale and the seed are swapped into the running controller
  // on the next computeControl, other changes (horizon, iteration count) create a new one.
  void setParams(const PARAMS_T& p) { params_ = p; }
  void setSamplingParams(const TEMPLATED_SAMPLING_PARAMS& p) { sampler_->setParams(p); }

  // --------------------------------------------------------------------- what the plant uses between solutions
  // interpolateControls (controllers/controller.cuh:383-394): linear between the two neighbouring steps
  control_array interpolateControls(double rel_time, const control_trajectory& c_traj) const
  {
    const double dt = params_.dt_;
    const int last = params_.num_timesteps_ - 1;
    const int lo = std::min(std::max((int)(rel_time / dt), 0), last), hi = std::min(lo + 1, last);
    const double a = (rel_time - lo * dt) / dt;
    control_array u;
    for (int c = 0; c < CONTROL_DIM; ++c) u[c] = (float)((1.0 - a) * c_traj(c, lo) + a * c_traj(c, hi));
    return u;
  }
  // interpolateState (controllers/controller.cuh:396-403): linear, per state
  state_array interpolateState(const state_trajectory& s_traj, double rel_time) const
  {
    const double dt = params_.dt_;
    const int last = params_.num_timesteps_ - 1;
    const int lo = std::min(std::max((int)(rel_time / dt), 0), last), hi = std::min(lo + 1, last);
    const double a = (rel_time - lo * dt) / dt;
    state_array x;
    for (int i = 0; i < STATE_DIM; ++i) x[i] = (float)((1.0 - a) * s_traj(i, lo) + a * s_traj(i, hi));
    return x;
  }
  // getCurrentControl (controllers/controller.cuh:347-363): the interpolated feed-forward control, constrained; the
  // feedback term is zero (NoFeedback)
  template <class FB_STATE>
  control_array getCurrentControl(const state_array& /*state*/, double rel_time,
                                  const state_array& /*target_nominal_state*/, const control_trajectory& c_traj,
                                  FB_STATE& /*fb_state*/) const
  {
    control_array u = interpolateControls(rel_time, c_traj);
    enforce(u);
    return u;
  }
  // only used by RMPPI in the reference (controllers/controller.cuh:327-329)
  void updateImportanceSamplingControl(const state_array&, int) {}
  void resetControls() { control_ = params_.init_control_traj_; }
  bool getFeedbackEnabled() const { return false; }
  void computeFeedback(const state_array&) {}
  void computeFeedbackPropagatedStateSeq() {}
  // outputs of the state trajectory (stateToOutput per step; the built-in models output their state)
  output_trajectory getTargetOutputSeq() const
  {
    const state_trajectory s = getTargetStateSeq();
    output_trajectory o;
    for (int t = 0; t < params_.num_timesteps_; ++t) {
      if constexpr (is_device_plugin<DYN_T>::value) {
        model_->stateToOutput(s.data() + (size_t)t * STATE_DIM, o.data() + (size_t)t * OUTPUT_DIM);
      } else {
        for (int i = 0; i < OUTPUT_DIM; ++i) o(i, t) = s(i, t);
      }
    }
    return o;
  }

  // the plugins, non-owning (controllers/controller.cuh:804-807); parameter changes made through them are picked up
  // by the next computeControl
  DYN_T* model_;
  COST_T* cost_;
  FB_T* fb_;
  SAMPLING_T* sampler_;
  void setSeedCUDARandomNumberGen(unsigned long long seed)
  {
    params_.seed_ = seed;
    if (ctrl_) check(mppi_controller_set_seed(ctrl_, seed), "setSeed");
  }
  int getNumTimesteps() const { return params_.num_timesteps_; }
  float getDt() const { return params_.dt_; }
  float getLambda() const { return params_.lambda_; }
  float getAlpha() const { return params_.alpha_; }
  // rebuild the device side from scratch on the next computeControl
  void resetPlugins() { release(); }

 protected:
  // the controller-kind fields of the descriptor (none for the vanilla controller)
  virtual void extendDesc(mppi_controller_desc&) const {}

 private:
  void enforce(control_array& u) const
  {
    if constexpr (is_device_plugin<DYN_T>::value) {
      model_->enforceConstraints(u.data());
    } else {
      state_array x = state_array::Zero();
      model_->enforceConstraints(x, u);
    }
  }

  // Everything the device side depends on, as bytes: the parameter blocks (for plugin models the objects themselves,
  // which are passed to the kernels by value), the controller parameters and (with_map) the cost map's version.
  std::string snapshot(const mppi_controller_desc& d, bool with_map = true) const
  {
    std::string s;
    auto add = [&s](const void* p, size_t n) { s.append(static_cast<const char*>(p), n); };
    add(&d.dynamics, sizeof(d.dynamics));
    add(&d.cost, sizeof(d.cost));
    add(&d.sampler, sizeof(d.sampler));
    add(&d.num_rollouts, sizeof(int32_t) * 2);
    add(&d.dt, sizeof(float) * 3);
    add(&d.num_iters, sizeof(int32_t) * 3);
    add(d.slide_control_scale, sizeof(d.slide_control_scale));
    add(&d.controller_kind, sizeof(int32_t) * 4 + sizeof(d.state_leash_dist));
    if constexpr (is_device_plugin<DYN_T>::value) {
      add(static_cast<const void*>(model_), sizeof(DYN_T));
      add(static_cast<const void*>(cost_), sizeof(COST_T));
    } else {
      const unsigned long long v = with_map ? cost_->costMapVersion() : 0ull;
      add(&v, sizeof(v));
      if (d.goal_traj_host) add(d.goal_traj_host, sizeof(float) * d.cost.goal_traj_len * DYN_T::OUTPUT_DIM);
    }
    if (d.std_dev_t_host) add(d.std_dev_t_host, sizeof(float) * d.num_timesteps * CONTROL_DIM);
    return s;
  }

  mppi_controller_desc make_desc() const
  {
    mppi_controller_desc d;
    std::memset(&d, 0, sizeof(d));
    d.sampler = sampler_->desc();
    d.sampler.seed = params_.seed_;
    d.num_rollouts = NUM_ROLLOUTS;
    d.num_timesteps = params_.num_timesteps_;
    d.dt = params_.dt_;
    d.lambda = params_.lambda_;
    d.alpha = params_.alpha_;
    d.num_iters = params_.num_iters_;
    d.use_graph = params_.use_graph_ ? 1 : 0;
    for (int c = 0; c < CONTROL_DIM; ++c) d.slide_control_scale[c] = params_.slide_control_scale_[c];
    d.init_control = control_.data();  // [T][C] row-major == the first T columns of the col-major C x MAX_T matrix
    d.std_dev_t_host = sampler_->stdDevTHost();
    if constexpr (is_device_plugin<DYN_T>::value) {
      d.dynamics = model_->desc();
    } else {
      d.dynamics = model_->desc();
      d.cost = cost_->desc(*model_);
      d.goal_traj_host = cost_->goalTrajectoryHost();
    }
    extendDesc(d);
    return d;
  }

  // Creates the device side on first use; afterwards swaps changed parameters into it (mppi_controller_set_params),
  // or creates a new one when the change needs new buffers.
  void ensure()
  {
    const int T = params_.num_timesteps_;
    if (T < 1 || T > MAX_TIMESTEPS) throw std::runtime_error("num_timesteps out of range");
    if (sampler_->stdDevTHost() && sampler_->stdDevTRows() < T)
      throw std::runtime_error("per-step std_dev table holds fewer rows than num_timesteps");
    mppi_controller_desc d = make_desc();
    std::string snap = snapshot(d);
    if (ctrl_) {
      if (snap == snapshot_) return;
      if constexpr (!is_device_plugin<DYN_T>::value) {
        // new values of an unchanged extent (and nothing else changed): staged in place, the graphs stay
        if (cost_->costMapVersion() != map_version_ && cost_->costMapHost() != nullptr &&
            snapshot(d, false) == snapshot_nomap_ && d.cost.map_width == map_w_ && d.cost.map_height == map_h_) {
          check(mppi_controller_update_cost_map(ctrl_, cost_->costMapHost(), map_w_, map_h_, 0), "updateTexture");
          map_version_ = cost_->costMapVersion();
          snapshot_ = std::move(snap);
          return;
        }
        if (cost_->costMapVersion() != map_version_) d.cost_map_host = cost_->costMapHost();
        d.nn_theta_host = model_->hostParameters();
      }
      if (mppi_controller_set_params(ctrl_, &d) == MPPI_OK) {
        snapshot_ = std::move(snap);
        snapshot_nomap_ = snapshot(d, false);
        map_w_ = d.cost.map_width;
        map_h_ = d.cost.map_height;
        if constexpr (!is_device_plugin<DYN_T>::value) map_version_ = cost_->costMapVersion();
        return;
      }
      release();
    }
    if constexpr (is_device_plugin<DYN_T>::value) {
      ctrl_ = PluginControllerFactory<DYN_T, COST_T>::create(d, model_, cost_, plugin_keep_);
    } else {
      d.cost_map_host = cost_->costMapHost();
      d.nn_theta_host = model_->hostParameters();
      check(mppi_controller_create(&d, &ctrl_), "mppi_controller_create");
      map_version_ = cost_->costMapVersion();
    }
    snapshot_ = std::move(snap);
    snapshot_nomap_ = snapshot(d, false);
    map_w_ = d.cost.map_width;
    map_h_ = d.cost.map_height;
  }
  void release()
  {
    if (ctrl_) mppi_controller_destroy(ctrl_);
    ctrl_ = nullptr;
    plugin_keep_.reset();
    state_fresh_ = false;
  }

 protected:
  PARAMS_T params_;

 private:
  mppi_controller* ctrl_ = nullptr;
  std::string snapshot_;              // the parameters the device side was last given
  std::string snapshot_nomap_;        // the same without the cost map's version
  int map_w_ = 0, map_h_ = 0;         // the extent of the map the device side holds
  unsigned long long map_version_ = 0;
  std::shared_ptr<void> plugin_keep_;  // the plugin callback table's state (plugin models only)
  control_trajectory control_;
  mutable state_trajectory state_;
  mutable bool state_fresh_ = false;
  float baseline_ = 0.0f, normalizer_ = 0.0f;
  MPPIFreeEnergyStatistics fe_;
};

// ColoredMPPIParams (controllers/ColoredMPPI/colored_mppi_controller.cuh:16-37)
template <int S_DIM, int C_DIM, int MAX_TIMESTEPS>
struct ColoredMPPIParams : ControllerParams<S_DIM, C_DIM, MAX_TIMESTEPS> {
  float r = 2.0f;
  float gamma = 0.0f;
  Mat<float, S_DIM, 1> state_leash_dist_ = Mat<float, S_DIM, 1>::Zero();
};

// ColoredMPPIController (controllers/ColoredMPPI/colored_mppi_controller.cuh:39-200, .cu:145-268): the vanilla
// iteration with Tsallis weights when gamma and r are both nonzero, the state leash on x0 (applied here on the host
// with the model's enforceLeash against the previous solution's state at the last slide's step count, before the
// state goes to the device), and only control 1 clamped after smoothing.
template <class DYN_T, class COST_T, class FB_T, int MAX_TIMESTEPS, int NUM_ROLLOUTS,
          class SAMPLING_T = sampling_distributions::ColoredNoiseDistributionC<DYN_T::CONTROL_DIM>,
          class PARAMS_T = ColoredMPPIParams<DYN_T::STATE_DIM, DYN_T::CONTROL_DIM, MAX_TIMESTEPS>>
class ColoredMPPIController
    : public VanillaMPPIController<DYN_T, COST_T, FB_T, MAX_TIMESTEPS, NUM_ROLLOUTS, SAMPLING_T, PARAMS_T> {
  using BASE = VanillaMPPIController<DYN_T, COST_T, FB_T, MAX_TIMESTEPS, NUM_ROLLOUTS, SAMPLING_T, PARAMS_T>;

 public:
  using state_array = typename BASE::state_array;
  using BASE::BASE;
  std::string getControllerName() const override { return "Colored MPPI"; }

  void computeControl(const state_array& state, int optimization_stride = 1) override
  {
    state_array local = state;
    if (leash_active_ && solved_) {
      const auto prev = this->getTargetStateSeq();
      const int j = std::min(std::max(leash_jump_, 0), this->getNumTimesteps() - 1);
      this->model_->enforceLeash(state, prev.col(j), this->params_.state_leash_dist_, local);
    }
    BASE::computeControl(local, optimization_stride);
    solved_ = true;
  }
  void slideControlSequence(int steps) override
  {
    leash_jump_ = steps;
    BASE::slideControlSequence(steps);
  }

  void setGamma(float gamma) { this->params_.gamma = gamma; }
  float getGamma() const { return this->params_.gamma; }
  void setRExp(float r) { this->params_.r = r; }
  float getRExp() const { return this->params_.r; }
  void setStateLeashLength(float new_state_leash, int index = 0) { this->params_.state_leash_dist_[index] = new_state_leash; }
  float getStateLeashLength(int index) const { return this->params_.state_leash_dist_[index]; }
  bool getLeashActive() const { return leash_active_; }
  void setLeashActive(bool a) { leash_active_ = a; }
  void setNoiseDecay(float decay)
  {
    auto p = this->sampler_->getParams();
    p.std_dev_decay = decay;
    this->sampler_->setParams(p);
  }
  void setOffsetDecayRate(float rate)
  {
    auto p = this->sampler_->getParams();
    p.offset_decay_rate = rate;
    this->sampler_->setParams(p);
  }
  float getOffsetDecayRate() const { return this->sampler_->getParams().offset_decay_rate; }
  void setColoredNoiseExponents(const std::vector<float>& e)
  {
    auto p = this->sampler_->getParams();
    for (size_t i = 0; i < e.size() && i < (size_t)DYN_T::CONTROL_DIM; ++i) p.exponents[i] = e[i];
    this->sampler_->setParams(p);
  }
  float getColoredNoiseExponent(int index) const { return this->sampler_->getParams().exponents[index]; }
  std::vector<float> getColoredNoiseExponents() const
  {
    std::vector<float> out;
    for (int i = 0; i < DYN_T::CONTROL_DIM; ++i) out.push_back(this->sampler_->getParams().exponents[i]);
    return out;
  }

 protected:
  void extendDesc(mppi_controller_desc& d) const override
  {
    d.controller_kind = MPPI_CONTROLLER_COLORED;
    d.tsallis_gamma = this->params_.gamma;
    d.tsallis_r = this->params_.r;
    d.leash_active = 0;  // leashed here
  }

 private:
  int leash_jump_ = 1;
  bool leash_active_ = false;
  bool solved_ = false;
};

// =================================================================================== two-system controllers
// Feedback gains K_t for RMPPI's real system and getFeedbackControl: gains [T][S][C] (the layout of
// DDPFeedbackState::fb_gain_traj_, feedback_controllers/DDP/ddp.cuh:30-40), set by the caller (the DDP solver that
// computes them in the reference is not part of this library). Fills the FB_T slot of the two-system controllers.
template <class DYN_T, int FB_TIMESTEPS>
class GainFeedback {
 public:
  static const int TEMPLATED_FB_TIMESTEPS = FB_TIMESTEPS;
  GainFeedback(DYN_T* = nullptr, float = 0.0f) {}
  void setGains(const std::vector<float>& g) { gains_ = g; }
  const std::vector<float>& getGains() const { return gains_; }

 private:
  std::vector<float> gains_;
};

template <class T>
struct is_gain_feedback : std::false_type {};
template <class D, int N>
struct is_gain_feedback<GainFeedback<D, N>> : std::true_type {};

// TubeMPPIParams (controllers/Tube-MPPI/tube_mppi_controller.cuh:17-21); value_function_threshold_ is RMPPI's
template <int S_DIM, int C_DIM, int MAX_TIMESTEPS>
struct TubeMPPIParams : ControllerParams<S_DIM, C_DIM, MAX_TIMESTEPS> {
  float nominal_threshold_ = 20.0f;
  float value_function_threshold_ = 1000.0f;
  int num_candidate_nominal_states_ = 9;  // RobustMPPIParams (robust_mppi_controller.cuh:46-53)
  int num_eval_samples_ = 64;             // eval_dyn_kernel_dim_.x
};

// The two-system controllers over mppi_tube_* (both systems' sequences, statistics and state trajectories live in the
// C-ABI object; parameter changes recreate it and restore the sequences and the nominal state).
template <int MODE, class DYN_T, class COST_T, class FB_T, int MAX_TIMESTEPS, int NUM_ROLLOUTS, class SAMPLING_T,
          class PARAMS_T>
class TwoSystemController {
 public:
  static const int STATE_DIM = DYN_T::STATE_DIM, CONTROL_DIM = DYN_T::CONTROL_DIM;
  using state_array = typename DYN_T::state_array;
  using control_array = typename DYN_T::control_array;
  using control_trajectory = Mat<float, CONTROL_DIM, MAX_TIMESTEPS>;
  using state_trajectory = Mat<float, STATE_DIM, MAX_TIMESTEPS>;
  using TEMPLATED_DYNAMICS = DYN_T;
  using TEMPLATED_COSTS = COST_T;
  using TEMPLATED_FEEDBACK = FB_T;
  using TEMPLATED_SAMPLING = SAMPLING_T;
  using TEMPLATED_PARAMS = PARAMS_T;

  TwoSystemController(DYN_T* model, COST_T* cost, FB_T* fb_controller, SAMPLING_T* sampler, float dt, int max_iter,
                      float lambda, float alpha, int num_timesteps = MAX_TIMESTEPS,
                      const control_trajectory& init_control_traj = control_trajectory::Zero(), void* = nullptr)
    : model_(model), cost_(cost), fb_(fb_controller), sampler_(sampler)
  {
    params_.dt_ = dt;
    params_.num_iters_ = max_iter;
    params_.lambda_ = lambda;
    params_.alpha_ = alpha;
    params_.num_timesteps_ = num_timesteps;
    params_.init_control_traj_ = init_control_traj;
    control_[0] = control_[1] = init_control_traj;
  }
  TwoSystemController(DYN_T* model, COST_T* cost, FB_T* fb_controller, SAMPLING_T* sampler, const PARAMS_T& params,
                      void* = nullptr)
    : model_(model), cost_(cost), fb_(fb_controller), sampler_(sampler), params_(params)
  {
    control_[0] = control_[1] = params.init_control_traj_;
  }
  virtual ~TwoSystemController() { release(); }
  TwoSystemController(const TwoSystemController&) = delete;
  TwoSystemController& operator=(const TwoSystemController&) = delete;

  void computeControl(const state_array& state, int optimization_stride = 1)
  {
    ensure();
    push_gains();
    check(mppi_tube_compute_control(ctrl_, state.data(), optimization_stride), "computeControl");
    for (int n = 0; n < 2; ++n) {
      check(mppi_tube_get_control_seq(ctrl_, n, control_[n].data()), "getControlSeq");
      check(mppi_tube_get_state_seq(ctrl_, n, state_[n].data()), "getStateSeq");
      check(mppi_tube_get_weight_stats(ctrl_, n, &stats_[n]), "weights");
      mppi_free_energy fe;
      check(mppi_tube_get_free_energy(ctrl_, n, &fe), "free energy");
      freeEnergyEstimate& e = n ? fe_.nominal_sys : fe_.real_sys;
      e.freeEnergyMean = fe.free_energy_mean;
      e.freeEnergyVariance = fe.free_energy_variance;
      e.freeEnergyModifiedVariance = fe.free_energy_modified_variance;
      e.normalizerPercent = fe.normalizer_percent;
      e.increase = fe.increase;
      e.previousBaseline = fe.previous_baseline;
    }
    int32_t used = 1;
    check(mppi_tube_get_nominal_state_used(ctrl_, &used), "nominal_state_used");
    fe_.nominal_state_used = used;
    check(mppi_tube_get_nominal_state(ctrl_, nominal_state_.data()), "getNominalState");
    have_nominal_ = true;
  }
  void slideControlSequence(int steps)
  {
    if (!ctrl_) return;
    check(mppi_tube_slide_control_seq(ctrl_, steps), "slideControlSequence");
    for (int n = 0; n < 2; ++n) check(mppi_tube_get_control_seq(ctrl_, n, control_[n].data()), "getControlSeq");
    check(mppi_tube_get_nominal_state(ctrl_, nominal_state_.data()), "getNominalState");
  }
  // the nominal state (RMPPI: nominal_state_; Tube: the nominal trajectory's start)
  void setNominalState(const state_array& x)
  {
    nominal_state_ = x;
    have_nominal_ = true;
    if (ctrl_) check(mppi_tube_set_nominal_state(ctrl_, x.data()), "setNominalState");
  }
  state_array getNominalState() const { return nominal_state_; }
  control_trajectory getNominalControlSeq() const { return control_[1]; }
  control_trajectory getActualControlSeq() const { return control_[0]; }
  state_trajectory getNominalStateSeq() const { return state_[1]; }
  state_trajectory getActualStateSeq() const { return state_[0]; }
  // getBaselineCost / getNormalizerCost in the reference's distribution indexing (Tube: 0 actual, 1 nominal;
  // RMPPI: 0 nominal, 1 real)
  float getBaselineCost(int idx = 0) const { return stats_[is_nominal(idx)].baseline; }
  float getNormalizerCost(int idx = 0) const { return (float)stats_[is_nominal(idx)].normalizer; }
  MPPIFreeEnergyStatistics getFreeEnergyStatistics() const { return fe_; }
  // getFeedbackControl (controllers/controller.cuh:567-571): K_t (x - x_goal) with the gains of the FB_T object
  control_array getFeedbackControl(const state_array& x, const state_array& goal, int t) const
  {
    control_array u = control_array::Zero();
    const std::vector<float>& g = gains();
    if (g.empty()) return u;
    for (int j = 0; j < CONTROL_DIM; ++j)
      for (int i = 0; i < STATE_DIM; ++i) u[j] += g[((size_t)t * STATE_DIM + i) * CONTROL_DIM + j] * (x[i] - goal[i]);
    return u;
  }
  PARAMS_T getParams() const { return params_; }
  void setParams(const PARAMS_T& p) { params_ = p; }
  int getNumTimesteps() const { return params_.num_timesteps_; }
  float getDt() const { return params_.dt_; }
  DYN_T* model_;
  COST_T* cost_;
  FB_T* fb_;
  SAMPLING_T* sampler_;

 protected:
  virtual bool is_nominal(int idx) const = 0;
  // the device side, created or refreshed, with the current gains
  mppi_tube_controller* device()
  {
    ensure();
    push_gains();
    return ctrl_;
  }
  // the sequences, state trajectories and nominal state of the device side
  void pull()
  {
    for (int n = 0; n < 2; ++n) {
      check(mppi_tube_get_control_seq(ctrl_, n, control_[n].data()), "getControlSeq");
      check(mppi_tube_get_state_seq(ctrl_, n, state_[n].data()), "getStateSeq");
    }
    check(mppi_tube_get_nominal_state(ctrl_, nominal_state_.data()), "getNominalState");
    have_nominal_ = true;
  }
  PARAMS_T params_;
  control_trajectory control_[2];  // [0] actual / real, [1] nominal
  state_trajectory state_[2];

 private:
  const std::vector<float>& gains() const
  {
    static const std::vector<float> none;
    if constexpr (is_gain_feedback<FB_T>::value) {
      return fb_ ? fb_->getGains() : none;
    } else {
      return none;
    }
  }
  void push_gains()
  {
    const std::vector<float>& g = gains();
    const size_t need = (size_t)params_.num_timesteps_ * STATE_DIM * CONTROL_DIM;
    if (!g.empty() && g.size() < need) throw std::runtime_error("feedback gains hold fewer than T x S x C values");
    check(mppi_tube_set_feedback_gains(ctrl_, g.empty() ? nullptr : g.data()), "setFeedbackGains");
  }
  mppi_tube_desc make_desc() const
  {
    mppi_tube_desc d;
    std::memset(&d, 0, sizeof(d));
    mppi_controller_desc& b = d.base;
    b.sampler = sampler_->desc();
    b.sampler.seed = params_.seed_;
    b.num_rollouts = NUM_ROLLOUTS;
    b.num_timesteps = params_.num_timesteps_;
    b.dt = params_.dt_;
    b.lambda = params_.lambda_;
    b.alpha = params_.alpha_;
    b.num_iters = params_.num_iters_;
    for (int c = 0; c < CONTROL_DIM; ++c) b.slide_control_scale[c] = params_.slide_control_scale_[c];
    b.std_dev_t_host = sampler_->stdDevTHost();
    b.dynamics = model_->desc();
    b.cost = cost_->desc(*model_);
    b.goal_traj_host = cost_->goalTrajectoryHost();
    d.mode = MODE;
    d.nominal_threshold = params_.nominal_threshold_;
    d.value_function_threshold = params_.value_function_threshold_;
    if (MODE == MPPI_TWO_SYSTEM_RMPPI) {
      d.num_candidates = params_.num_candidate_nominal_states_;
      d.num_eval_samples = params_.num_eval_samples_;
    }
    return d;
  }
  std::string snapshot(const mppi_tube_desc& d) const
  {
    std::string s(reinterpret_cast<const char*>(&d), sizeof(d));
    const unsigned long long v = cost_->costMapVersion();
    s.append(reinterpret_cast<const char*>(&v), sizeof(v));
    const mppi_controller_desc& b = d.base;
    if (b.goal_traj_host) s.append((const char*)b.goal_traj_host, sizeof(float) * b.cost.goal_traj_len * DYN_T::OUTPUT_DIM);
    if (b.std_dev_t_host) s.append((const char*)b.std_dev_t_host, sizeof(float) * b.num_timesteps * CONTROL_DIM);
    return s;
  }
  // creates the device side, or recreates it after a parameter change with the sequences and nominal state restored
  void ensure()
  {
    const int T = params_.num_timesteps_;
    if (T < 1 || T > MAX_TIMESTEPS) throw std::runtime_error("num_timesteps out of range");
    mppi_tube_desc d = make_desc();
    d.base.init_control = control_[0].data();
    std::string snap = snapshot(d);
    if (ctrl_ && snap == snapshot_) return;
    release();
    d.base.cost_map_host = cost_->costMapHost();
    d.base.nn_theta_host = model_->hostParameters();
    check(mppi_tube_create(&d, &ctrl_), "mppi_tube_create");
    check(mppi_tube_set_control_seq(ctrl_, 1, control_[1].data()), "setControlSeq");
    if (have_nominal_) check(mppi_tube_set_nominal_state(ctrl_, nominal_state_.data()), "setNominalState");
    snapshot_ = std::move(snap);
  }
  void release()
  {
    if (ctrl_) mppi_tube_destroy(ctrl_);
    ctrl_ = nullptr;
  }
  mppi_tube_controller* ctrl_ = nullptr;
  std::string snapshot_;
  mppi_weight_stats stats_[2]{};
  MPPIFreeEnergyStatistics fe_;
  state_array nominal_state_ = state_array::Zero();
  bool have_nominal_ = false;
};

// TubeMPPIController (controllers/Tube-MPPI/tube_mppi_controller.cuh:23-116): getControlSeq / getTargetStateSeq are
// the nominal system's, getActualControlSeq the actual one's.
template <class DYN_T, class COST_T, class FB_T, int MAX_TIMESTEPS, int NUM_ROLLOUTS,
          class SAMPLING_T = sampling_distributions::GaussianDistributionC<DYN_T::CONTROL_DIM>,
          class PARAMS_T = TubeMPPIParams<DYN_T::STATE_DIM, DYN_T::CONTROL_DIM, MAX_TIMESTEPS>>
class TubeMPPIController : public TwoSystemController<MPPI_TWO_SYSTEM_TUBE, DYN_T, COST_T, FB_T, MAX_TIMESTEPS,
                                                      NUM_ROLLOUTS, SAMPLING_T, PARAMS_T> {
  using BASE =
      TwoSystemController<MPPI_TWO_SYSTEM_TUBE, DYN_T, COST_T, FB_T, MAX_TIMESTEPS, NUM_ROLLOUTS, SAMPLING_T, PARAMS_T>;

 public:
  using BASE::BASE;
  std::string getControllerName() const { return "Tube MPPI"; }
  typename BASE::control_trajectory getControlSeq() const { return this->control_[1]; }
  typename BASE::state_trajectory getTargetStateSeq() const { return this->state_[1]; }
  float getNominalThreshold() const { return this->params_.nominal_threshold_; }
  void setNominalThreshold(float v) { this->params_.nominal_threshold_ = v; }

 protected:
  bool is_nominal(int idx) const override { return idx == 1; }
};

// RobustMPPIController's optimisation (controllers/R-MPPI/robust_mppi_controller.cu:635-756) with the nominal state
// supplied by the caller (setNominalState); getControlSeq is the real system's control.
template <class DYN_T, class COST_T, class FB_T, int MAX_TIMESTEPS, int NUM_ROLLOUTS,
          class SAMPLING_T = sampling_distributions::GaussianDistributionC<DYN_T::CONTROL_DIM>,
          class PARAMS_T = TubeMPPIParams<DYN_T::STATE_DIM, DYN_T::CONTROL_DIM, MAX_TIMESTEPS>>
class RobustMPPIController : public TwoSystemController<MPPI_TWO_SYSTEM_RMPPI, DYN_T, COST_T, FB_T, MAX_TIMESTEPS,
                                                        NUM_ROLLOUTS, SAMPLING_T, PARAMS_T> {
  using BASE = TwoSystemController<MPPI_TWO_SYSTEM_RMPPI, DYN_T, COST_T, FB_T, MAX_TIMESTEPS, NUM_ROLLOUTS,
                                   SAMPLING_T, PARAMS_T>;

 public:
  using BASE::BASE;
  std::string getControllerName() const { return "RMPPI"; }
  typename BASE::control_trajectory getControlSeq() const { return this->control_[0]; }
  typename BASE::state_trajectory getTargetStateSeq() const { return this->state_[1]; }
  float getValueFunctionThreshold() const { return this->params_.value_function_threshold_; }
  void setValueFunctionThreshold(float v) { this->params_.value_function_threshold_ = v; }
  // updateImportanceSamplingControl (robust_mppi_controller.cu:548-628): the candidate evaluation that chooses the
  // nominal state and slides the nominal control, before computeControl
  void updateImportanceSamplingControl(const typename BASE::state_array& state, int stride)
  {
    mppi_tube_controller* c = this->device();
    check(mppi_tube_update_importance_sampling_control(c, state.data(), stride), "updateImportanceSamplingControl");
    this->pull();
    int32_t best = 0;
    check(mppi_tube_get_candidates(c, nullptr, nullptr, nullptr, nullptr, &best), "getCandidates");
    best_index_ = best;
  }
  int getBestIndex() const { return best_index_; }

 protected:
  bool is_nominal(int idx) const override { return idx == 0; }

 private:
  int best_index_ = 0;
};

}  // namespace mppi_amd
