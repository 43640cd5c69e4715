// C-ABI of the engine: parameter defaults and validation, kernel-level entry points and the controller object that
// runs whole MPPI iterations (VanillaMPPIController::computeControl, controllers/MPPI/mppi_controller.cu:152-237).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <utility>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "fused.hpp"
#include "host.hpp"
#include "proposal.hpp"
#include "kernels.hpp"
#include "models.hpp"

namespace mppi_amd {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg)
{
  set_last_error(msg);
  return status;
}

int hip_fail(hipError_t e, const char* what)
{
  set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  return MPPI_E_HIP;
}

int wait_host_signal(const uint32_t* flag, uint32_t want, hipEvent_t ev)
{
  const volatile uint32_t* f = flag;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t n = 1;; ++n) {
    if (*f == want) {
      std::atomic_thread_fence(std::memory_order_acquire);
      return MPPI_OK;
    }
    if ((n & 1023u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(HOST_SPIN_MS)) break;
    __builtin_ia32_pause();
  }
  MPPI_HIP_CHECK(hipEventSynchronize(ev));
  if (*f != want) return fail(MPPI_E_HIP, "control call completed without its completion signal");
  return MPPI_OK;
}

// The attribute belongs to the kernel on the device that is current when it is set, so it is recorded per
// (device, kernel); the lock serialises controllers created from several host threads.
int allow_max_dynamic_lds(const void* fn)
{
  static std::mutex mu;
  static std::set<std::pair<int, const void*>> done;
  int dev = 0;
  MPPI_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (done.count({dev, fn}) != 0) return MPPI_OK;
  MPPI_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_LIMIT));
  done.insert({dev, fn});
  return MPPI_OK;
}

int pure_noise_start(int K, float pct)
{
  const float thr = (1.0f - pct) * (float)K;
  // smallest k with (float)k >= thr, found without a K-long scan
  int k = (int)std::ceil((double)thr);
  if (k < 0) k = 0;
  while (k > 0 && (float)(k - 1) >= thr) --k;
  while (k < K && !((float)k >= thr)) ++k;
  return std::min(k, K);
}

int model_dims(int model, int* S, int* C, int* O)
{
  switch (model) {
    case MPPI_MODEL_CARTPOLE: *S = 4; *C = 1; *O = 4; return MPPI_OK;
    case MPPI_MODEL_DOUBLE_INTEGRATOR: *S = 4; *C = 2; *O = 4; return MPPI_OK;
    case MPPI_MODEL_RACER_DUBINS:
    case MPPI_MODEL_RACER_DUBINS_NN: *S = 7; *C = 2; *O = 7; return MPPI_OK;
    default: return fail(MPPI_E_UNSUPPORTED, "unknown dynamics model id " + std::to_string(model));
  }
}

int validate_dynamics(const mppi_dynamics_params& d)
{
  int S, C, O;
  MPPI_CHECK(model_dims(d.model, &S, &C, &O));
  if (d.state_dim != S || d.control_dim != C || d.output_dim != O)
    return fail(MPPI_E_INVALID, "dynamics params: state/control/output dims do not match the model");
  for (int i = 0; i < C; ++i)
    if (!(d.control_min[i] <= d.control_max[i])) return fail(MPPI_E_INVALID, "dynamics params: control_min > control_max");
  if (d.model == MPPI_MODEL_RACER_DUBINS_NN) {
    const int32_t want[4] = {NN_IN, NN_H1, NN_H2, NN_OUT};
    if (d.nn_num_layers != 4 || std::memcmp(d.nn_layers, want, sizeof(want)) != 0)
      return fail(MPPI_E_UNSUPPORTED, "residual MLP: only the 9-64-64-4 shape is compiled into the kernel library");
    if (d.nn_num_params != NN_PARAMS) return fail(MPPI_E_INVALID, "residual MLP: nn_num_params mismatch");
    for (int i = 0; i < NN_OUT; ++i)
      if (d.nn_output_states[i] < 0 || d.nn_output_states[i] >= S)
        return fail(MPPI_E_INVALID, "residual MLP: nn_output_states out of range");
    if (d.nn_precision != MPPI_NN_FP32 && d.nn_precision != MPPI_NN_BF16)
      return fail(MPPI_E_INVALID, "residual MLP: unknown precision");
    if (d.nn_precision == MPPI_NN_BF16)
      for (int i = 0; i < NN_OUT; ++i)
        for (int j = 0; j < i; ++j)
          if (d.nn_output_states[i] == d.nn_output_states[j])
            return fail(MPPI_E_INVALID, "residual MLP: the BF16 path needs distinct nn_output_states");
  }
  return MPPI_OK;
}

int validate_cost(const mppi_cost_params& c, const mppi_dynamics_params& d, bool need_map_ptr)
{
  if (c.cost != MPPI_COST_CARTPOLE_QUADRATIC && c.cost != MPPI_COST_QUADRATIC_MAP)
    return fail(MPPI_E_UNSUPPORTED, "unknown cost id " + std::to_string(c.cost));
  if (c.output_dim != d.output_dim) return fail(MPPI_E_INVALID, "cost params: output_dim does not match the model");
  if (c.cost == MPPI_COST_QUADRATIC_MAP) {
    if (c.map != nullptr || !need_map_ptr) {
      if (c.map_coeff != 0.0f && (c.map != nullptr || !need_map_ptr)) {
        if (c.map_width < 2 || c.map_height < 2) return fail(MPPI_E_INVALID, "cost map: width/height must be >= 2");
        // the kernels index texels with 32-bit offsets
        if ((int64_t)c.map_width * c.map_height > ((int64_t)1 << 30))
          return fail(MPPI_E_SIZE, "cost map: at most 2^30 texels");
        if (!(c.map_resolution > 0.0f)) return fail(MPPI_E_INVALID, "cost map: resolution must be > 0");
        if (c.map_x_index < 0 || c.map_x_index >= c.output_dim || c.map_y_index < 0 || c.map_y_index >= c.output_dim)
          return fail(MPPI_E_INVALID, "cost map: position indices out of range");
      }
    }
    if (c.speed_coeff != 0.0f && (c.speed_index < 0 || c.speed_index >= c.output_dim))
      return fail(MPPI_E_INVALID, "speed term: speed_index out of range");
  }
  return validate_goal_traj(c);
}

int validate_tsallis(float gamma, float r, bool required)
{
  const bool on = gamma != 0.0f && r != 0.0f;
  if (!on) return required ? fail(MPPI_E_INVALID, "Tsallis weights: gamma and r must be nonzero") : MPPI_OK;
  if (!(gamma > 0.0f)) return fail(MPPI_E_INVALID, "Tsallis weights: gamma must be > 0");
  if (r == 1.0f || !(r == r)) return fail(MPPI_E_INVALID, "Tsallis weights: r must differ from 1");
  return MPPI_OK;
}

int validate_goal_traj(const mppi_cost_params& c)
{
  if (c.goal_traj == nullptr) return MPPI_OK;
  if (c.cost != MPPI_COST_QUADRATIC_MAP)
    return fail(MPPI_E_UNSUPPORTED, "goal trajectory: only the quadratic cost takes one");
  if (c.goal_traj_len < 1) return fail(MPPI_E_INVALID, "goal trajectory: goal_traj_len must be >= 1");
  if (c.goal_time < 0) return fail(MPPI_E_INVALID, "goal trajectory: goal_time must be >= 0");
  return MPPI_OK;
}

int validate_sampler(const mppi_sampler_params& s, const mppi_dynamics_params& d)
{
  if (s.kind != MPPI_SAMPLER_GAUSSIAN && s.kind != MPPI_SAMPLER_COLORED)
    return fail(MPPI_E_UNSUPPORTED, "unknown sampler kind " + std::to_string(s.kind));
  if (s.control_dim != d.control_dim) return fail(MPPI_E_INVALID, "sampler params: control_dim does not match the model");
  for (int i = 0; i < s.control_dim; ++i)
    if (!(s.std_dev[i] > 0.0f)) return fail(MPPI_E_INVALID, "sampler params: std_dev must be > 0");
  if (s.kind == MPPI_SAMPLER_COLORED) {
    if (s.fft_len_mode != MPPI_FFT_LEN_2T && s.fft_len_mode != MPPI_FFT_LEN_T)
      return fail(MPPI_E_INVALID, "colored sampler: unknown fft_len_mode");
    if (!(s.fmin >= 0.0f && s.fmin <= 0.5f)) return fail(MPPI_E_INVALID, "colored sampler: fmin must be in [0, 0.5]");
  }
  return MPPI_OK;
}

int validate_sizes(int K, int T)
{
  if (K < 1) return fail(MPPI_E_INVALID, "num_rollouts must be >= 1");
  if (T < 1) return fail(MPPI_E_INVALID, "num_timesteps must be >= 1");
  if (T > MPPI_MAX_TIMESTEPS) return fail(MPPI_E_SIZE, "num_timesteps exceeds MPPI_MAX_TIMESTEPS");
  if ((int64_t)K * T * MPPI_MAX_CONTROL > ((int64_t)1 << 31)) return fail(MPPI_E_SIZE, "sample buffer too large");
  return MPPI_OK;
}

// ------------------------------------------------------------------------------------ host post-processing
// Savitzky-Golay [-3, 12, 17, 12, -3] / 35 with edge replication, u[0] kept (controllers/controller.cuh:573-607).
void smooth_control(std::vector<float>& u, int T, int C)
{
  const float coef[5] = {-3.0f / 35.0f, 12.0f / 35.0f, 17.0f / 35.0f, 12.0f / 35.0f, -3.0f / 35.0f};
  std::vector<float> buf((size_t)(T + 4) * C);
  for (int c = 0; c < C; ++c) {
    buf[c] = u[c];
    buf[C + c] = u[c];
    for (int t = 0; t < T; ++t) buf[(t + 2) * C + c] = u[t * C + c];
    buf[(T + 2) * C + c] = u[(T - 1) * C + c];
    buf[(T + 3) * C + c] = u[(T - 1) * C + c];
  }
  for (int i = 1; i < T; ++i)
    for (int c = 0; c < C; ++c) {
      float s = 0.0f;
      for (int j = 0; j < 5; ++j) s += coef[j] * buf[(i + j) * C + c];
      u[i * C + c] = s;
    }
}

template <int MODEL>
static void propagate_states(const mppi_dynamics_params& p, const float* x0, const float* u, int T, float dt,
                             float* states)
{
  using M = Model<MODEL>;
  constexpr int S = M::S, C = M::C;
  for (int i = 0; i < S; ++i) states[i] = x0[i];
  for (int t = 0; t + 1 < T; ++t) {
    float uc[C], xd[S];
    for (int c = 0; c < C; ++c) uc[c] = u[t * C + c];
    enforce_constraints<C>(p, uc);
    M::deriv(p, states + t * S, uc, xd);
    M::update(p, states + t * S, states + (t + 1) * S, xd, dt);
  }
}

// computeOutputTrajectoryHelper (controllers/controller.cuh:664-684) on the host.
void propagate(const mppi_dynamics_params& p, const float* x0, const float* u, int T, float dt, float* states)
{
  switch (p.model) {
    case MPPI_MODEL_CARTPOLE: propagate_states<MPPI_MODEL_CARTPOLE>(p, x0, u, T, dt, states); break;
    case MPPI_MODEL_DOUBLE_INTEGRATOR: propagate_states<MPPI_MODEL_DOUBLE_INTEGRATOR>(p, x0, u, T, dt, states); break;
    case MPPI_MODEL_RACER_DUBINS: propagate_states<MPPI_MODEL_RACER_DUBINS>(p, x0, u, T, dt, states); break;
    case MPPI_MODEL_RACER_DUBINS_NN: propagate_states<MPPI_MODEL_RACER_DUBINS_NN>(p, x0, u, T, dt, states); break;
    default: break;
  }
}

template <int C>
static void clamp_sequence(const mppi_dynamics_params& p, float* u, int T)
{
  for (int t = 0; t < T; ++t) enforce_constraints<C>(p, u + t * C);
}

void clamp_controls_seq(const mppi_dynamics_params& p, float* u, int T)
{
  if (p.control_dim == 1) clamp_sequence<1>(p, u, T);
  else clamp_sequence<2>(p, u, T);
}

void free_energy_from_stats(const mppi_weight_stats& st, double lambda, int K, mppi_free_energy* fe)
{
  // core/mppi_common.cu:1101-1117, from the FP64 device sums
  const double norm = st.normalizer / K;
  fe->free_energy_mean = (float)(-lambda * std::log(norm) + st.baseline);
  fe->free_energy_variance = (float)(lambda * (st.sum_w2 / K - norm * norm));
  const double weird = fe->free_energy_variance / (norm * std::sqrt((double)K));
  fe->free_energy_modified_variance = (float)(lambda * (weird + 0.5 * weird * weird));
  fe->normalizer_percent = (float)norm;
  fe->increase = st.baseline - fe->previous_baseline;
}

template <int MODEL>
static void host_step_t(const mppi_dynamics_params& p, const float* x, const float* u_in, float dt, float* xn,
                        float* xd, float* y)
{
  using M = Model<MODEL>;
  float u[M::C], xdl[M::S], xnl[M::S];
  for (int i = 0; i < M::C; ++i) u[i] = u_in[i];
  M::deriv(p, x, u, xdl);
  M::update(p, x, xnl, xdl, dt);
  for (int i = 0; i < M::S; ++i) {
    if (xd) xd[i] = xdl[i];
    if (xn) xn[i] = xnl[i];
  }
  if (y)
    for (int i = 0; i < M::O; ++i) y[i] = xnl[i];
}

}  // namespace mppi_amd

using namespace mppi_amd;

// ======================================================================================================= C-ABI
struct mppi_controller {
  mppi_controller_desc d{};
  bool has_plugin = false;  // rollouts run a user plugin's kernels (mppi_controller_create_plugin)
  mppi_plugin plug{};
  int S = 0, C = 0, O = 0, K = 0, T = 0;
  hipStream_t stream = nullptr;
  float* x0_d = nullptr;
  float* mean_d = nullptr;
  float* samples_d = nullptr;
  float* costs_d = nullptr;
  mppi_weight_stats* stats_d = nullptr;
  void* ws_d = nullptr;
  IterArgs* iter_d = nullptr;
  float* map_d = nullptr;
  float* nn_d = nullptr;
  float* std_t_d = nullptr;  // [T][C] copy of the descriptor's per-step standard deviations
  float* goal_d = nullptr;   // [goal_len][O] copy of the descriptor's goal trajectory
  int goal_len = 0;
  float* w_d = nullptr;      // [K] weights of the Tsallis iteration (the costs stay in costs_d)
  int leash_jump = 1;        // the colored controller's leash index: the last slide's step count
  // staged cost-map updates, applied at the start of the next call (map_mu guards the staging fields)
  std::mutex map_mu;
  std::vector<float> map_stage;  // [stage_h][stage_w]
  int stage_w = 0, stage_h = 0;
  bool stage_values = false, stage_frame = false;
  float stage_origin[2] = {0.0f, 0.0f}, stage_res = 1.0f, stage_rot[4] = {1.0f, 0.0f, 0.0f, 1.0f};
  int stage_rotated = 0;
  float* map_pinned = nullptr;  // pinned source of the in-place upload
  size_t map_pinned_n = 0;
  hipEvent_t map_ev = nullptr;  // the last in-place upload (its source is reused only after it)
  // per-call exchange buffers: mapped fine-grained host memory (the *_hd members are their device-side addresses),
  // moved by copy_words launches on the controller's stream
  IterArgs* iter_h = nullptr;
  float* x0_h = nullptr;
  float* nominal_h = nullptr;
  float* mean_out_h = nullptr;
  mppi_weight_stats* stats_h = nullptr;
  IterArgs* iter_hd = nullptr;
  float* x0_hd = nullptr;
  float* nominal_hd = nullptr;
  float* mean_out_hd = nullptr;
  mppi_weight_stats* stats_hd = nullptr;
  std::vector<float> control, raw_mean, nn_host;
  // getTargetStateSeq: propagated on the first request after a call, from that call's x0 and smoothed (not yet
  // constrained) controls
  mutable std::vector<float> state_seq;
  std::vector<float> state_x0;
  mutable std::vector<float> state_u;
  mutable bool state_stale = false;
  mppi_weight_stats stats{};
  mppi_free_energy fe{};
  uint64_t rng_counter = 0;
  int eager_runs = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool computed = false;            // computeControl has run once (the device loop starts from its x0 and mean)
  bool loop_ready = false;          // loop_iter_d holds the device loop's arguments
  IterArgs* loop_iter_d = nullptr;  // free-running iteration arguments, one per iteration of a call (rng offsets
                                    // advanced on the device)
  hipGraph_t loop_graph = nullptr;
  hipGraphExec_t loop_exec = nullptr;
  hipGraph_t loop_group_graph = nullptr;  // LOOP_GROUP consecutive loop calls in one graph
  hipGraphExec_t loop_group_exec = nullptr;
  hipGraph_t loop_big_graph = nullptr;  // LOOP_BIG consecutive loop calls
  hipGraphExec_t loop_big_exec = nullptr;
  // learned proposal (optional; the network object is owned by the caller)
  ProposalNet* prop = nullptr;
  int prop_K = 0;
  float prop_tolerance = -1.0f;  // < 0: no fallback
  float* prop_dmu_d = nullptr;
  float* prop_sig_d = nullptr;
  float* sigma_base_d = nullptr;
  float* cost0_h = nullptr;
  float* cost0_hd = nullptr;
  int prop_runs = 0;
  hipGraph_t prop_graph = nullptr;
  hipGraphExec_t prop_exec = nullptr;
  int32_t n_prop = 0, n_fallback = 0;
  mppi_weight_stats prop_stats{};
  float prop_cost0 = 0.0f;
  // sampled-trajectory export: the last iteration's constrained samples are kept and its sampling mean saved
  int vis_random = 0, vis_top = 0;
  uint64_t vis_draws = 0;
  float* vis_mean_d = nullptr;
  int* vis_idx_d = nullptr;
  float* vis_buf_d = nullptr;  // controls [n][T][C], outputs [n][T][O], step costs [n][T], terminal [n]
  // stream time of the last computeControl from the first upload to the last download
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  mutable float last_gpu_ms = 0.0f;
  mutable bool timing_pending = false;  // last_gpu_ms is read from ev0 / ev1 on request
  // completion signal of a call: the device workgroup count and epoch (sig_d[0], sig_d[1]) and the mapped host word
  // the call's last kernel publishes the epoch to; `signals` counts the signals enqueued
  uint32_t* sig_d = nullptr;
  uint32_t* done_h = nullptr;
  uint32_t* done_hd = nullptr;
  uint32_t signals = 0;
  // direct calls (no graph, built-in model and cost, normExp weights): the last combine smooths and constrains the new
  // mean into control_out (mapped) and nominal_d, the next call's sampling mean, which the host uploads only when it
  // has changed its sequence since (nominal_dirty)
  float* nominal_d = nullptr;
  float* control_out_h = nullptr;
  float* control_out_hd = nullptr;
  bool nominal_dirty = true;
  bool x0_d_stale = false;  // a direct call read x0 from host memory: x0_d is refreshed before its next use
  // state_u holds the raw mean, smoothed when the trajectory is first requested
  mutable bool state_smooth_pending = false;
};

extern "C" {

const char* mppi_last_error(void) { return g_last_error.c_str(); }

const char* mppi_status_string(int status)
{
  switch (status) {
    case MPPI_OK: return "ok";
    case MPPI_E_INVALID: return "invalid argument";
    case MPPI_E_HIP: return "HIP runtime error";
    case MPPI_E_UNSUPPORTED: return "unsupported combination";
    case MPPI_E_NOMEM: return "out of device memory";
    case MPPI_E_SIZE: return "size exceeds a compiled limit";
    default: return "unknown status";
  }
}

const char* mppi_version(void) { return "0.1.0"; }

int mppi_dynamics_defaults(int32_t model, mppi_dynamics_params* out)
{
  if (!out) return fail(MPPI_E_INVALID, "null output");
  mppi_dynamics_params d;
  std::memset(&d, 0, sizeof(d));
  d.model = model;
  MPPI_CHECK(model_dims(model, &d.state_dim, &d.control_dim, &d.output_dim));
  for (int i = 0; i < MPPI_MAX_CONTROL; ++i) {
    d.control_min[i] = -FLT_MAX;  // Dynamics::setDefaultControlRanges (dynamics/dynamics.cu:20-27)
    d.control_max[i] = FLT_MAX;
  }
  d.cart_mass = 1.0f;
  d.pole_mass = 1.0f;
  d.pole_length = 1.0f;
  d.gravity = 9.81f;
  // RacerDubinsParams defaults (dynamics/racer_dubins/racer_dubins.cuh:81-103)
  d.c_t = 1.3f;
  d.c_b = 2.5f;
  d.c_v = 3.7f;
  d.c_0 = 4.9f;
  d.steering_constant = 0.6f;
  d.steer_command_angle_scale = 5.0f;
  d.steer_angle_scale = -9.1f;
  d.max_steer_angle = 0.5f;
  d.max_steer_rate = 5.0f;
  d.brake_delay_constant = 6.6f;
  d.max_brake_rate_neg = 0.9f;
  d.max_brake_rate_pos = 0.33f;
  d.wheel_base = 0.3f;
  d.gear_sign = 1;
  if (model == MPPI_MODEL_RACER_DUBINS || model == MPPI_MODEL_RACER_DUBINS_NN) {
    // throttle/brake and steering commands in [-1, 1]; the brake state is then clamped to [0, 1] on host and device
    for (int i = 0; i < 2; ++i) {
      d.control_min[i] = -1.0f;
      d.control_max[i] = 1.0f;
    }
  }
  if (model == MPPI_MODEL_RACER_DUBINS_NN) {
    d.nn_num_layers = 4;
    d.nn_layers[0] = NN_IN;
    d.nn_layers[1] = NN_H1;
    d.nn_layers[2] = NN_H2;
    d.nn_layers[3] = NN_OUT;
    // residual on VEL_X, YAW, STEER_ANGLE, BRAKE_STATE
    d.nn_output_states[0] = 0;
    d.nn_output_states[1] = 1;
    d.nn_output_states[2] = 4;
    d.nn_output_states[3] = 5;
    d.nn_precision = MPPI_NN_FP32;
    d.nn_num_params = NN_PARAMS;
  }
  *out = d;
  return MPPI_OK;
}

int mppi_cost_defaults(int32_t cost, const mppi_dynamics_params* dyn, mppi_cost_params* out)
{
  if (!out || !dyn) return fail(MPPI_E_INVALID, "null argument");
  mppi_cost_params c;
  std::memset(&c, 0, sizeof(c));
  c.cost = cost;
  c.output_dim = dyn->output_dim;
  if (cost == MPPI_COST_CARTPOLE_QUADRATIC) {
    // CartpoleQuadraticCostParams (cost_functions/cartpole/cartpole_quadratic_cost.cuh:10-23)
    c.coeff[0] = 1000.0f;
    c.coeff[1] = 100.0f;
    c.coeff[2] = 2000.0f;
    c.coeff[3] = 100.0f;
    c.goal[2] = 3.14159265358979323846f;
  } else if (cost == MPPI_COST_QUADRATIC_MAP) {
    for (int i = 0; i < MPPI_MAX_OUTPUT; ++i) c.coeff[i] = (i < c.output_dim) ? 1.0f : 0.0f;
    c.map_resolution = 1.0f;
    c.map_coeff = 1.0f;
    const bool bicycle = dyn->model == MPPI_MODEL_RACER_DUBINS || dyn->model == MPPI_MODEL_RACER_DUBINS_NN;
    c.map_x_index = bicycle ? 2 : 0;
    c.map_y_index = bicycle ? 3 : 1;
    c.speed_index = 0;
  } else {
    return fail(MPPI_E_UNSUPPORTED, "unknown cost id");
  }
  *out = c;
  return MPPI_OK;
}

int mppi_sampler_defaults(int32_t kind, int32_t control_dim, mppi_sampler_params* out)
{
  if (!out) return fail(MPPI_E_INVALID, "null output");
  if (control_dim < 1 || control_dim > MPPI_MAX_CONTROL) return fail(MPPI_E_INVALID, "control_dim out of range");
  mppi_sampler_params s;
  std::memset(&s, 0, sizeof(s));
  s.kind = kind;
  s.control_dim = control_dim;
  for (int i = 0; i < MPPI_MAX_CONTROL; ++i) s.std_dev[i] = 1.0f;
  s.pure_noise_trajectories_percentage = 0.01f;
  s.std_dev_decay = 1.0f;
  s.offset_decay_rate = 0.97f;
  s.fmin = 0.0f;
  s.fft_len_mode = MPPI_FFT_LEN_2T;
  *out = s;
  return MPPI_OK;
}

int mppi_nn_num_params(int32_t num_layers, const int32_t* layers, int32_t* out_num_params)
{
  if (!layers || !out_num_params || num_layers < 2) return fail(MPPI_E_INVALID, "bad layer description");
  int32_t n = 0;
  for (int i = 1; i < num_layers; ++i) n += (layers[i - 1] + 1) * layers[i];
  *out_num_params = n;
  return MPPI_OK;
}

int mppi_pure_noise_start(int32_t num_rollouts, float pct, int32_t* out_index)
{
  if (!out_index || num_rollouts < 1) return fail(MPPI_E_INVALID, "bad argument");
  *out_index = pure_noise_start(num_rollouts, pct);
  return MPPI_OK;
}

// ----------------------------------------------------------------------------------- host-side plugin methods
int mppi_host_step(const mppi_dynamics_params* p, const float* x, const float* u, float dt, float* x_next,
                   float* x_dot, float* y)
{
  if (!p || !x || !u) return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(validate_dynamics(*p));
  if (p->model == MPPI_MODEL_RACER_DUBINS_NN && !p->nn_theta)
    return fail(MPPI_E_INVALID, "residual MLP: nn_theta (host pointer) is null");
  switch (p->model) {
    case MPPI_MODEL_CARTPOLE: host_step_t<MPPI_MODEL_CARTPOLE>(*p, x, u, dt, x_next, x_dot, y); break;
    case MPPI_MODEL_DOUBLE_INTEGRATOR: host_step_t<MPPI_MODEL_DOUBLE_INTEGRATOR>(*p, x, u, dt, x_next, x_dot, y); break;
    case MPPI_MODEL_RACER_DUBINS: host_step_t<MPPI_MODEL_RACER_DUBINS>(*p, x, u, dt, x_next, x_dot, y); break;
    case MPPI_MODEL_RACER_DUBINS_NN: host_step_t<MPPI_MODEL_RACER_DUBINS_NN>(*p, x, u, dt, x_next, x_dot, y); break;
    default: return fail(MPPI_E_UNSUPPORTED, "unknown model");
  }
  return MPPI_OK;
}

int mppi_host_running_cost_at(const mppi_cost_params* c, const float* y, int32_t t, float* out)
{
  if (!c || !y || !out) return fail(MPPI_E_INVALID, "null argument");
  if (t < 0) return fail(MPPI_E_INVALID, "running cost: step index must be >= 0");
  if (c->cost == MPPI_COST_QUADRATIC_MAP) {
    if (c->map == nullptr && c->map_coeff != 0.0f) return fail(MPPI_E_INVALID, "running cost: map is null");
    MPPI_CHECK(validate_goal_traj(*c));
    switch (c->output_dim) {
      case 4: *out = Cost<MPPI_COST_QUADRATIC_MAP, 4>::running(*c, y, t); return MPPI_OK;
      case 7: *out = Cost<MPPI_COST_QUADRATIC_MAP, 7>::running(*c, y, t); return MPPI_OK;
      default: return fail(MPPI_E_UNSUPPORTED, "running cost: output_dim");
    }
  }
  if (c->cost == MPPI_COST_CARTPOLE_QUADRATIC && c->output_dim == 4) {
    *out = Cost<MPPI_COST_CARTPOLE_QUADRATIC, 4>::running(*c, y, t);
    return MPPI_OK;
  }
  return fail(MPPI_E_UNSUPPORTED, "running cost: unknown cost id");
}

int mppi_host_running_cost(const mppi_cost_params* c, const float* y, float* out)
{
  return mppi_host_running_cost_at(c, y, 0, out);
}

int mppi_host_enforce_constraints(const mppi_dynamics_params* p, float* u)
{
  if (!p || !u) return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(validate_dynamics(*p));
  if (p->control_dim == 1) enforce_constraints<1>(*p, u);
  else if (p->control_dim == 2) enforce_constraints<2>(*p, u);
  else return fail(MPPI_E_UNSUPPORTED, "control_dim");
  return MPPI_OK;
}

// normalizeAngle / shortestAngularDistance, float versions (utils/angle_utils.cuh:23-44)
static float host_normalize_angle(float a)
{
  const float res = fmodf(a + (float)M_PI, (float)(2.0 * M_PI));
  return res <= 0.0f ? res + (float)M_PI : res - (float)M_PI;
}

static void enforce_leash(const mppi_dynamics_params& p, const float* xt, const float* xn, const float* lv, float* out)
{
  const int S = p.state_dim;
  const bool bicycle = p.model == MPPI_MODEL_RACER_DUBINS || p.model == MPPI_MODEL_RACER_DUBINS_NN;
  for (int i = 0; i < S; ++i) out[i] = xt[i];
  constexpr int YAW = 1, PX = 2, PY = 3;
  if (bicycle) {
    // position offset to the nominal in the body frame of the true yaw, clamped there, rotated back
    const float yaw = xt[YAW];
    float dx = xn[PX] - xt[PX], dy = xn[PY] - xt[PY];
    float bx = dx * cosf(yaw) + dy * sinf(yaw);
    float by = -dx * sinf(yaw) + dy * cosf(yaw);
    bx = fminf(fmaxf(bx, -lv[PX]), lv[PX]);
    by = fminf(fmaxf(by, -lv[PY]), lv[PY]);
    dx = bx * cosf(yaw) + -by * sinf(yaw);
    dy = bx * sinf(yaw) + by * cosf(yaw);
    out[PX] += dx;
    out[PY] += dy;
  }
  for (int i = 0; i < S; ++i) {
    if (bicycle && (i == PX || i == PY)) continue;
    const bool yaw = bicycle && i == YAW;
    const float diff = yaw ? host_normalize_angle(xn[i] - xt[i]) : xn[i] - xt[i];
    if (lv[i] < fabsf(diff)) {
      out[i] = xt[i] + fminf(fmaxf(diff, -lv[i]), lv[i]);
      if (yaw) out[i] = host_normalize_angle(out[i]);
    } else {
      out[i] = xn[i];
    }
  }
}

int mppi_host_enforce_leash(const mppi_dynamics_params* p, const float* state_true, const float* state_nominal,
                            const float* leash_values, float* state_out)
{
  if (!p || !state_true || !state_nominal || !leash_values || !state_out) return fail(MPPI_E_INVALID, "null argument");
  if (p->state_dim < 1 || p->state_dim > MPPI_MAX_STATE) return fail(MPPI_E_INVALID, "state_dim out of range");
  enforce_leash(*p, state_true, state_nominal, leash_values, state_out);
  return MPPI_OK;
}

int mppi_generate_normals(uint64_t seed, uint64_t offset, float* out_d, int64_t n, void* stream)
{
  if (n < 0 || (n > 0 && !out_d)) return fail(MPPI_E_INVALID, "bad normals buffer");
  return launch_generate_normals(seed, offset, out_d, n, (hipStream_t)stream);
}

int mppi_sample_gaussian(const mppi_sampler_params* sampler, int32_t K, int32_t T, const float* mean_d,
                         float* samples_d, int32_t stride, int32_t iteration, uint64_t rng_offset,
                         const float* eps_in_d, void* stream)
{
  if (!sampler || !mean_d || !samples_d) return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(validate_sizes(K, T));
  if (sampler->control_dim < 1 || sampler->control_dim > MPPI_MAX_CONTROL)
    return fail(MPPI_E_INVALID, "sampler control_dim out of range");
  if (stride < 0) return fail(MPPI_E_INVALID, "optimization_stride must be >= 0");
  IterRef it{nullptr, {rng_offset, stride, iteration}};
  return launch_sample_gaussian(*sampler, K, T, sampler->control_dim, mean_d, samples_d, it, eps_in_d,
                                (hipStream_t)stream);
}

int mppi_sample_colored(const mppi_sampler_params* sampler, int32_t K, int32_t T, const float* mean_d,
                        float* samples_d, int32_t stride, int32_t iteration, uint64_t rng_offset,
                        const float* spectra_in_d, void* stream)
{
  if (!sampler || !mean_d || !samples_d) return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(validate_sizes(K, T));
  if (sampler->control_dim < 1 || sampler->control_dim > MPPI_MAX_CONTROL)
    return fail(MPPI_E_INVALID, "sampler control_dim out of range");
  if (sampler->fft_len_mode != MPPI_FFT_LEN_2T && sampler->fft_len_mode != MPPI_FFT_LEN_T)
    return fail(MPPI_E_INVALID, "colored sampler: unknown fft_len_mode");
  const int n = (sampler->fft_len_mode == MPPI_FFT_LEN_T) ? T : 2 * T;
  if (n < 2) return fail(MPPI_E_INVALID, "colored sampler: transform length must be >= 2");
  if (stride < 0 || stride >= n) return fail(MPPI_E_INVALID, "colored sampler: optimization_stride out of range");
  IterRef it{nullptr, {rng_offset, stride, iteration}};
  return launch_sample_colored(*sampler, K, T, sampler->control_dim, mean_d, samples_d, it, spectra_in_d,
                               (hipStream_t)stream);
}

int mppi_rollout(const mppi_dynamics_params* dyn, const mppi_cost_params* cost, const mppi_sampler_params* sampler,
                 int32_t K, int32_t T, float dt, float lambda, float alpha, const float* x0_d, const float* mean_d,
                 float* samples_d, float* costs_d, void* stream)
{
  if (!dyn || !cost || !sampler || !x0_d || !mean_d || !samples_d || !costs_d)
    return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(validate_sizes(K, T));
  MPPI_CHECK(validate_dynamics(*dyn));
  MPPI_CHECK(validate_cost(*cost, *dyn, true));
  MPPI_CHECK(validate_sampler(*sampler, *dyn));
  if (dyn->model == MPPI_MODEL_RACER_DUBINS_NN && dyn->nn_theta == nullptr)
    return fail(MPPI_E_INVALID, "residual MLP: nn_theta is null");
  return launch_rollout(*dyn, *cost, *sampler, K, T, dt, lambda, alpha, x0_d, mean_d, samples_d, costs_d,
                        (hipStream_t)stream);
}

int mppi_rollout_two_system(const mppi_dynamics_params* dyn, const mppi_cost_params* cost,
                            const mppi_sampler_params* sampler0, const mppi_sampler_params* sampler1,
                            const mppi_two_system_params* params, int32_t K, int32_t T, float dt, float lambda,
                            float alpha, const float* x0_d, const float* mean_d, float* samples_d, float* costs_d,
                            void* stream)
{
  if (!dyn || !cost || !sampler0 || !params || !x0_d || !mean_d || !samples_d || !costs_d)
    return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(validate_sizes(K, T));
  MPPI_CHECK(validate_dynamics(*dyn));
  MPPI_CHECK(validate_cost(*cost, *dyn, true));
  MPPI_CHECK(validate_sampler(*sampler0, *dyn));
  if (dyn->model == MPPI_MODEL_RACER_DUBINS_NN && dyn->nn_theta == nullptr)
    return fail(MPPI_E_INVALID, "residual MLP: nn_theta is null");
  // distribution 1 contributes only its standard deviations
  mppi_sampler_params sp1 = *sampler0;
  if (sampler1 != nullptr) {
    std::memcpy(sp1.std_dev, sampler1->std_dev, sizeof(sp1.std_dev));
    sp1.std_dev_t = sampler1->std_dev_t;
    MPPI_CHECK(validate_sampler(sp1, *dyn));
  }
  if (params->mode == MPPI_TWO_SYSTEM_RMPPI && !(params->value_function_threshold == params->value_function_threshold))
    return fail(MPPI_E_INVALID, "two-system rollout: value_function_threshold is NaN");
  return launch_two_system(*dyn, *cost, *sampler0, sp1, *params, K, T, dt, lambda, alpha, x0_d, mean_d, samples_d,
                           costs_d, (hipStream_t)stream);
}

int mppi_compute_weights(float* costs_d, int32_t K, float lambda, mppi_weight_stats* stats_d, void* stream)
{
  if (!costs_d || !stats_d) return fail(MPPI_E_INVALID, "null argument");
  if (K < 1) return fail(MPPI_E_INVALID, "num_rollouts must be >= 1");
  if (!(lambda > 0.0f)) return fail(MPPI_E_INVALID, "lambda must be > 0");
  return launch_compute_weights(costs_d, K, lambda, stats_d, (hipStream_t)stream);
}

int mppi_compute_weights_tsallis(float* costs_d, int32_t K, float lambda, float gamma, float r,
                                 mppi_weight_stats* stats_d, void* stream)
{
  if (!costs_d || !stats_d) return fail(MPPI_E_INVALID, "null argument");
  if (K < 1) return fail(MPPI_E_INVALID, "num_rollouts must be >= 1");
  if (!(lambda > 0.0f)) return fail(MPPI_E_INVALID, "lambda must be > 0");
  MPPI_CHECK(validate_tsallis(gamma, r, true));
  return launch_weights(costs_d, costs_d, K, lambda, gamma, r, stats_d, nullptr, nullptr, 0, (hipStream_t)stream);
}

int mppi_weighted_update(const float* weights_d, const mppi_weight_stats* stats_d, const float* samples_d, int32_t K,
                         int32_t T, int32_t C, float* mean_out_d, void* stream)
{
  if (!weights_d || !stats_d || !samples_d || !mean_out_d) return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(validate_sizes(K, T));
  if (C < 1 || C > MPPI_MAX_CONTROL) return fail(MPPI_E_INVALID, "control_dim out of range");
  // internal workspace of the partial rows, kept per device and grown on demand (a growth allocates, so the first
  // call at a new maximum size must not be inside a stream capture)
  static std::mutex mu;
  static std::vector<std::pair<void*, size_t>> per_dev(64, {nullptr, 0});
  int dev = 0;
  MPPI_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= (int)per_dev.size()) return fail(MPPI_E_INVALID, "device index out of range");
  const size_t need = weighted_update_workspace_bytes(K, T, C);
  void* ws = nullptr;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto& slot = per_dev[dev];
    if (slot.second < need) {
      if (slot.first) MPPI_HIP_CHECK(hipFree(slot.first));
      slot = {nullptr, 0};
      MPPI_CHECK(dev_alloc(&slot.first, need));
      slot.second = need;
    }
    ws = slot.first;
  }
  return launch_weighted_update(weights_d, stats_d, samples_d, K, T, C, mean_out_d, ws, (hipStream_t)stream);
}

// --------------------------------------------------------------------------------------------------- controller
// The sample buffer receives the constrained samples of every iteration with keep_samples or the trajectory export.
// The colored controller's Tsallis weights (ColoredMPPIParams gamma and r both nonzero).
static bool uses_tsallis(const mppi_controller_desc& d)
{
  return d.controller_kind == MPPI_CONTROLLER_COLORED && d.tsallis_gamma != 0.0f && d.tsallis_r != 0.0f;
}

// The colored controller's own parameters: kind, Tsallis gamma / r, leash distances.
static int validate_controller_kind(const mppi_controller_desc& d, bool plugin)
{
  if (d.controller_kind != MPPI_CONTROLLER_VANILLA && d.controller_kind != MPPI_CONTROLLER_COLORED)
    return fail(MPPI_E_INVALID, "unknown controller_kind");
  if (d.controller_kind != MPPI_CONTROLLER_COLORED) return MPPI_OK;
  MPPI_CHECK(validate_tsallis(d.tsallis_gamma, d.tsallis_r, false));
  if (plugin && uses_tsallis(d))
    return fail(MPPI_E_UNSUPPORTED, "Tsallis weights: not available with a dynamics plugin");
  for (int i = 0; i < MPPI_MAX_STATE; ++i)
    if (!(d.state_leash_dist[i] >= 0.0f)) return fail(MPPI_E_INVALID, "state_leash_dist must be >= 0");
  return MPPI_OK;
}

static bool keeps_samples(const mppi_controller* c) { return c->d.keep_samples != 0 || c->vis_random + c->vis_top > 0; }

// Host post-processing of a plugin controller: the plugin's own constraint and step (Dynamics::enforceConstraints and
// Dynamics::step on the host) in place of the built-in models'.
static void plugin_clamp_sequence(const mppi_controller* c, float* u)
{
  for (int t = 0; t < c->T; ++t) (void)c->plug.host_enforce_constraints(c->plug.user, u + (size_t)t * c->C);
}

static void plugin_propagate(const mppi_controller* c, const float* x0, const float* u, float* states)
{
  const int S = c->S, C = c->C;
  for (int i = 0; i < S; ++i) states[i] = x0[i];
  std::vector<float> uc((size_t)C);
  for (int t = 0; t + 1 < c->T; ++t) {
    for (int j = 0; j < C; ++j) uc[j] = u[(size_t)t * C + j];
    (void)c->plug.host_enforce_constraints(c->plug.user, uc.data());
    (void)c->plug.host_step(c->plug.user, states + (size_t)t * S, uc.data(), c->d.dt, states + (size_t)(t + 1) * S);
  }
}

// Destroys every captured graph of the controller (the next use captures again); the caller has synchronised the
// stream, so no replay of them is still queued.
static void drop_graphs(mppi_controller* c)
{
  hipGraphExec_t* execs[] = {&c->graph_exec, &c->loop_exec, &c->loop_group_exec, &c->loop_big_exec, &c->prop_exec};
  hipGraph_t* graphs[] = {&c->graph, &c->loop_graph, &c->loop_group_graph, &c->loop_big_graph, &c->prop_graph};
  for (hipGraphExec_t* e : execs) {
    if (*e) (void)hipGraphExecDestroy(*e);
    *e = nullptr;
  }
  for (hipGraph_t* g : graphs) {
    if (*g) (void)hipGraphDestroy(*g);
    *g = nullptr;
  }
}

static int controller_free(mppi_controller* c)
{
  if (!c) return MPPI_OK;
  drop_graphs(c);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  void* dev[] = {c->x0_d, c->mean_d, c->samples_d, c->costs_d, c->stats_d, c->ws_d, c->iter_d,
                 c->map_d, c->nn_d, c->loop_iter_d, c->prop_dmu_d, c->prop_sig_d, c->sigma_base_d,
                 c->vis_mean_d, c->vis_idx_d, c->vis_buf_d, c->std_t_d, c->goal_d, c->w_d, c->sig_d,
                 c->nominal_d};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (c->map_ev) (void)hipEventDestroy(c->map_ev);
  void* host[] = {c->iter_h,  c->x0_h,       c->nominal_h, c->mean_out_h,   c->stats_h,
                  c->cost0_h, c->map_pinned, c->done_h,    c->control_out_h};
  for (void* p : host)
    if (p) (void)hipHostFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MPPI_OK;
}

// Controller-owned device copies of the descriptor's host tables -- per-step standard deviations [T][C] and the goal
// trajectory [goal_traj_len][O] -- pointed to from the parameter blocks of `n`. Without a host table the block's own
// pointer stays (a caller's device table, or NULL for none).
static int upload_tables(mppi_controller* c, const mppi_controller_desc* desc, mppi_controller_desc& n)
{
  if (desc->std_dev_t_host != nullptr) {
    const size_t cnt = (size_t)c->T * c->C;
    for (size_t i = 0; i < cnt; ++i)
      if (!(desc->std_dev_t_host[i] > 0.0f)) return fail(MPPI_E_INVALID, "per-step std_dev must be > 0");
    if (c->std_t_d == nullptr) MPPI_CHECK(dev_alloc(&c->std_t_d, sizeof(float) * cnt));
    MPPI_HIP_CHECK(hipMemcpy(c->std_t_d, desc->std_dev_t_host, sizeof(float) * cnt, hipMemcpyHostToDevice));
    n.sampler.std_dev_t = c->std_t_d;
  }
  if (desc->goal_traj_host != nullptr && !c->has_plugin) {
    const int H = desc->cost.goal_traj_len;
    if (H < 1) return fail(MPPI_E_INVALID, "goal trajectory: goal_traj_len must be >= 1");
    const size_t cnt = (size_t)H * c->O;
    if (c->goal_d == nullptr || H != c->goal_len) {
      if (c->goal_d) MPPI_HIP_CHECK(hipFree(c->goal_d));
      if (c->d.cost.goal_traj == c->goal_d) c->d.cost.goal_traj = nullptr;
      c->goal_d = nullptr;
      MPPI_CHECK(dev_alloc(&c->goal_d, sizeof(float) * cnt));
      c->goal_len = H;
    }
    MPPI_HIP_CHECK(hipMemcpy(c->goal_d, desc->goal_traj_host, sizeof(float) * cnt, hipMemcpyHostToDevice));
    n.cost.goal_traj = c->goal_d;
  }
  n.std_dev_t_host = nullptr;
  n.goal_traj_host = nullptr;
  return MPPI_OK;
}

static int controller_setup(mppi_controller* c, const mppi_controller_desc* desc)
{
  c->d = *desc;
  c->K = desc->num_rollouts;
  c->T = desc->num_timesteps;
  MPPI_CHECK(validate_sizes(c->K, c->T));
  if (c->has_plugin) {
    const mppi_plugin& p = c->plug;
    if (p.state_dim < 1 || p.state_dim > MPPI_MAX_STATE || p.control_dim < 1 || p.control_dim > MPPI_MAX_CONTROL ||
        p.output_dim < 1)
      return fail(MPPI_E_INVALID, "plugin: dimensions out of range");
    if (!p.enqueue_iteration || !p.host_step || !p.host_enforce_constraints)
      return fail(MPPI_E_INVALID, "plugin: missing callback");
    c->d.dynamics.model = 0;
    c->d.dynamics.state_dim = p.state_dim;
    c->d.dynamics.control_dim = p.control_dim;
    c->d.dynamics.output_dim = p.output_dim;
    c->d.dynamics.nn_theta = nullptr;
    c->d.cost.map = nullptr;
  } else {
    MPPI_CHECK(validate_dynamics(desc->dynamics));
  }
  MPPI_CHECK(validate_sampler(desc->sampler, c->d.dynamics));
  MPPI_CHECK(validate_controller_kind(*desc, c->has_plugin));
  if (!(desc->lambda > 0.0f)) return fail(MPPI_E_INVALID, "lambda must be > 0");
  if (desc->num_iters < 1) return fail(MPPI_E_INVALID, "num_iters must be >= 1");
  c->S = c->d.dynamics.state_dim;
  c->C = c->d.dynamics.control_dim;
  c->O = c->d.dynamics.output_dim;
  const int K = c->K, T = c->T, C = c->C, S = c->S;
  MPPI_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  MPPI_CHECK(dev_alloc(&c->x0_d, sizeof(float) * MPPI_MAX_STATE));
  MPPI_CHECK(dev_alloc(&c->mean_d, sizeof(float) * T * C));
  MPPI_CHECK(dev_alloc(&c->samples_d, sizeof(float) * (size_t)K * T * C));
  MPPI_CHECK(dev_alloc(&c->costs_d, sizeof(float) * K));
  MPPI_CHECK(dev_alloc(&c->stats_d, sizeof(mppi_weight_stats)));
  // the workspace serves the fused iteration's partials and, for the colored controller, the weighted update's
  const size_t ws = std::max(fused_iteration_workspace_bytes(K, T, C),
                             desc->controller_kind == MPPI_CONTROLLER_COLORED ? weighted_update_workspace_bytes(K, T, C)
                                                                              : (size_t)0);
  MPPI_CHECK(dev_alloc(&c->ws_d, ws));
  if (desc->controller_kind == MPPI_CONTROLLER_COLORED) MPPI_CHECK(dev_alloc(&c->w_d, sizeof(float) * K));
  MPPI_CHECK(dev_alloc(&c->iter_d, sizeof(IterArgs) * desc->num_iters));
  MPPI_CHECK(dev_alloc(&c->loop_iter_d, sizeof(IterArgs) * desc->num_iters));
  MPPI_CHECK(host_alloc_mapped(&c->iter_h, &c->iter_hd, sizeof(IterArgs) * desc->num_iters));
  MPPI_CHECK(host_alloc_mapped(&c->x0_h, &c->x0_hd, sizeof(float) * MPPI_MAX_STATE));
  MPPI_CHECK(host_alloc_mapped(&c->nominal_h, &c->nominal_hd, sizeof(float) * T * C));
  MPPI_CHECK(host_alloc_mapped(&c->mean_out_h, &c->mean_out_hd, sizeof(float) * T * C));
  MPPI_CHECK(host_alloc_mapped(&c->stats_h, &c->stats_hd, sizeof(mppi_weight_stats)));
  MPPI_CHECK(host_alloc_mapped(&c->done_h, &c->done_hd, sizeof(uint32_t)));
  MPPI_CHECK(host_alloc_mapped(&c->control_out_h, &c->control_out_hd, sizeof(float) * T * C));
  MPPI_CHECK(dev_alloc(&c->nominal_d, sizeof(float) * T * C));
  MPPI_CHECK(dev_alloc(&c->sig_d, 2 * sizeof(uint32_t)));
  MPPI_HIP_CHECK(hipMemsetAsync(c->sig_d, 0, 2 * sizeof(uint32_t), c->stream));
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  // plugin-owned device copies of the cost map and the MLP parameters
  if (!c->has_plugin && desc->cost_map_host != nullptr) {
    const size_t n = (size_t)desc->cost.map_width * desc->cost.map_height;
    if (n < 4) return fail(MPPI_E_INVALID, "cost map: width/height must be >= 2");
    MPPI_CHECK(dev_alloc(&c->map_d, sizeof(float) * n));
    MPPI_HIP_CHECK(hipMemcpy(c->map_d, desc->cost_map_host, sizeof(float) * n, hipMemcpyHostToDevice));
    c->d.cost.map = c->map_d;
  }
  MPPI_CHECK(upload_tables(c, desc, c->d));
  if (!c->has_plugin) MPPI_CHECK(validate_cost(c->d.cost, c->d.dynamics, true));
  if (!c->has_plugin && desc->dynamics.model == MPPI_MODEL_RACER_DUBINS_NN) {
    if (desc->nn_theta_host == nullptr) return fail(MPPI_E_INVALID, "residual MLP: nn_theta_host is required");
    c->nn_host.assign(desc->nn_theta_host, desc->nn_theta_host + desc->dynamics.nn_num_params);
    MPPI_CHECK(dev_alloc(&c->nn_d, sizeof(float) * c->nn_host.size()));
    MPPI_HIP_CHECK(hipMemcpy(c->nn_d, c->nn_host.data(), sizeof(float) * c->nn_host.size(), hipMemcpyHostToDevice));
    c->d.dynamics.nn_theta = c->nn_d;
  }
  c->d.init_control = nullptr;
  c->d.cost_map_host = nullptr;
  c->d.nn_theta_host = nullptr;
  c->control.assign((size_t)T * C, 0.0f);
  if (desc->init_control) std::copy(desc->init_control, desc->init_control + (size_t)T * C, c->control.begin());
  c->raw_mean = c->control;
  c->state_seq.assign((size_t)T * S, 0.0f);
  c->stats_h->baseline = 0.0f;
  c->rng_counter = 0;
  return MPPI_OK;
}

static int controller_create(const mppi_controller_desc* desc, const mppi_plugin* plugin, mppi_controller** out)
{
  if (!desc || !out) return fail(MPPI_E_INVALID, "null argument");
  *out = nullptr;
  mppi_controller* c = new (std::nothrow) mppi_controller();
  if (!c) return fail(MPPI_E_NOMEM, "host allocation failed");
  if (plugin != nullptr) {
    c->has_plugin = true;
    c->plug = *plugin;
  }
  const int st = controller_setup(c, desc);
  if (st != MPPI_OK) {
    const std::string msg = g_last_error;
    controller_free(c);
    set_last_error(msg);
    return st;
  }
  *out = c;
  return MPPI_OK;
}

int mppi_controller_create(const mppi_controller_desc* desc, mppi_controller** out)
{
  return controller_create(desc, nullptr, out);
}

int mppi_controller_create_plugin(const mppi_controller_desc* desc, const mppi_plugin* plugin, mppi_controller** out)
{
  if (!plugin) return fail(MPPI_E_INVALID, "null plugin");
  return controller_create(desc, plugin, out);
}

int mppi_controller_destroy(mppi_controller* ctrl)
{
  if (ctrl && ctrl->stream) (void)hipStreamSynchronize(ctrl->stream);
  return controller_free(ctrl);
}

// One iteration of a plugin controller: colored samples (if any) drawn by the library's sampler into the sample
// buffer, the plugin's fused rollout kernel writing per-workgroup partials, then the library's combine. The partial
// rows live after room for ceil(K / 16) partials in the workspace (the plugin picks its own workgroup size).
static int enqueue_plugin_iteration(mppi_controller* c, IterRef ref, IterArgs* advance, const ResultCopies* copies,
                                    int advance_step, hipEvent_t between, int* partials_out)
{
  const mppi_controller_desc& d = c->d;
  const bool colored = d.sampler.kind == MPPI_SAMPLER_COLORED;
  if (ref.ptr == nullptr) return fail(MPPI_E_INVALID, "plugin iteration: device iteration arguments required");
  if (colored) MPPI_CHECK(launch_sample_colored(d.sampler, c->K, c->T, c->C, c->mean_d, c->samples_d, ref, nullptr,
                                                c->stream));
  const int gmax = (c->K + 15) / 16;
  mppi_plugin_iteration it{};
  it.x0_d = c->x0_d;
  it.mean_d = c->mean_d;
  it.samples_d = c->samples_d;
  it.costs_d = c->costs_d;
  it.iter_d = ref.ptr;
  it.partials_d = c->ws_d;
  it.sums_d = reinterpret_cast<float*>(static_cast<char*>(c->ws_d) + (size_t)gmax * sizeof(Partial));
  it.max_partials = gmax;
  it.num_rollouts = c->K;
  it.num_timesteps = c->T;
  it.generate = colored ? 0 : 1;
  it.write_samples = keeps_samples(c) ? 1 : 0;
  it.dt = d.dt;
  it.lambda = d.lambda;
  it.alpha = d.alpha;
  it.sampler = d.sampler;
  it.stream = (void*)c->stream;
  int32_t G = 0;
  const int st = c->plug.enqueue_iteration(c->plug.user, &it, &G);
  if (st != MPPI_OK) return fail(st, "plugin enqueue_iteration failed with status " + std::to_string(st));
  if (G < 1 || G > gmax) return fail(MPPI_E_INVALID, "plugin enqueue_iteration: partial count out of range");
  if (partials_out != nullptr) *partials_out = G;
  if (between != nullptr) MPPI_HIP_CHECK(hipEventRecord(between, c->stream));
  return launch_combine(static_cast<const Partial*>(c->ws_d), it.sums_d, G, c->K, c->T * c->C, d.lambda, c->stats_d,
                        c->mean_d, advance, advance_step, copies, c->stream);
}

// One iteration of the colored controller with Tsallis weights, which need the global baseline before any weight
// (so the fused kernel's per-workgroup log-sum-exp partials do not apply): samples into the sample buffer, the rollout
// (costs, constrained samples written back), the weight transform against the global baseline, the weighted update
// into the mean, then the mean to the host copy on the last iteration.
static int enqueue_tsallis_iteration(mppi_controller* c, IterRef ref, IterArgs* advance, const ResultCopies* copies,
                                     int advance_step, hipEvent_t between, int* partials_out)
{
  const mppi_controller_desc& d = c->d;
  const int K = c->K, T = c->T, C = c->C;
  if (d.sampler.kind == MPPI_SAMPLER_COLORED)
    MPPI_CHECK(launch_sample_colored(d.sampler, K, T, C, c->mean_d, c->samples_d, ref, nullptr, c->stream));
  else
    MPPI_CHECK(launch_sample_gaussian(d.sampler, K, T, C, c->mean_d, c->samples_d, ref, nullptr, c->stream));
  MPPI_CHECK(launch_rollout(d.dynamics, d.cost, d.sampler, K, T, d.dt, d.lambda, d.alpha, c->x0_d, c->mean_d,
                            c->samples_d, c->costs_d, c->stream));
  if (between != nullptr) MPPI_HIP_CHECK(hipEventRecord(between, c->stream));
  if (partials_out != nullptr) *partials_out = 0;
  MPPI_CHECK(launch_weights(c->costs_d, c->w_d, K, d.lambda, d.tsallis_gamma, d.tsallis_r, c->stats_d,
                            copies != nullptr ? copies->stats : nullptr, advance, advance_step, c->stream));
  MPPI_CHECK(launch_weighted_update(c->w_d, c->stats_d, c->samples_d, K, T, C, c->mean_d, c->ws_d, c->stream));
  if (copies != nullptr && copies->mean != nullptr) {
    const WordSpan m[] = {{c->mean_d, copies->mean, T * C}};
    MPPI_CHECK(launch_copy_words(m, 1, c->stream, &copies->signal));
  }
  return MPPI_OK;
}

// One MPPI iteration on the controller's stream: the fused sampling + rollout + weights + update (Gaussian samples
// drawn inside the rollout kernel; colored samples too where the kernel's workgroup shape allows, else by the colored
// sampler into the sample buffer first).
static int enqueue_iteration(mppi_controller* c, IterRef ref, IterArgs* advance,
                             const ResultCopies* copies = nullptr, int advance_step = 1,
                             hipEvent_t between = nullptr, int* partials_out = nullptr)
{
  const mppi_controller_desc& d = c->d;
  const bool colored = d.sampler.kind == MPPI_SAMPLER_COLORED;
  if (c->has_plugin) return enqueue_plugin_iteration(c, ref, advance, copies, advance_step, between, partials_out);
  if (uses_tsallis(d)) return enqueue_tsallis_iteration(c, ref, advance, copies, advance_step, between, partials_out);
  ColoredPlan plan;
  if (colored) MPPI_CHECK(colored_plan(d.sampler, c->K, c->T, c->C, c->stream, &plan));
  // the combine kernel writes the new mean over the sampling mean, as updateDistributionParamsFromDevice writes
  // control_means_d_ (sampling_distributions/gaussian/gaussian.cu:434-467); the next iteration samples around it
  return launch_fused_iteration(d.dynamics, d.cost, d.sampler, c->K, c->T, d.dt, d.lambda, d.alpha, c->x0_d,
                                c->mean_d, c->samples_d, !colored, keeps_samples(c), ref, c->costs_d, c->ws_d,
                                c->stats_d, c->mean_d, advance, c->stream, nullptr, nullptr, copies,
                                colored ? &plan : nullptr, advance_step, between, partials_out);
}

// The completion signal of a control call (given by its last kernel).
static HostSignal call_signal(const mppi_controller* c)
{
  HostSignal h;
  h.count = c->sig_d;
  h.epoch = c->sig_d + 1;
  h.flag = c->done_hd;
  return h;
}

// The work of one computeControl call, enqueued on the controller's stream (captured once when use_graph is set).
static int controller_enqueue(mppi_controller* c)
{
  const int K = c->K, T = c->T, C = c->C, S = c->S;
  const hipStream_t s = c->stream;
  const mppi_controller_desc& d = c->d;
  const bool vis = c->vis_random + c->vis_top > 0;
  const WordSpan in[] = {{c->iter_hd, c->iter_d, words_of<IterArgs>(d.num_iters)},
                         {c->x0_hd, c->x0_d, S},
                         {c->nominal_hd, c->mean_d, T * C},
                         {c->nominal_hd, c->vis_mean_d, T * C}};
  // with the trajectory export and one iteration, the nominal is also the last iteration's sampling mean
  MPPI_CHECK(launch_copy_words(in, vis && d.num_iters == 1 ? 4 : 3, s));
  // the last combine also writes the new mean and the statistics to the host buffers
  ResultCopies out;
  out.mean = c->mean_out_hd;
  out.stats = c->stats_hd;
  out.signal = call_signal(c);
  for (int it = 0; it < d.num_iters; ++it) {
    if (vis && d.num_iters > 1 && it + 1 == d.num_iters) {
      const WordSpan m[] = {{c->mean_d, c->vis_mean_d, T * C}};
      MPPI_CHECK(launch_copy_words(m, 1, s));
    }
    const IterRef ref{c->iter_d + it, {0, 0, 0}};
    MPPI_CHECK(enqueue_iteration(c, ref, nullptr, it + 1 == d.num_iters ? &out : nullptr));
  }
  return MPPI_OK;
}

// x0_d from the last call's x0 (host staging) where a direct call left it behind, on the controller's stream.
static int refresh_x0_d(mppi_controller* c)
{
  if (!c->x0_d_stale) return MPPI_OK;
  const WordSpan w[] = {{c->x0_hd, c->x0_d, c->S}};
  MPPI_CHECK(launch_copy_words(w, 1, c->stream));
  c->x0_d_stale = false;
  return MPPI_OK;
}

// Whether computeControl takes the direct path: launches without a graph, the built-in models and costs and the
// fused normExp iteration (the device post-processing mirrors the vanilla controller's).
static bool direct_path(const mppi_controller* c)
{
  const mppi_controller_desc& d = c->d;
  return !d.use_graph && !c->has_plugin && d.controller_kind == MPPI_CONTROLLER_VANILLA && !uses_tsallis(d);
}

// The direct path's call: the iteration arguments by value and x0 read from mapped host memory by the rollout, the
// nominal uploaded only when the host changed it (else the one the last call left in nominal_d), and the last
// combine's post-processing and completion signal.
static int controller_enqueue_direct(mppi_controller* c)
{
  const int T = c->T, C = c->C;
  const mppi_controller_desc& d = c->d;
  const bool vis = c->vis_random + c->vis_top > 0;
  const bool colored = d.sampler.kind == MPPI_SAMPLER_COLORED;
  ColoredPlan plan;
  if (colored) MPPI_CHECK(colored_plan(d.sampler, c->K, T, C, c->stream, &plan));
  WordSpan in[2];
  int n = 0;
  if (c->nominal_dirty) in[n++] = WordSpan{c->nominal_hd, c->nominal_d, T * C};
  // with the trajectory export and one iteration, the nominal is also the last iteration's sampling mean
  if (vis && d.num_iters == 1)
    in[n++] = WordSpan{c->nominal_dirty ? (const void*)c->nominal_hd : (const void*)c->nominal_d, c->vis_mean_d, T * C};
  if (n > 0) MPPI_CHECK(launch_copy_words(in, n, c->stream));
  ResultCopies out;
  out.mean = c->mean_out_hd;
  out.stats = c->stats_hd;
  out.signal = call_signal(c);
  out.control = c->control_out_hd;
  out.nominal = c->nominal_d;
  out.limits = control_limits(d.dynamics);
  out.T = T;
  out.C = C;
  c->x0_d_stale = true;
  for (int it = 0; it < d.num_iters; ++it) {
    if (vis && d.num_iters > 1 && it + 1 == d.num_iters) {
      const WordSpan m[] = {{c->mean_d, c->vis_mean_d, T * C}};
      MPPI_CHECK(launch_copy_words(m, 1, c->stream));
    }
    const IterRef ref{nullptr, c->iter_h[it]};
    MPPI_CHECK(launch_fused_iteration(d.dynamics, d.cost, d.sampler, c->K, T, d.dt, d.lambda, d.alpha, c->x0_hd,
                                      it == 0 ? c->nominal_d : c->mean_d, c->samples_d, !colored, keeps_samples(c),
                                      ref, c->costs_d, c->ws_d, c->stats_d, c->mean_d, nullptr, c->stream, nullptr,
                                      nullptr, it + 1 == d.num_iters ? &out : nullptr, colored ? &plan : nullptr));
  }
  return MPPI_OK;
}

// The proposal call: network inference on the uploaded state and nominal sequence, then one proposal iteration of
// prop_K samples (importance-weight corrected costs), the new mean, statistics and the warm start's cost to the host.
static int controller_enqueue_proposal(mppi_controller* c)
{
  const int T = c->T, C = c->C, S = c->S;
  const hipStream_t s = c->stream;
  const mppi_controller_desc& d = c->d;
  // the inference reads x0 and the nominal sequence from the host buffers and copies them (and the iteration
  // arguments) on for the rollout
  const InferStaging stg{c->x0_d, c->mean_d, c->iter_hd, c->iter_d};
  MPPI_CHECK(proposal_infer(c->prop, c->x0_hd, c->nominal_hd, c->sigma_base_d, c->prop_dmu_d, c->prop_sig_d, s, &stg));
  ResultCopies out;
  out.mean = c->mean_out_hd;
  out.stats = c->stats_hd;
  out.cost0_src = c->costs_d;
  out.cost0 = c->cost0_hd;
  out.signal = call_signal(c);
  return launch_fused_iteration(d.dynamics, d.cost, d.sampler, c->prop_K, T, d.dt, d.lambda, d.alpha, c->x0_d,
                                c->mean_d, c->samples_d, true, d.keep_samples != 0, IterRef{c->iter_d, {0, 0, 0}},
                                c->costs_d, c->ws_d, c->stats_d, c->mean_d, nullptr, s, c->prop_dmu_d, c->prop_sig_d,
                                &out);
}

// Runs one enqueue function eagerly on its first use and from a captured hipGraph afterwards (when use_graph), records
// ev1 on the stream (the end of the call's GPU work) and waits for the call's completion signal.
static int controller_run(mppi_controller* c, int (*enqueue)(mppi_controller*), hipGraph_t* graph,
                          hipGraphExec_t* exec, int* runs)
{
  if (c->d.use_graph && graph != nullptr && *runs > 0) {
    if (!*exec) {
      MPPI_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int st = enqueue(c);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (st != MPPI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      *graph = g;
      MPPI_HIP_CHECK(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
    }
    MPPI_HIP_CHECK(hipGraphLaunch(*exec, c->stream));
  } else {
    MPPI_CHECK(enqueue(c));
    (*runs)++;
  }
  const uint32_t want = ++c->signals;
  MPPI_HIP_CHECK(hipEventRecord(c->ev1, c->stream));
  c->timing_pending = true;
  const int st = wait_host_signal(c->done_h, want, c->ev1);
  if (st != MPPI_OK) c->signals = *c->done_h;  // resynchronised with what the device gave
  return st;
}

static void controller_set_iters(mppi_controller* c, int stride, int n)
{
  for (int it = 0; it < n; ++it) {
    c->iter_h[it].rng_offset = c->rng_counter++;
    c->iter_h[it].stride = stride;
    c->iter_h[it].iteration = it;
  }
}

// Applies the staged cost-map updates: values of the current extent are copied in place on the controller's stream
// (ordered after the work already queued there and before the call's own), so the captured graphs, which hold the map
// pointer and frame by value, stay valid; a new extent or frame changes the cost block, so the graphs are dropped and
// captured again on their next use.
static int apply_staged_map(mppi_controller* c)
{
  std::lock_guard<std::mutex> lock(c->map_mu);
  if (!c->stage_values && !c->stage_frame) return MPPI_OK;
  mppi_cost_params& cost = c->d.cost;
  const bool same_extent =
      !c->stage_values || (c->map_d != nullptr && c->stage_w == cost.map_width && c->stage_h == cost.map_height);
  if (c->stage_frame || !same_extent) {
    MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  if (c->stage_values) {
    const size_t n = (size_t)c->stage_w * c->stage_h;
    if (!same_extent) {
      if (c->map_d) MPPI_HIP_CHECK(hipFree(c->map_d));
      c->map_d = nullptr;
      cost.map = nullptr;
      MPPI_CHECK(dev_alloc(&c->map_d, sizeof(float) * n));
    }
    if (!c->map_ev) MPPI_HIP_CHECK(hipEventCreateWithFlags(&c->map_ev, hipEventDisableTiming));
    else MPPI_HIP_CHECK(hipEventSynchronize(c->map_ev));  // the previous upload has read the pinned source
    if (c->map_pinned_n < n) {
      if (c->map_pinned) MPPI_HIP_CHECK(hipHostFree(c->map_pinned));
      c->map_pinned = nullptr;
      c->map_pinned_n = 0;
      MPPI_CHECK(host_alloc(&c->map_pinned, sizeof(float) * n));
      c->map_pinned_n = n;
    }
    std::memcpy(c->map_pinned, c->map_stage.data(), sizeof(float) * n);
    MPPI_HIP_CHECK(hipMemcpyAsync(c->map_d, c->map_pinned, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    MPPI_HIP_CHECK(hipEventRecord(c->map_ev, c->stream));
    cost.map = c->map_d;
    cost.map_width = c->stage_w;
    cost.map_height = c->stage_h;
    c->stage_values = false;
  }
  if (c->stage_frame) {
    cost.map_origin_x = c->stage_origin[0];
    cost.map_origin_y = c->stage_origin[1];
    cost.map_resolution = c->stage_res;
    cost.map_rotated = c->stage_rotated;
    for (int i = 0; i < 4; ++i) cost.map_rotation[i] = c->stage_rot[i];
    c->stage_frame = false;
  }
  return MPPI_OK;
}

int mppi_controller_update_cost_map(mppi_controller* c, const float* values, int32_t w, int32_t h,
                                    int32_t column_major)
{
  if (!c || !values) return fail(MPPI_E_INVALID, "null argument");
  if (c->has_plugin) return fail(MPPI_E_UNSUPPORTED, "cost map update: plugin controllers have no built-in map");
  if (w < 2 || h < 2) return fail(MPPI_E_INVALID, "cost map update: width/height must be >= 2");
  if ((int64_t)w * h > ((int64_t)1 << 30)) return fail(MPPI_E_SIZE, "cost map update: map too large");
  std::vector<float> v((size_t)w * h);
  if (column_major) {
    for (int j = 0; j < w; ++j)
      for (int i = 0; i < h; ++i) v[(size_t)i * w + j] = values[(size_t)j * h + i];
  } else {
    std::memcpy(v.data(), values, sizeof(float) * v.size());
  }
  std::lock_guard<std::mutex> lock(c->map_mu);
  c->map_stage.swap(v);
  c->stage_w = w;
  c->stage_h = h;
  c->stage_values = true;
  return MPPI_OK;
}

int mppi_controller_update_cost_map_frame(mppi_controller* c, float origin_x, float origin_y, float resolution,
                                          const float* rotation)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (c->has_plugin) return fail(MPPI_E_UNSUPPORTED, "cost map update: plugin controllers have no built-in map");
  if (!(resolution > 0.0f)) return fail(MPPI_E_INVALID, "cost map update: resolution must be > 0");
  std::lock_guard<std::mutex> lock(c->map_mu);
  c->stage_origin[0] = origin_x;
  c->stage_origin[1] = origin_y;
  c->stage_res = resolution;
  c->stage_rotated = rotation != nullptr;
  const float id[4] = {1.0f, 0.0f, 0.0f, 1.0f};
  for (int i = 0; i < 4; ++i) c->stage_rot[i] = rotation ? rotation[i] : id[i];
  c->stage_frame = true;
  return MPPI_OK;
}

int mppi_controller_compute_control(mppi_controller* c, const float* x0_host, int32_t stride)
{
  if (!c || !x0_host) return fail(MPPI_E_INVALID, "null argument");
  if (stride < 0 || stride > c->T) return fail(MPPI_E_INVALID, "optimization_stride out of range");
  const int T = c->T, C = c->C, S = c->S, K = c->K;
  const mppi_controller_desc& d = c->d;
  if (d.sampler.kind == MPPI_SAMPLER_COLORED) {
    const int n = (d.sampler.fft_len_mode == MPPI_FFT_LEN_T) ? T : 2 * T;
    if (stride >= n) return fail(MPPI_E_INVALID, "colored sampler: optimization_stride out of range");
  }
  MPPI_CHECK(apply_staged_map(c));
  c->fe.previous_baseline = c->stats.baseline;
  const bool colored_ctrl = d.controller_kind == MPPI_CONTROLLER_COLORED;
  if (colored_ctrl && d.leash_active && c->computed && !c->has_plugin) {
    // the leash pulls x0 towards the previous solution's state at the last slide's step count
    // (ColoredMPPIController::computeControl, colored_mppi_controller.cu:151-154)
    std::vector<float> prev((size_t)T * S);
    MPPI_CHECK(mppi_controller_get_state_seq(c, prev.data()));
    const int j = std::min(std::max(c->leash_jump, 0), T - 1);
    enforce_leash(d.dynamics, x0_host, prev.data() + (size_t)j * S, d.state_leash_dist, c->x0_h);
  } else {
    for (int i = 0; i < S; ++i) c->x0_h[i] = x0_host[i];
  }
  std::copy(c->control.begin(), c->control.end(), c->nominal_h);
  if (!c->ev0) {
    MPPI_HIP_CHECK(hipEventCreate(&c->ev0));
    MPPI_HIP_CHECK(hipEventCreate(&c->ev1));
  }
  MPPI_HIP_CHECK(hipEventRecord(c->ev0, c->stream));
  bool plain = true;
  int K_used = K;
  if (c->prop) {
    // learned proposal: one iteration of prop_K samples; with the fallback enabled it is rejected when its
    // weighted mean cost exceeds the cost of the warm start (sample 0) by more than the tolerance, and the plain
    // iterations run instead
    controller_set_iters(c, stride, 1);
    MPPI_CHECK(controller_run(c, controller_enqueue_proposal, &c->prop_graph, &c->prop_exec, &c->prop_runs));
    c->n_prop++;
    c->prop_stats = *c->stats_h;
    c->prop_cost0 = *c->cost0_h;
    const double c0 = c->prop_cost0;
    plain = c->prop_tolerance >= 0.0f &&
            !(c->prop_stats.weighted_cost <= c0 + (double)c->prop_tolerance * std::fabs(c0));
    if (plain) c->n_fallback++;
    K_used = c->prop_K;
  }
  bool direct = false;
  if (plain) {
    controller_set_iters(c, stride, d.num_iters);
    direct = direct_path(c);
    if (direct) MPPI_CHECK(controller_run(c, controller_enqueue_direct, nullptr, nullptr, &c->eager_runs));
    else MPPI_CHECK(controller_run(c, controller_enqueue, &c->graph, &c->graph_exec, &c->eager_runs));
    K_used = K;
  }
  c->stats = *c->stats_h;
  c->raw_mean.assign(c->mean_out_h, c->mean_out_h + (size_t)T * C);
  free_energy_from_stats(c->stats, d.lambda, K_used, &c->fe);
  // the state trajectory of the smoothed controls (computeStateTrajectory) is propagated when first requested
  c->state_x0.assign(c->x0_h, c->x0_h + S);
  c->state_stale = true;
  if (direct) {
    // smoothed and constrained on the device, bit-identical to the host functions; the device holds it as the next
    // nominal
    c->control.assign(c->control_out_h, c->control_out_h + (size_t)T * C);
    c->state_u = c->raw_mean;
    c->state_smooth_pending = true;
    c->nominal_dirty = false;
    c->computed = true;
    return MPPI_OK;
  }
  c->nominal_dirty = true;
  c->control = c->raw_mean;
  smooth_control(c->control, T, C);
  c->state_u = c->control;
  c->state_smooth_pending = false;
  if (colored_ctrl) {
    // the colored controller clamps control 1 alone to its range (colored_mppi_controller.cu:226-231; a model with
    // one control has none to clamp)
    if (C >= 2)
      for (int t = 0; t < T; ++t) {
        float& u = c->control[(size_t)t * C + 1];
        u = fminf(fmaxf(u, d.dynamics.control_min[1]), d.dynamics.control_max[1]);
      }
  } else if (c->has_plugin) {
    plugin_clamp_sequence(c, c->control.data());
  } else {
    clamp_controls_seq(d.dynamics, c->control.data(), T);
  }
  c->computed = true;
  return MPPI_OK;
}

int mppi_controller_set_proposal(mppi_controller* c, mppi_proposal* proposal, int32_t num_rollouts,
                                 float fallback_tolerance)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (c->prop_exec) (void)hipGraphExecDestroy(c->prop_exec);
  if (c->prop_graph) (void)hipGraphDestroy(c->prop_graph);
  c->prop_exec = nullptr;
  c->prop_graph = nullptr;
  c->prop_runs = 0;
  if (!proposal) {
    c->prop = nullptr;
    return MPPI_OK;
  }
  if (c->has_plugin) return fail(MPPI_E_UNSUPPORTED, "proposal: not available with a dynamics plugin");
  if (c->vis_random + c->vis_top > 0)
    return fail(MPPI_E_UNSUPPORTED, "proposal: turn the sampled-trajectory export off first");
  const PropDims pd = proposal_dims(proposal_net_of(proposal));
  if (pd.S != c->S || pd.C != c->C || pd.T != c->T)
    return fail(MPPI_E_INVALID, "proposal: state/control/horizon dims do not match the controller");
  if (c->d.sampler.kind != MPPI_SAMPLER_GAUSSIAN) return fail(MPPI_E_UNSUPPORTED, "proposal: Gaussian sampler only");
  if (c->d.sampler.std_dev_t != nullptr)
    return fail(MPPI_E_UNSUPPORTED, "proposal: the base sampler has per-step std_dev (uniform std_dev only)");
  if (uses_tsallis(c->d)) return fail(MPPI_E_UNSUPPORTED, "proposal: not available with Tsallis weights");
  if (num_rollouts < 2 || num_rollouts > c->K)
    return fail(MPPI_E_INVALID, "proposal: num_rollouts must be in [2, controller num_rollouts]");
  const size_t tc = sizeof(float) * c->T * c->C;
  if (!c->prop_dmu_d) {
    MPPI_CHECK(dev_alloc(&c->prop_dmu_d, tc));
    MPPI_CHECK(dev_alloc(&c->prop_sig_d, tc));
    MPPI_CHECK(dev_alloc(&c->sigma_base_d, sizeof(float) * MPPI_MAX_CONTROL));
    MPPI_CHECK(host_alloc_mapped(&c->cost0_h, &c->cost0_hd, sizeof(float)));
  }
  MPPI_HIP_CHECK(hipMemcpy(c->sigma_base_d, c->d.sampler.std_dev, sizeof(float) * MPPI_MAX_CONTROL,
                           hipMemcpyHostToDevice));
  c->prop = proposal_net_of(proposal);
  c->prop_K = num_rollouts;
  c->prop_tolerance = fallback_tolerance;
  return MPPI_OK;
}

int mppi_controller_get_last_call_ms(const mppi_controller* c, float* ms)
{
  if (!c || !ms) return fail(MPPI_E_INVALID, "null argument");
  if (c->timing_pending) {
    // stream time of the call: from ev0 (before the first launch) to ev1 (recorded right after the last one)
    MPPI_HIP_CHECK(hipEventSynchronize(c->ev1));
    MPPI_HIP_CHECK(hipEventElapsedTime(&c->last_gpu_ms, c->ev0, c->ev1));
    c->timing_pending = false;
  }
  *ms = c->last_gpu_ms;
  return MPPI_OK;
}

int mppi_controller_get_proposal_stats(const mppi_controller* c, int32_t* proposal_calls, int32_t* fallbacks,
                                       mppi_weight_stats* last_proposal_stats, float* last_warm_start_cost)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (proposal_calls) *proposal_calls = c->n_prop;
  if (fallbacks) *fallbacks = c->n_fallback;
  if (last_proposal_stats) *last_proposal_stats = c->prop_stats;
  if (last_warm_start_cost) *last_warm_start_cost = c->prop_cost0;
  return MPPI_OK;
}

// One device-resident iteration sequence: the same kernels as computeControl, reading x0 and the mean already in
// device memory, warm-starting the next call from the new mean, with the rng offset advanced on the device.
// Iteration `it` of every call reads slot `it` of loop_iter_d (its own iteration index, so std_dev_decay^it applies as
// in computeControl) and advances that slot's rng offset by num_iters, so that call n, iteration it draws with offset
// base + n num_iters + it, the numbering of consecutive computeControl calls.
static int controller_enqueue_loop(mppi_controller* c)
{
  const int n = c->d.num_iters;
  for (int it = 0; it < n; ++it)
    MPPI_CHECK(enqueue_iteration(c, IterRef{c->loop_iter_d + it, {0, 0, 0}}, c->loop_iter_d + it, nullptr, n));
  return MPPI_OK;
}

// Device-resident calls replay in graphs of LOOP_BIG, then LOOP_GROUP calls (the same kernels in the same order as
// that many single-call replays), which leaves fewer graph-launch boundaries between the calls' kernels.
constexpr int LOOP_GROUP = 8;
constexpr int LOOP_BIG = 32;

static int capture_loop(mppi_controller* c, int calls, hipGraph_t* graph, hipGraphExec_t* exec)
{
  MPPI_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int st = MPPI_OK;
  for (int i = 0; i < calls && st == MPPI_OK; ++i) st = controller_enqueue_loop(c);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (st != MPPI_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  *graph = g;
  MPPI_HIP_CHECK(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
  return MPPI_OK;
}

// Arguments of the free-running loop (once after creation or a reseed): x0 as last uploaded, the mean left by the
// last call, stride 1, iteration it in slot it, rng offsets from a range computeControl does not reach.
static int loop_args_init(mppi_controller* c)
{
  if (c->loop_ready) return MPPI_OK;
  const int n = c->d.num_iters;
  std::vector<IterArgs> a((size_t)n);
  for (int it = 0; it < n; ++it) a[it] = IterArgs{c->rng_counter + ((uint64_t)1 << 40) + (uint64_t)it, 1, it};
  MPPI_HIP_CHECK(hipMemcpyAsync(c->loop_iter_d, a.data(), sizeof(IterArgs) * n, hipMemcpyHostToDevice, c->stream));
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  c->loop_ready = true;
  return MPPI_OK;
}

int mppi_controller_step_device(mppi_controller* c, int32_t num_calls)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (num_calls < 0) return fail(MPPI_E_INVALID, "num_calls must be >= 0");
  if (!c->computed)
    return fail(MPPI_E_INVALID, "step_device: call computeControl once first (the loop starts from its x0 and mean)");
  if (num_calls == 0) return MPPI_OK;
  MPPI_CHECK(apply_staged_map(c));
  MPPI_CHECK(refresh_x0_d(c));
  if (!c->loop_exec) {
    MPPI_CHECK(loop_args_init(c));
    MPPI_CHECK(controller_enqueue_loop(c));  // eager pass (sets kernel attributes before capture)
    MPPI_CHECK(capture_loop(c, 1, &c->loop_graph, &c->loop_exec));
    MPPI_CHECK(capture_loop(c, LOOP_GROUP, &c->loop_group_graph, &c->loop_group_exec));
    MPPI_CHECK(capture_loop(c, LOOP_BIG, &c->loop_big_graph, &c->loop_big_exec));
    --num_calls;
  }
  for (; c->loop_big_exec != nullptr && num_calls >= LOOP_BIG; num_calls -= LOOP_BIG)
    MPPI_HIP_CHECK(hipGraphLaunch(c->loop_big_exec, c->stream));
  for (; c->loop_group_exec != nullptr && num_calls >= LOOP_GROUP; num_calls -= LOOP_GROUP)
    MPPI_HIP_CHECK(hipGraphLaunch(c->loop_group_exec, c->stream));
  for (int i = 0; i < num_calls; ++i) MPPI_HIP_CHECK(hipGraphLaunch(c->loop_exec, c->stream));
  return MPPI_OK;
}

int mppi_controller_time_kernels(mppi_controller* c, int32_t reps, float* rollout_ms, float* combine_ms,
                                 int32_t* num_partials)
{
  if (!c || !rollout_ms || !combine_ms || reps < 1) return fail(MPPI_E_INVALID, "bad argument");
  if (!c->computed) return fail(MPPI_E_INVALID, "time_kernels: call computeControl once first");
  MPPI_CHECK(loop_args_init(c));
  MPPI_CHECK(refresh_x0_d(c));
  std::vector<hipEvent_t> ev((size_t)(2 * reps + 1), nullptr);
  int st = MPPI_OK;
  for (auto& e : ev)
    if (st == MPPI_OK && hipEventCreate(&e) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventCreate");
  if (st == MPPI_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventRecord");
  // iteration 0 of the device loop, launch after launch: rollout between ev[2r] and ev[2r+1], combine to ev[2r+2]
  for (int r = 0; r < reps && st == MPPI_OK; ++r) {
    int G = 0;
    st = enqueue_iteration(c, IterRef{c->loop_iter_d, {0, 0, 0}}, c->loop_iter_d, nullptr, c->d.num_iters,
                           ev[2 * r + 1], &G);
    if (num_partials != nullptr) *num_partials = G;
    if (st == MPPI_OK && hipEventRecord(ev[2 * r + 2], c->stream) != hipSuccess)
      st = fail(MPPI_E_HIP, "hipEventRecord");
  }
  float roll = 0.0f, comb = 0.0f;
  if (st == MPPI_OK && hipEventSynchronize(ev.back()) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventSynchronize");
  for (int r = 0; r < reps && st == MPPI_OK; ++r) {
    float a = 0.0f, b = 0.0f;
    if (hipEventElapsedTime(&a, ev[2 * r], ev[2 * r + 1]) != hipSuccess ||
        hipEventElapsedTime(&b, ev[2 * r + 1], ev[2 * r + 2]) != hipSuccess)
      st = fail(MPPI_E_HIP, "hipEventElapsedTime");
    roll += a;
    comb += b;
  }
  (void)hipStreamSynchronize(c->stream);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (st != MPPI_OK) return st;
  *rollout_ms = roll / reps;
  *combine_ms = comb / reps;
  return MPPI_OK;
}

int mppi_controller_get_control_seq(const mppi_controller* c, float* out)
{
  if (!c || !out) return fail(MPPI_E_INVALID, "null argument");
  std::copy(c->control.begin(), c->control.end(), out);
  return MPPI_OK;
}

int mppi_controller_get_raw_mean(const mppi_controller* c, float* out)
{
  if (!c || !out) return fail(MPPI_E_INVALID, "null argument");
  std::copy(c->raw_mean.begin(), c->raw_mean.end(), out);
  return MPPI_OK;
}

int mppi_controller_get_state_seq(const mppi_controller* c, float* out)
{
  if (!c || !out) return fail(MPPI_E_INVALID, "null argument");
  if (c->state_stale) {
    if (c->state_smooth_pending) {
      smooth_control(c->state_u, c->T, c->C);
      c->state_smooth_pending = false;
    }
    if (c->has_plugin) {
      plugin_propagate(c, c->state_x0.data(), c->state_u.data(), c->state_seq.data());
    } else {
      mppi_dynamics_params hostdyn = c->d.dynamics;
      if (!c->nn_host.empty()) hostdyn.nn_theta = c->nn_host.data();
      propagate(hostdyn, c->state_x0.data(), c->state_u.data(), c->T, c->d.dt, c->state_seq.data());
    }
    c->state_stale = false;
  }
  std::copy(c->state_seq.begin(), c->state_seq.end(), out);
  return MPPI_OK;
}

int mppi_controller_get_weight_stats(const mppi_controller* c, mppi_weight_stats* out)
{
  if (!c || !out) return fail(MPPI_E_INVALID, "null argument");
  *out = c->stats;
  return MPPI_OK;
}

int mppi_controller_get_free_energy(const mppi_controller* c, mppi_free_energy* out)
{
  if (!c || !out) return fail(MPPI_E_INVALID, "null argument");
  *out = c->fe;
  return MPPI_OK;
}

int mppi_controller_set_control_seq(mppi_controller* c, const float* u)
{
  if (!c || !u) return fail(MPPI_E_INVALID, "null argument");
  c->control.assign(u, u + (size_t)c->T * c->C);
  c->nominal_dirty = true;
  return MPPI_OK;
}

int mppi_controller_slide_control_seq(mppi_controller* c, int32_t steps)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (steps < 0) return fail(MPPI_E_INVALID, "steps must be >= 0");
  const int T = c->T, C = c->C;
  const float* zero = c->d.dynamics.zero_control;
  const float* scale = c->d.slide_control_scale;
  c->leash_jump = steps;
  c->nominal_dirty = true;
  for (int i = 0; i < T; ++i) {
    const int ind = std::min(i + steps, T - 1);
    for (int j = 0; j < C; ++j) {
      float v = c->control[(size_t)ind * C + j];
      if (i + steps > T - 1) v = (v - zero[j]) * scale[j] + zero[j];
      c->control[(size_t)i * C + j] = v;
    }
  }
  return MPPI_OK;
}

int mppi_controller_device_buffers(const mppi_controller* c, float** samples_d, float** costs_d, float** mean_d)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (samples_d) *samples_d = c->samples_d;
  if (costs_d) *costs_d = c->costs_d;
  if (mean_d) *mean_d = c->mean_d;
  return MPPI_OK;
}

int mppi_controller_set_seed(mppi_controller* c, uint64_t seed)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  // the seed is a by-value kernel argument of every captured graph (the control call's, the proposal call's and the
  // device loop's): wait for queued replays, then drop all of them so that each is captured again with the new seed
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->d.sampler.seed = seed;
  c->rng_counter = 0;
  c->loop_ready = false;
  return MPPI_OK;
}

int mppi_controller_set_params(mppi_controller* c, const mppi_controller_desc* desc)
{
  if (!c || !desc) return fail(MPPI_E_INVALID, "null argument");
  const mppi_controller_desc& o = c->d;
  if (desc->num_rollouts != o.num_rollouts || desc->num_timesteps != o.num_timesteps ||
      desc->num_iters != o.num_iters || desc->sampler.kind != o.sampler.kind ||
      desc->controller_kind != o.controller_kind ||
      (!c->has_plugin && (desc->dynamics.model != o.dynamics.model || desc->cost.cost != o.cost.cost)))
    return fail(MPPI_E_INVALID, "set_params: sizes, iteration count, model, cost, sampler or controller kind "
                                "changed (create a new controller)");
  if (!(desc->lambda > 0.0f)) return fail(MPPI_E_INVALID, "lambda must be > 0");
  MPPI_CHECK(validate_controller_kind(*desc, c->has_plugin));
  c->nominal_dirty = true;  // the next call uploads the host's sequence under the new parameters
  if (c->prop != nullptr && uses_tsallis(*desc))
    return fail(MPPI_E_UNSUPPORTED, "set_params: Tsallis weights with a learned proposal attached");
  mppi_controller_desc n = *desc;
  if (c->has_plugin) {
    if (desc->dynamics.control_dim != c->C) return fail(MPPI_E_INVALID, "set_params: control_dim changed");
    n.dynamics.model = 0;
    n.dynamics.state_dim = c->S;
    n.dynamics.output_dim = c->O;
    n.dynamics.nn_theta = nullptr;
    n.cost.map = nullptr;
  } else {
    MPPI_CHECK(validate_dynamics(n.dynamics));
  }
  MPPI_CHECK(validate_sampler(n.sampler, n.dynamics));
  // everything is checked before any device copy changes, so that a refused call leaves the controller as it was:
  // the blocks are validated with the host tables standing in for the device copies they will become
  const size_t nm = (size_t)n.cost.map_width * n.cost.map_height;
  if (!c->has_plugin) {
    if (desc->cost_map_host != nullptr && nm < 4) return fail(MPPI_E_INVALID, "cost map: width/height must be >= 2");
    if (n.dynamics.model == MPPI_MODEL_RACER_DUBINS_NN && desc->nn_theta_host != nullptr &&
        n.dynamics.nn_num_params != (int)c->nn_host.size())
      return fail(MPPI_E_INVALID, "set_params: residual MLP parameter count changed");
    mppi_cost_params probe = n.cost;
    // the controller's map copy stays in use unless the caller passes a device map of its own
    if (desc->cost_map_host != nullptr) probe.map = desc->cost_map_host;
    else if (c->map_d != nullptr && n.cost.map == nullptr) probe.map = c->map_d;
    if (desc->goal_traj_host != nullptr) probe.goal_traj = desc->goal_traj_host;
    MPPI_CHECK(validate_cost(probe, n.dynamics, true));
  }
  if (desc->std_dev_t_host != nullptr)
    for (size_t i = 0; i < (size_t)c->T * c->C; ++i)
      if (!(desc->std_dev_t_host[i] > 0.0f)) return fail(MPPI_E_INVALID, "per-step std_dev must be > 0");
  if (c->prop != nullptr && (desc->std_dev_t_host != nullptr || n.sampler.std_dev_t != nullptr))
    return fail(MPPI_E_UNSUPPORTED, "set_params: per-step std_dev with a learned proposal attached");
  // queued replays read the old parameters: wait for them; every captured graph holds the old parameter blocks as
  // kernel arguments (and may point at copies replaced below)
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  if (!c->has_plugin) {
    if (desc->cost_map_host != nullptr) {
      const size_t old = (size_t)o.cost.map_width * o.cost.map_height;
      if (c->map_d == nullptr || nm != old) {
        if (c->map_d) MPPI_HIP_CHECK(hipFree(c->map_d));
        c->map_d = nullptr;
        c->d.cost.map = nullptr;
        MPPI_CHECK(dev_alloc(&c->map_d, sizeof(float) * nm));
      }
      MPPI_HIP_CHECK(hipMemcpy(c->map_d, desc->cost_map_host, sizeof(float) * nm, hipMemcpyHostToDevice));
    }
    if (c->map_d != nullptr && desc->cost_map_host == nullptr && n.cost.map == nullptr) {
      // the controller's own copy stays, with its own extent (the one it was uploaded with)
      n.cost.map = c->map_d;
      n.cost.map_width = o.cost.map_width;
      n.cost.map_height = o.cost.map_height;
    } else if (c->map_d != nullptr && desc->cost_map_host != nullptr) {
      n.cost.map = c->map_d;
    }
    if (n.dynamics.model == MPPI_MODEL_RACER_DUBINS_NN) {
      if (desc->nn_theta_host != nullptr) {
        c->nn_host.assign(desc->nn_theta_host, desc->nn_theta_host + c->nn_host.size());
        MPPI_HIP_CHECK(hipMemcpy(c->nn_d, c->nn_host.data(), sizeof(float) * c->nn_host.size(),
                                 hipMemcpyHostToDevice));
      }
      n.dynamics.nn_theta = c->nn_d;
    }
  }
  MPPI_CHECK(upload_tables(c, desc, n));
  n.init_control = nullptr;
  n.cost_map_host = nullptr;
  n.nn_theta_host = nullptr;
  const bool reseed = n.sampler.seed != o.sampler.seed;
  c->d = n;
  if (reseed) {
    c->rng_counter = 0;
    c->loop_ready = false;
  }
  if (c->prop != nullptr) {
    MPPI_HIP_CHECK(hipMemcpy(c->sigma_base_d, c->d.sampler.std_dev, sizeof(float) * MPPI_MAX_CONTROL,
                             hipMemcpyHostToDevice));
  }
  return MPPI_OK;
}

int mppi_controller_set_sampled_trajectories(mppi_controller* c, int32_t num_random, int32_t num_top)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (num_random < 0 || num_top < 0) return fail(MPPI_E_INVALID, "sampled trajectories: negative count");
  const int n = 1 + num_random + num_top;
  const int pool = (int)(0.98 * c->K) - 1;  // random picks come from samples 1 .. 0.98 K - 1
  if (n > std::min(c->K, MPPI_MAX_SAMPLED_TRAJECTORIES) || num_random > std::max(pool, 0) || num_top > c->K)
    return fail(MPPI_E_SIZE, "sampled trajectories: too many for the sample count");
  if (c->prop != nullptr || c->has_plugin)
    return fail(MPPI_E_UNSUPPORTED, "sampled trajectories: not available with a learned proposal or a plugin");
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  // the control call's graph changes (samples kept, mean saved): capture it again
  drop_graphs(c);
  for (void* p : {(void*)c->vis_mean_d, (void*)c->vis_idx_d, (void*)c->vis_buf_d})
    if (p) MPPI_HIP_CHECK(hipFree(p));
  c->vis_mean_d = nullptr;
  c->vis_idx_d = nullptr;
  c->vis_buf_d = nullptr;
  c->vis_random = c->vis_top = 0;
  if (num_random + num_top == 0) return MPPI_OK;
  const int T = c->T;
  MPPI_CHECK(dev_alloc(&c->vis_mean_d, sizeof(float) * T * c->C));
  MPPI_CHECK(dev_alloc(&c->vis_idx_d, sizeof(int) * n));
  MPPI_CHECK(dev_alloc(&c->vis_buf_d, sizeof(float) * (size_t)n * (T * c->C + T * c->O + T + 1)));
  c->vis_random = num_random;
  c->vis_top = num_top;
  return MPPI_OK;
}

int mppi_controller_get_sampled_trajectories(mppi_controller* c, int32_t* count, int32_t* indices, float* controls,
                                             float* outputs, float* step_costs, float* terminal_costs,
                                             float* weights)
{
  if (!c || !count) return fail(MPPI_E_INVALID, "null argument");
  if (c->vis_random + c->vis_top == 0)
    return fail(MPPI_E_INVALID, "sampled trajectories: enable them with mppi_controller_set_sampled_trajectories");
  if (!c->computed) return fail(MPPI_E_INVALID, "sampled trajectories: call computeControl first");
  const int K = c->K, T = c->T, C = c->C, O = c->O;
  const int n = 1 + c->vis_random + c->vis_top;
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  std::vector<float> cost((size_t)K);
  MPPI_HIP_CHECK(hipMemcpy(cost.data(), c->costs_d, sizeof(float) * K, hipMemcpyDeviceToHost));
  for (auto& v : cost)
    if (!(v <= FLT_MAX)) v = INFINITY;
  std::vector<int> sel;
  sel.reserve((size_t)n);
  sel.push_back(0);  // sample 0 is the nominal sequence itself
  // random samples without replacement from 1 .. 0.98 K - 1 (the last ones are pure noise), partial Fisher-Yates
  {
    std::mt19937_64 gen(c->d.sampler.seed * 0x9E3779B97F4A7C15ull + 0x5851F42D4C957F2Dull * ++c->vis_draws);
    std::vector<int> pool((size_t)((int)(0.98 * K) - 1));
    std::iota(pool.begin(), pool.end(), 1);
    for (int i = 0; i < c->vis_random; ++i) {
      std::uniform_int_distribution<size_t> pick((size_t)i, pool.size() - 1);
      std::swap(pool[(size_t)i], pool[pick(gen)]);
      sel.push_back(pool[(size_t)i]);
    }
  }
  // the num_top samples of largest weight (lowest cost, the lower index first on ties), by decreasing weight
  {
    std::vector<int> order((size_t)K);
    std::iota(order.begin(), order.end(), 0);
    std::partial_sort(order.begin(), order.begin() + c->vis_top, order.end(), [&](int a, int b) {
      return cost[(size_t)a] < cost[(size_t)b] || (cost[(size_t)a] == cost[(size_t)b] && a < b);
    });
    for (int i = 0; i < c->vis_top; ++i) sel.push_back(order[(size_t)i]);
  }
  MPPI_HIP_CHECK(hipMemcpy(c->vis_idx_d, sel.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  const mppi_controller_desc& d = c->d;
  float* ctl = c->vis_buf_d;
  float* out = ctl + (size_t)n * T * C;
  float* stp = out + (size_t)n * T * O;
  float* term = stp + (size_t)n * T;
  MPPI_CHECK(refresh_x0_d(c));
  MPPI_CHECK(launch_vis_rollouts(d.dynamics, d.cost, d.sampler, K, T, d.dt, d.lambda, d.alpha, c->x0_d, c->vis_mean_d,
                                 c->samples_d, c->vis_idx_d, n, ctl, out, stp, term, c->stream));
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (controls) MPPI_HIP_CHECK(hipMemcpy(controls, ctl, sizeof(float) * n * T * C, hipMemcpyDeviceToHost));
  if (outputs) MPPI_HIP_CHECK(hipMemcpy(outputs, out, sizeof(float) * n * T * O, hipMemcpyDeviceToHost));
  if (step_costs) MPPI_HIP_CHECK(hipMemcpy(step_costs, stp, sizeof(float) * n * T, hipMemcpyDeviceToHost));
  if (terminal_costs) MPPI_HIP_CHECK(hipMemcpy(terminal_costs, term, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (indices) std::copy(sel.begin(), sel.end(), indices);
  if (weights) {
    // normalized weights w_k / eta of the last iteration: the device's own weights with the Tsallis transform, else
    // exp(-(c_k - b) / lambda) / eta over its costs (core/mppi_common.cu:994-1002)
    const double b = c->stats.baseline, eta = c->stats.normalizer, inv = 1.0 / (double)d.lambda;
    std::vector<float> w;
    if (uses_tsallis(d)) {
      w.resize(c->K);
      MPPI_HIP_CHECK(hipMemcpy(w.data(), c->w_d, sizeof(float) * c->K, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < n; ++i) {
      const size_t k = (size_t)sel[(size_t)i];
      const double ck = cost[k];
      weights[i] = w.empty() ? (float)((ck <= FLT_MAX) ? std::exp(-inv * (ck - b)) / eta : 0.0)
                             : (float)((double)w[k] / eta);
    }
  }
  *count = n;
  return MPPI_OK;
}

int mppi_controller_stream(const mppi_controller* c, void** stream)
{
  if (!c || !stream) return fail(MPPI_E_INVALID, "null argument");
  *stream = (void*)c->stream;
  return MPPI_OK;
}

}  // extern "C"
