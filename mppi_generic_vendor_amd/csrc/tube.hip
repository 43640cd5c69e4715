// Two-system controllers: TubeMPPIController::computeControl (controllers/Tube-MPPI/tube_mppi_controller.cu:158-295)
// and the optimisation of RobustMPPIController::computeControl (controllers/R-MPPI/robust_mppi_controller.cu:635-756).
//
// Per iteration, on the controller's stream: the two distributions' samples (the same normals, each around its own
// mean), one two-system rollout launch (both systems per sample in one kernel), then per system the weight transform
// (baseline, normExp weights, FP64 normalizer) and the weighted update; the new means, the statistics and the costs
// come back in one synchronisation. The nominal-state rule, the smoothing and the state trajectories are host work on
// [T][C] / [T][S] arrays, as in the reference.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fused.hpp"
#include "host.hpp"
#include "kernels.hpp"
#include "models.hpp"

using namespace mppi_amd;

struct mppi_tube_controller {
  mppi_tube_desc d{};
  mppi_controller_desc& b() { return d.base; }
  const mppi_controller_desc& b() const { return d.base; }
  int S = 0, C = 0, O = 0, K = 0, T = 0;
  hipStream_t stream = nullptr;
  float* x0_d = nullptr;       // [2][S]
  float* mean_d = nullptr;     // [2][T][C] sampling means
  float* out_d = nullptr;      // [2][T][C] updated means
  float* samples_d = nullptr;  // [2][K][T][C]
  float* costs_d = nullptr;    // [2][K]
  float* w_d = nullptr;        // [2][K]
  float* gains_d = nullptr;    // [T][S][C]
  float* map_d = nullptr;
  float* nn_d = nullptr;
  float* std_t_d = nullptr;
  float* goal_d = nullptr;
  mppi_weight_stats* stats_d = nullptr;  // [2]
  void* ws_d = nullptr;
  // mapped pinned exchange buffers (the *_hd members are their device-side addresses): x0 [2][MAX_STATE], means
  // [2][T][C], outs [2][T][C], stats [2]; for RMPPI's candidate evaluation candidates [N][S], strides [N], costs [N][M]
  float* host_io = nullptr;
  float* host_io_hd = nullptr;
  float* isc_io = nullptr;
  float* isc_io_hd = nullptr;
  bool gains_dirty = false;  // the host gains changed since they were uploaded
  // fused iteration (Gaussian samples): per system, the workgroups' Partials [2][gstride] and rows [2][gstride][T C]
  Partial* parts_d = nullptr;
  float* psums_d = nullptr;
  int gstride = 0;
  // completion signal of a call (device workgroup count and epoch, mapped host word), signals enqueued, fallback event
  uint32_t* sig_d = nullptr;
  uint32_t* done_h = nullptr;
  uint32_t* done_hd = nullptr;
  uint32_t signals = 0;
  hipEvent_t ev = nullptr;
  std::vector<float> control[2];  // system d's control sequence (Tube: 0 actual, 1 nominal; RMPPI: 0 nominal, 1 real)
  std::vector<float> states[2];   // system d's state trajectory
  std::vector<float> gains;       // host copy [T][S][C] (empty: none)
  std::vector<float> nominal_x;   // RMPPI nominal state / Tube nominal trajectory start
  std::vector<float> nn_host;     // host copy of the network parameters (host state propagation)
  mppi_dynamics_params dyn_host{};  // the dynamics with host pointers, for the host-side steps
  bool nominal_init = false;
  int nominal_state_used = 1;
  mppi_weight_stats stats[2]{};
  mppi_free_energy fe[2]{};
  uint64_t rng_counter = 0;
  IterArgs last_args{};  // generator offset, stride and iteration of the last iteration run
  bool have_samples = false;
  // RMPPI candidate evaluation
  int N = 0, M = 0;
  float* cand_d = nullptr;   // [N][S]
  int* stride_d = nullptr;   // [N]
  float* eval_d = nullptr;   // [N][M]
  std::vector<float> lsw;    // line-search weights [3][N]
  std::vector<float> cand_h, eval_h, cand_fe;
  std::vector<int> strides_h;
  int best_index = 0, nominal_stride = 0;
  bool isc_init = false;
  bool rmppi() const { return d.mode == MPPI_TWO_SYSTEM_RMPPI; }
  // index of the nominal system in the buffers
  int nom() const { return rmppi() ? 0 : 1; }
};

namespace {

int tube_free(mppi_tube_controller* c)
{
  if (!c) return MPPI_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* devs[] = {c->x0_d, c->mean_d, c->out_d, c->samples_d, c->costs_d, c->w_d, c->gains_d, c->map_d, c->nn_d,
                  c->std_t_d, c->goal_d, c->stats_d, c->ws_d, c->cand_d, c->stride_d, c->eval_d, c->sig_d,
                  c->parts_d, c->psums_d};
  for (void* p : devs)
    if (p) (void)hipFree(p);
  for (void* p : {(void*)c->host_io, (void*)c->isc_io, (void*)c->done_h})
    if (p) (void)hipHostFree(p);
  if (c->ev) (void)hipEventDestroy(c->ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MPPI_OK;
}

size_t io_floats(int T, int C) { return 2 * MPPI_MAX_STATE + 4 * (size_t)T * C; }

// The completion signal of the call's last kernel.
HostSignal tube_signal(const mppi_tube_controller* c)
{
  HostSignal h;
  h.count = c->sig_d;
  h.epoch = c->sig_d + 1;
  h.flag = c->done_hd;
  return h;
}

// Waits for the last of the `n` signals of the work just enqueued.
int tube_wait(mppi_tube_controller* c, uint32_t n = 1)
{
  c->signals += n;
  const uint32_t want = c->signals;
  MPPI_HIP_CHECK(hipEventRecord(c->ev, c->stream));
  const int st = wait_host_signal(c->done_h, want, c->ev);
  if (st != MPPI_OK) c->signals = *c->done_h;
  return st;
}

int tube_setup(mppi_tube_controller* c, const mppi_tube_desc* desc)
{
  c->d = *desc;
  mppi_controller_desc& b = c->b();
  if (desc->mode != MPPI_TWO_SYSTEM_TUBE && desc->mode != MPPI_TWO_SYSTEM_RMPPI)
    return fail(MPPI_E_INVALID, "two-system controller: mode must be MPPI_TWO_SYSTEM_TUBE or MPPI_TWO_SYSTEM_RMPPI");
  c->K = b.num_rollouts;
  c->T = b.num_timesteps;
  MPPI_CHECK(validate_sizes(c->K, c->T));
  MPPI_CHECK(validate_dynamics(b.dynamics));
  MPPI_CHECK(validate_sampler(b.sampler, b.dynamics));
  if (b.dynamics.model == MPPI_MODEL_RACER_DUBINS_NN && b.dynamics.nn_precision == MPPI_NN_BF16)
    return fail(MPPI_E_UNSUPPORTED, "two-system controller: the BF16 network model is not supported");
  if (!(b.lambda > 0.0f)) return fail(MPPI_E_INVALID, "lambda must be > 0");
  if (b.num_iters < 1) return fail(MPPI_E_INVALID, "num_iters must be >= 1");
  if (!(desc->nominal_threshold == desc->nominal_threshold) ||
      !(desc->value_function_threshold == desc->value_function_threshold))
    return fail(MPPI_E_INVALID, "two-system controller: thresholds must not be NaN");
  c->S = b.dynamics.state_dim;
  c->C = b.dynamics.control_dim;
  c->O = b.dynamics.output_dim;
  const int K = c->K, T = c->T, C = c->C, S = c->S;
  MPPI_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  MPPI_CHECK(dev_alloc(&c->x0_d, sizeof(float) * 2 * S));
  MPPI_CHECK(dev_alloc(&c->mean_d, sizeof(float) * 2 * T * C));
  MPPI_CHECK(dev_alloc(&c->out_d, sizeof(float) * 2 * T * C));
  MPPI_CHECK(dev_alloc(&c->samples_d, sizeof(float) * 2 * (size_t)K * T * C));
  MPPI_CHECK(dev_alloc(&c->costs_d, sizeof(float) * 2 * K));
  MPPI_CHECK(dev_alloc(&c->w_d, sizeof(float) * 2 * K));
  MPPI_CHECK(dev_alloc(&c->stats_d, sizeof(mppi_weight_stats) * 2));
  MPPI_CHECK(dev_alloc(&c->ws_d, weighted_update_workspace_bytes(K, T, C)));
  // partials of the fused iteration: at most ceil(K / 16) workgroups per system
  c->gstride = (K + 15) / 16;
  MPPI_CHECK(dev_alloc(&c->parts_d, sizeof(Partial) * 2 * (size_t)c->gstride));
  MPPI_CHECK(dev_alloc(&c->psums_d, sizeof(float) * 2 * (size_t)c->gstride * T * C));
  MPPI_CHECK(dev_alloc(&c->gains_d, sizeof(float) * (size_t)T * S * C));
  MPPI_HIP_CHECK(hipMemset(c->gains_d, 0, sizeof(float) * (size_t)T * S * C));
  MPPI_CHECK(host_alloc_mapped(&c->host_io, &c->host_io_hd,
                               sizeof(float) * io_floats(T, C) + 2 * sizeof(mppi_weight_stats)));
  MPPI_CHECK(host_alloc_mapped(&c->done_h, &c->done_hd, sizeof(uint32_t)));
  MPPI_CHECK(dev_alloc(&c->sig_d, 2 * sizeof(uint32_t)));
  MPPI_HIP_CHECK(hipMemset(c->sig_d, 0, 2 * sizeof(uint32_t)));
  MPPI_HIP_CHECK(hipEventCreateWithFlags(&c->ev, hipEventDisableTiming));
  if (b.cost_map_host != nullptr) {
    const size_t n = (size_t)b.cost.map_width * b.cost.map_height;
    if (n < 4) return fail(MPPI_E_INVALID, "cost map: width/height must be >= 2");
    MPPI_CHECK(dev_alloc(&c->map_d, sizeof(float) * n));
    MPPI_HIP_CHECK(hipMemcpy(c->map_d, b.cost_map_host, sizeof(float) * n, hipMemcpyHostToDevice));
    b.cost.map = c->map_d;
  }
  if (b.std_dev_t_host != nullptr) {
    const size_t n = (size_t)T * C;
    for (size_t i = 0; i < n; ++i)
      if (!(b.std_dev_t_host[i] > 0.0f)) return fail(MPPI_E_INVALID, "per-step std_dev must be > 0");
    MPPI_CHECK(dev_alloc(&c->std_t_d, sizeof(float) * n));
    MPPI_HIP_CHECK(hipMemcpy(c->std_t_d, b.std_dev_t_host, sizeof(float) * n, hipMemcpyHostToDevice));
    b.sampler.std_dev_t = c->std_t_d;
  }
  if (b.goal_traj_host != nullptr) {
    const int H = b.cost.goal_traj_len;
    if (H < 1) return fail(MPPI_E_INVALID, "goal trajectory: goal_traj_len must be >= 1");
    MPPI_CHECK(dev_alloc(&c->goal_d, sizeof(float) * (size_t)H * c->O));
    MPPI_HIP_CHECK(hipMemcpy(c->goal_d, b.goal_traj_host, sizeof(float) * (size_t)H * c->O, hipMemcpyHostToDevice));
    b.cost.goal_traj = c->goal_d;
  }
  MPPI_CHECK(validate_cost(b.cost, b.dynamics, true));
  if (b.dynamics.model == MPPI_MODEL_RACER_DUBINS_NN) {
    if (b.nn_theta_host == nullptr) return fail(MPPI_E_INVALID, "residual MLP: nn_theta_host is required");
    const size_t n = (size_t)b.dynamics.nn_num_params;
    MPPI_CHECK(dev_alloc(&c->nn_d, sizeof(float) * n));
    MPPI_HIP_CHECK(hipMemcpy(c->nn_d, b.nn_theta_host, sizeof(float) * n, hipMemcpyHostToDevice));
    c->nn_host.assign(b.nn_theta_host, b.nn_theta_host + n);
    b.dynamics.nn_theta = c->nn_d;
  }
  c->dyn_host = b.dynamics;
  if (!c->nn_host.empty()) c->dyn_host.nn_theta = c->nn_host.data();
  for (int d = 0; d < 2; ++d) {
    c->control[d].assign((size_t)T * C, 0.0f);
    if (b.init_control) std::copy(b.init_control, b.init_control + (size_t)T * C, c->control[d].begin());
    c->states[d].assign((size_t)T * S, 0.0f);
  }
  c->nominal_x.assign(S, 0.0f);
  if (c->rmppi()) {
    c->N = desc->num_candidates ? desc->num_candidates : 9;
    c->M = desc->num_eval_samples ? desc->num_eval_samples : 64;
    if (c->N < 3 || c->N % 2 == 0 || c->N > 1023)
      return fail(MPPI_E_INVALID, "RMPPI: num_candidates must be odd, >= 3 and < 1024");
    if (c->M < 1 || c->M > K) return fail(MPPI_E_INVALID, "RMPPI: num_eval_samples must be in [1, num_rollouts]");
    MPPI_CHECK(dev_alloc(&c->cand_d, sizeof(float) * (size_t)c->N * S));
    MPPI_CHECK(dev_alloc(&c->stride_d, sizeof(int) * c->N));
    MPPI_CHECK(dev_alloc(&c->eval_d, sizeof(float) * (size_t)c->N * c->M));
    // computeLineSearchWeights (robust_mppi_controller.cu:480-498): [3][N]
    const int h = c->N / 2;
    c->lsw.assign((size_t)3 * c->N, 0.0f);
    for (int i = 0; i < h + 1; ++i) {
      c->lsw[0 * c->N + i] = 1 - i / float(h);
      c->lsw[1 * c->N + i] = i / float(h);
    }
    for (int i = 1; i < h + 1; ++i) {
      c->lsw[1 * c->N + h + i] = 1 - i / float(h);
      c->lsw[2 * c->N + h + i] = i / float(h);
    }
    MPPI_CHECK(host_alloc_mapped(&c->isc_io, &c->isc_io_hd,
                                 sizeof(float) * ((size_t)c->N * S + (size_t)c->N + (size_t)c->N * c->M)));
    c->cand_h.assign((size_t)c->N * S, 0.0f);
    c->strides_h.assign(c->N, 0);
    c->eval_h.assign((size_t)c->N * c->M, 0.0f);
    c->cand_fe.assign(c->N, 0.0f);
  }
  b.init_control = nullptr;
  b.cost_map_host = nullptr;
  b.nn_theta_host = nullptr;
  b.std_dev_t_host = nullptr;
  b.goal_traj_host = nullptr;
  return MPPI_OK;
}

// One optimisation iteration: uploads, samples, two-system rollout, weights and updates of both systems, downloads.
int tube_iteration(mppi_tube_controller* c, const float* x0_host, int stride, int iter)
{
  const mppi_controller_desc& b = c->b();
  const int K = c->K, T = c->T, C = c->C, S = c->S;
  const size_t TC = (size_t)T * C;
  float* x0_h = c->host_io;
  float* mean_h = x0_h + 2 * MPPI_MAX_STATE;
  float* out_h = mean_h + 2 * TC;
  mppi_weight_stats* st_h = reinterpret_cast<mppi_weight_stats*>(c->host_io + io_floats(T, C));
  float* x0_hd = c->host_io_hd;
  float* mean_hd = x0_hd + 2 * MPPI_MAX_STATE;
  float* out_hd = mean_hd + 2 * TC;
  mppi_weight_stats* st_hd = reinterpret_cast<mppi_weight_stats*>(c->host_io_hd + io_floats(T, C));
  const int nom = c->nom(), act = 1 - nom;
  // initial states: the measured state for the actual / real system, the nominal state for the nominal one
  for (int i = 0; i < S; ++i) {
    x0_h[act * S + i] = x0_host[i];
    x0_h[nom * S + i] = c->rmppi() ? c->nominal_x[i] : c->states[nom][i];
  }
  // sampling means: Tube -- each system's own control; RMPPI -- the nominal control for both (:657-659)
  for (int d = 0; d < 2; ++d) {
    const std::vector<float>& m = c->rmppi() ? c->control[nom] : c->control[d];
    std::copy(m.begin(), m.end(), mean_h + d * TC);
  }
  // the states and means from the mapped buffers by one copy kernel; the gains (pageable) only when they changed
  const WordSpan in[] = {{x0_hd, c->x0_d, 2 * S}, {mean_hd, c->mean_d, (int)(2 * TC)}};
  MPPI_CHECK(launch_copy_words(in, 2, c->stream));
  if (c->rmppi() && !c->gains.empty() && c->gains_dirty) {
    MPPI_HIP_CHECK(hipMemcpyAsync(c->gains_d, c->gains.data(), sizeof(float) * c->gains.size(),
                                  hipMemcpyHostToDevice, c->stream));
    c->gains_dirty = false;
  }
  const IterRef ref{nullptr, {c->rng_counter, stride, iter}};
  c->last_args = ref.val;
  mppi_two_system_params tp{};
  tp.mode = c->d.mode;
  tp.nominal_index = nom;
  tp.value_function_threshold = c->d.value_function_threshold;
  tp.feedback_gains = (c->rmppi() && !c->gains.empty()) ? c->gains_d : nullptr;
  if (b.sampler.kind == MPPI_SAMPLER_GAUSSIAN) {
    // the fused iteration: both systems' samples drawn in the rollout kernel, per-workgroup partials, and one combine
    // launch over both systems writing the new means and statistics to the mapped buffers, then the signal
    int G = 0;
    MPPI_CHECK(launch_two_system_fused(b.dynamics, b.cost, b.sampler, tp, K, T, b.dt, b.lambda, b.alpha, c->x0_d,
                                       c->mean_d, c->samples_d, c->costs_d, ref, c->parts_d, c->psums_d, c->gstride,
                                       &G, c->stream));
    ResultCopies rc;
    rc.mean = out_hd;
    rc.stats = st_hd;
    rc.signal = tube_signal(c);
    MPPI_CHECK(launch_combine(c->parts_d, c->psums_d, G, K, (int)TC, b.lambda, c->stats_d, c->out_d, nullptr, 1, &rc,
                              c->stream, 2, c->gstride));
    MPPI_CHECK(tube_wait(c));
  } else {
    for (int d = 0; d < 2; ++d) {
      float* sd = c->samples_d + (size_t)d * K * TC;
      if (b.sampler.kind == MPPI_SAMPLER_COLORED)
        MPPI_CHECK(launch_sample_colored(b.sampler, K, T, C, c->mean_d + d * TC, sd, ref, nullptr, c->stream));
      else
        MPPI_CHECK(launch_sample_gaussian(b.sampler, K, T, C, c->mean_d + d * TC, sd, ref, nullptr, c->stream));
    }
    MPPI_CHECK(launch_two_system(b.dynamics, b.cost, b.sampler, b.sampler, tp, K, T, b.dt, b.lambda, b.alpha, c->x0_d,
                                 c->mean_d, c->samples_d, c->costs_d, c->stream));
    for (int d = 0; d < 2; ++d) {
      MPPI_CHECK(launch_weights(c->costs_d + (size_t)d * K, c->w_d + (size_t)d * K, K, b.lambda, 0.0f, 0.0f,
                                c->stats_d + d, nullptr, nullptr, 0, c->stream));
      MPPI_CHECK(launch_weighted_update(c->w_d + (size_t)d * K, c->stats_d + d, c->samples_d + (size_t)d * K * TC, K, T,
                                        C, c->out_d + d * TC, c->ws_d, c->stream));
    }
    // both new means and statistics to the mapped buffers, then the completion signal
    const WordSpan res[] = {{c->out_d, out_hd, (int)(2 * TC)}, {c->stats_d, st_hd, words_of<mppi_weight_stats>(2)}};
    const HostSignal sig = tube_signal(c);
    MPPI_CHECK(launch_copy_words(res, 2, c->stream, &sig));
    MPPI_CHECK(tube_wait(c));
  }
  c->rng_counter++;
  c->have_samples = true;
  for (int d = 0; d < 2; ++d) {
    std::copy(out_h + d * TC, out_h + (d + 1) * TC, c->control[d].begin());
    c->stats[d] = st_h[d];
    free_energy_from_stats(st_h[d], b.lambda, K, &c->fe[d]);
  }
  return MPPI_OK;
}

bool check_nominal(int32_t v) { return v == 0 || v == 1; }

}  // namespace

extern "C" {

int mppi_tube_create(const mppi_tube_desc* desc, mppi_tube_controller** out)
{
  if (!desc || !out) return fail(MPPI_E_INVALID, "null argument");
  *out = nullptr;
  mppi_tube_controller* c = new (std::nothrow) mppi_tube_controller();
  if (!c) return fail(MPPI_E_NOMEM, "host allocation failed");
  const int st = tube_setup(c, desc);
  if (st != MPPI_OK) {
    const std::string msg = mppi_last_error();
    tube_free(c);
    set_last_error(msg);
    return st;
  }
  *out = c;
  return MPPI_OK;
}

int mppi_tube_destroy(mppi_tube_controller* ctrl) { return tube_free(ctrl); }

int mppi_tube_compute_control(mppi_tube_controller* c, const float* x0_host, int32_t stride)
{
  if (!c || !x0_host) return fail(MPPI_E_INVALID, "null argument");
  if (stride < 0) return fail(MPPI_E_INVALID, "optimization_stride must be >= 0");
  const mppi_controller_desc& b = c->b();
  const int T = c->T, S = c->S, C = c->C;
  const int nom = c->nom(), act = 1 - nom;
  for (int d = 0; d < 2; ++d) c->fe[d].previous_baseline = c->stats[d].baseline;
  if (!c->rmppi() && !c->nominal_init) {
    // the nominal trajectory starts at the measured state (:161-166)
    std::copy(x0_host, x0_host + S, c->states[nom].begin());
    c->nominal_init = true;
  }
  for (int it = 0; it < b.num_iters; ++it) {
    MPPI_CHECK(tube_iteration(c, x0_host, stride, it));
    if (!c->rmppi()) {
      // computeStateTrajectory(state), then the nominal-state rule (:258-276)
      const std::vector<float> x_nom(c->states[nom].begin(), c->states[nom].begin() + S);
      propagate(c->dyn_host, x_nom.data(), c->control[nom].data(), T, b.dt, c->states[nom].data());
      propagate(c->dyn_host, x0_host, c->control[act].data(), T, b.dt, c->states[act].data());
      if (c->stats[act].baseline < c->stats[nom].baseline + c->d.nominal_threshold) {
        c->nominal_state_used = 0;
        c->states[nom] = c->states[act];
        c->control[nom] = c->control[act];
      } else {
        c->nominal_state_used = 1;
      }
    }
  }
  if (!c->rmppi()) {
    smooth_control(c->control[nom], T, C);
    const std::vector<float> x_nom(c->states[nom].begin(), c->states[nom].begin() + S);
    propagate(c->dyn_host, x_nom.data(), c->control[nom].data(), T, b.dt, c->states[nom].data());
    propagate(c->dyn_host, x0_host, c->control[act].data(), T, b.dt, c->states[act].data());
  } else {
    smooth_control(c->control[act], T, C);
    smooth_control(c->control[nom], T, C);
    propagate(c->dyn_host, c->nominal_x.data(), c->control[nom].data(), T, b.dt, c->states[nom].data());
    propagate(c->dyn_host, x0_host, c->control[act].data(), T, b.dt, c->states[act].data());
  }
  for (int d = 0; d < 2; ++d) c->fe[d].increase = c->stats[d].baseline - c->fe[d].previous_baseline;
  return MPPI_OK;
}

#define TUBE_SYS(nominal) ((nominal) ? c->nom() : 1 - c->nom())

int mppi_tube_get_control_seq(const mppi_tube_controller* c, int32_t nominal, float* out)
{
  if (!c || !out || !check_nominal(nominal)) return fail(MPPI_E_INVALID, "invalid argument");
  const std::vector<float>& u = c->control[TUBE_SYS(nominal)];
  std::copy(u.begin(), u.end(), out);
  return MPPI_OK;
}

int mppi_tube_set_control_seq(mppi_tube_controller* c, int32_t nominal, const float* u)
{
  if (!c || !u || !check_nominal(nominal)) return fail(MPPI_E_INVALID, "invalid argument");
  std::vector<float>& v = c->control[TUBE_SYS(nominal)];
  std::copy(u, u + v.size(), v.begin());
  return MPPI_OK;
}

int mppi_tube_get_state_seq(const mppi_tube_controller* c, int32_t nominal, float* out)
{
  if (!c || !out || !check_nominal(nominal)) return fail(MPPI_E_INVALID, "invalid argument");
  const std::vector<float>& s = c->states[TUBE_SYS(nominal)];
  std::copy(s.begin(), s.end(), out);
  return MPPI_OK;
}

int mppi_tube_get_weight_stats(const mppi_tube_controller* c, int32_t nominal, mppi_weight_stats* out)
{
  if (!c || !out || !check_nominal(nominal)) return fail(MPPI_E_INVALID, "invalid argument");
  *out = c->stats[TUBE_SYS(nominal)];
  return MPPI_OK;
}

int mppi_tube_get_free_energy(const mppi_tube_controller* c, int32_t nominal, mppi_free_energy* out)
{
  if (!c || !out || !check_nominal(nominal)) return fail(MPPI_E_INVALID, "invalid argument");
  *out = c->fe[TUBE_SYS(nominal)];
  return MPPI_OK;
}

int mppi_tube_get_nominal_state_used(const mppi_tube_controller* c, int32_t* out)
{
  if (!c || !out) return fail(MPPI_E_INVALID, "null argument");
  *out = c->nominal_state_used;
  return MPPI_OK;
}

int mppi_tube_slide_control_seq(mppi_tube_controller* c, int32_t steps)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (steps < 0) return fail(MPPI_E_INVALID, "steps must be >= 0");
  if (c->rmppi()) return MPPI_OK;  // RobustMPPIController::slideControlSequence does nothing
  const mppi_controller_desc& b = c->b();
  const int T = c->T, C = c->C, S = c->S, nom = c->nom();
  // updateNominalState(nominal_control_.col(0)): one model step (Dynamics::step, no constraint) of the nominal
  // trajectory's first state (:345-351)
  {
    float xn[MPPI_MAX_STATE], xd[MPPI_MAX_STATE], y[MPPI_MAX_OUTPUT];
    MPPI_CHECK(mppi_host_step(&c->dyn_host, c->states[nom].data(), c->control[nom].data(), b.dt, xn, xd, y));
    std::copy(xn, xn + S, c->states[nom].begin());
  }
  const float* zero = b.dynamics.zero_control;
  const float* scale = b.slide_control_scale;
  for (int d = 0; d < 2; ++d) {
    std::vector<float>& u = c->control[d];
    for (int i = 0; i < T; ++i) {
      const int ind = std::min(i + steps, T - 1);
      for (int j = 0; j < C; ++j) {
        float v = u[(size_t)ind * C + j];
        if (i + steps > T - 1) v = (v - zero[j]) * scale[j] + zero[j];
        u[(size_t)i * C + j] = v;
      }
    }
  }
  return MPPI_OK;
}

int mppi_tube_set_nominal_state(mppi_tube_controller* c, const float* x)
{
  if (!c || !x) return fail(MPPI_E_INVALID, "null argument");
  if (c->rmppi()) {
    std::copy(x, x + c->S, c->nominal_x.begin());
  } else {
    std::copy(x, x + c->S, c->states[c->nom()].begin());
    c->nominal_init = true;
  }
  return MPPI_OK;
}

int mppi_tube_get_nominal_state(const mppi_tube_controller* c, float* x)
{
  if (!c || !x) return fail(MPPI_E_INVALID, "null argument");
  const float* src = c->rmppi() ? c->nominal_x.data() : c->states[c->nom()].data();
  std::copy(src, src + c->S, x);
  return MPPI_OK;
}

int mppi_tube_set_feedback_gains(mppi_tube_controller* c, const float* gains)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (gains == nullptr) {
    c->gains.clear();
    return MPPI_OK;
  }
  const size_t n = (size_t)c->T * c->S * c->C;
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(gains[i])) return fail(MPPI_E_INVALID, "feedback gains must be finite");
  c->gains.assign(gains, gains + n);
  c->gains_dirty = true;
  return MPPI_OK;
}

int mppi_tube_get_feedback_control(const mppi_tube_controller* c, const float* x, const float* x_nominal, int32_t t,
                                   float* u_out)
{
  if (!c || !x || !x_nominal || !u_out) return fail(MPPI_E_INVALID, "null argument");
  if (t < 0 || t >= c->T) return fail(MPPI_E_INVALID, "time index out of range");
  const int S = c->S, C = c->C;
  for (int j = 0; j < C; ++j) {
    float s = 0.0f;
    if (!c->gains.empty())
      for (int i = 0; i < S; ++i) s += c->gains[((size_t)t * S + i) * C + j] * (x[i] - x_nominal[i]);
    u_out[j] = s;
  }
  return MPPI_OK;
}

int mppi_tube_set_thresholds(mppi_tube_controller* c, float nominal_threshold, float value_function_threshold)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (!(nominal_threshold == nominal_threshold) || !(value_function_threshold == value_function_threshold))
    return fail(MPPI_E_INVALID, "thresholds must not be NaN");
  c->d.nominal_threshold = nominal_threshold;
  c->d.value_function_threshold = value_function_threshold;
  return MPPI_OK;
}

int mppi_tube_update_importance_sampling_control(mppi_tube_controller* c, const float* x, int32_t stride)
{
  if (!c || !x) return fail(MPPI_E_INVALID, "null argument");
  if (!c->rmppi()) return fail(MPPI_E_INVALID, "updateImportanceSamplingControl: RMPPI controllers only");
  if (stride < 0) return fail(MPPI_E_INVALID, "stride must be >= 0");
  const mppi_controller_desc& b = c->b();
  const int K = c->K, T = c->T, C = c->C, S = c->S, N = c->N, M = c->M, nom = c->nom();
  const size_t TC = (size_t)T * C;
  if (!c->isc_init) {
    std::copy(x, x + S, c->nominal_x.begin());
    c->isc_init = true;
    c->nominal_stride = 0;
  } else {
    // candidates on the line through the nominal trajectory's first two states and the measured state (:351-363)
    const float* xk = c->states[nom].data();
    const float* xkp1 = c->states[nom].data() + (T > 1 ? S : 0);
    for (int i = 0; i < N; ++i) {
      const float w0 = c->lsw[i], w1 = c->lsw[N + i], w2 = c->lsw[2 * N + i];
      for (int s = 0; s < S; ++s) c->cand_h[(size_t)i * S + s] = xk[s] * w0 + xkp1[s] * w1 + x[s] * w2;
      // computeImportanceSamplerStride (:501-511): round((0, s, s) . w), half away from zero
      c->strides_h[i] = (int)std::round(0.0f * w0 + (float)stride * w1 + (float)stride * w2);
    }
    // candidates, strides and the nominal control through the mapped buffers
    float* mean_h = c->host_io + 2 * MPPI_MAX_STATE;
    float* cand_io = c->isc_io;
    int* stride_io = reinterpret_cast<int*>(cand_io + (size_t)N * S);
    float* eval_io = cand_io + (size_t)N * S + N;
    float* cand_iod = c->isc_io_hd;
    std::copy(c->control[nom].begin(), c->control[nom].end(), mean_h);
    std::copy(c->cand_h.begin(), c->cand_h.end(), cand_io);
    std::copy(c->strides_h.begin(), c->strides_h.end(), stride_io);
    const WordSpan in[] = {{cand_iod, c->cand_d, N * S},
                           {cand_iod + (size_t)N * S, c->stride_d, N},
                           {c->host_io_hd + 2 * MPPI_MAX_STATE, c->mean_d, (int)TC}};
    MPPI_CHECK(launch_copy_words(in, 3, c->stream));
    // generateSamples(stride, 0) around the nominal control (:593-597); the evaluation reads samples j < M, so of a
    // Gaussian draw only those rows are written
    const IterRef ref{nullptr, {c->rng_counter, stride, 0}};
    if (b.sampler.kind == MPPI_SAMPLER_COLORED)
      MPPI_CHECK(launch_sample_colored(b.sampler, K, T, C, c->mean_d, c->samples_d, ref, nullptr, c->stream));
    else
      MPPI_CHECK(launch_sample_gaussian(b.sampler, K, T, C, c->mean_d, c->samples_d, ref, nullptr, c->stream,
                                        std::min(M, K)));
    MPPI_CHECK(launch_candidate_eval(b.dynamics, b.cost, b.sampler, N, M, K, T, b.dt, b.lambda, b.alpha, c->cand_d,
                                     c->stride_d, c->mean_d, c->samples_d, c->eval_d, c->stream));
    const WordSpan res[] = {{c->eval_d, cand_iod + (size_t)N * S + N, N * M}};
    const HostSignal sig = tube_signal(c);
    MPPI_CHECK(launch_copy_words(res, 1, c->stream, &sig));
    MPPI_CHECK(tube_wait(c));
    std::copy(eval_io, eval_io + (size_t)N * M, c->eval_h.begin());
    c->rng_counter++;
    // computeCandidateBaseline + computeBestIndex (:513-546), in the reference's arithmetic
    float base = c->eval_h[0];
    for (int i = 1; i < N * M; ++i)
      if (c->eval_h[i] < base) base = c->eval_h[i];
    for (int i = 0; i < N; ++i) {
      float fe = 0.0f;
      for (int j = 0; j < M; ++j)
        fe += expf((float)(-1.0 / (double)b.lambda * (double)(c->eval_h[(size_t)i * M + j] - base)));
      fe = (float)(fe / (1.0 * M));
      fe = -b.lambda * logf(fe) + base;
      c->cand_fe[i] = fe;
      if (fe < c->d.value_function_threshold) c->best_index = i;
    }
    c->nominal_state_used = c->best_index;
    c->nominal_stride = c->strides_h[c->best_index];
    std::copy(c->cand_h.begin() + (size_t)c->best_index * S, c->cand_h.begin() + (size_t)(c->best_index + 1) * S,
              c->nominal_x.begin());
  }
  // slideControlSequenceHelper(nominal_stride, nominal control) and the nominal trajectory (:558-566)
  {
    std::vector<float>& u = c->control[nom];
    const float* zero = b.dynamics.zero_control;
    const float* scale = b.slide_control_scale;
    const int steps = c->nominal_stride;
    for (int i = 0; i < T; ++i) {
      const int ind = std::min(i + steps, T - 1);
      for (int j = 0; j < C; ++j) {
        float v = u[(size_t)ind * C + j];
        if (i + steps > T - 1) v = (v - zero[j]) * scale[j] + zero[j];
        u[(size_t)i * C + j] = v;
      }
    }
  }
  propagate(c->dyn_host, c->nominal_x.data(), c->control[nom].data(), T, b.dt, c->states[nom].data());
  return MPPI_OK;
}

int mppi_tube_get_candidates(const mppi_tube_controller* c, float* states, int32_t* strides, float* costs,
                             float* free_energies, int32_t* best_index)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (!c->rmppi()) return fail(MPPI_E_INVALID, "candidates: RMPPI controllers only");
  if (states) std::copy(c->cand_h.begin(), c->cand_h.end(), states);
  if (strides) std::copy(c->strides_h.begin(), c->strides_h.end(), strides);
  if (costs) std::copy(c->eval_h.begin(), c->eval_h.end(), costs);
  if (free_energies) std::copy(c->cand_fe.begin(), c->cand_fe.end(), free_energies);
  if (best_index) *best_index = c->best_index;
  return MPPI_OK;
}

int mppi_tube_get_samples(const mppi_tube_controller* c, float* samples_host, float* costs_host)
{
  if (!c) return fail(MPPI_E_INVALID, "null argument");
  if (!c->have_samples) return fail(MPPI_E_INVALID, "no computeControl call yet");
  const size_t n = 2 * (size_t)c->K * c->T * c->C;
  if (samples_host)
    MPPI_HIP_CHECK(hipMemcpyAsync(samples_host, c->samples_d, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  if (costs_host)
    MPPI_HIP_CHECK(hipMemcpyAsync(costs_host, c->costs_d, sizeof(float) * 2 * c->K, hipMemcpyDeviceToHost, c->stream));
  MPPI_HIP_CHECK(hipStreamSynchronize(c->stream));
  return MPPI_OK;
}

int mppi_tube_time_kernels(mppi_tube_controller* c, int32_t reps, float* rollout_ms, float* combine_ms,
                           int32_t* num_partials)
{
  if (!c || !rollout_ms || !combine_ms || reps < 1) return fail(MPPI_E_INVALID, "bad argument");
  if (!c->have_samples) return fail(MPPI_E_INVALID, "time_kernels: call computeControl once first");
  const mppi_controller_desc& b = c->b();
  if (b.sampler.kind != MPPI_SAMPLER_GAUSSIAN)
    return fail(MPPI_E_INVALID, "time_kernels: the fused two-system iteration takes the Gaussian sampler");
  const int K = c->K;
  const size_t TC = (size_t)c->T * c->C;
  // the last iteration again: x0_d and mean_d still hold its inputs
  const IterRef ref{nullptr, c->last_args};
  mppi_two_system_params tp{};
  tp.mode = c->d.mode;
  tp.nominal_index = c->nom();
  tp.value_function_threshold = c->d.value_function_threshold;
  tp.feedback_gains = (c->rmppi() && !c->gains.empty()) ? c->gains_d : nullptr;
  std::vector<hipEvent_t> ev((size_t)(2 * reps + 1), nullptr);
  int st = MPPI_OK;
  for (auto& e : ev)
    if (st == MPPI_OK && hipEventCreate(&e) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventCreate");
  if (st == MPPI_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventRecord");
  for (int r = 0; r < reps && st == MPPI_OK; ++r) {
    int G = 0;
    st = launch_two_system_fused(b.dynamics, b.cost, b.sampler, tp, K, c->T, b.dt, b.lambda, b.alpha, c->x0_d,
                                 c->mean_d, c->samples_d, c->costs_d, ref, c->parts_d, c->psums_d, c->gstride, &G,
                                 c->stream);
    if (num_partials != nullptr) *num_partials = G;
    if (st == MPPI_OK && hipEventRecord(ev[2 * r + 1], c->stream) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventRecord");
    if (st == MPPI_OK)
      st = launch_combine(c->parts_d, c->psums_d, G, K, (int)TC, b.lambda, c->stats_d, c->out_d, nullptr, 1, nullptr,
                          c->stream, 2, c->gstride);
    if (st == MPPI_OK && hipEventRecord(ev[2 * r + 2], c->stream) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventRecord");
  }
  float roll = 0.0f, comb = 0.0f;
  if (st == MPPI_OK && hipEventSynchronize(ev.back()) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventSynchronize");
  for (int r = 0; r < reps && st == MPPI_OK; ++r) {
    float x = 0.0f, y = 0.0f;
    if (hipEventElapsedTime(&x, ev[2 * r], ev[2 * r + 1]) != hipSuccess ||
        hipEventElapsedTime(&y, ev[2 * r + 1], ev[2 * r + 2]) != hipSuccess)
      st = fail(MPPI_E_HIP, "hipEventElapsedTime");
    roll += x;
    comb += y;
  }
  (void)hipStreamSynchronize(c->stream);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (st != MPPI_OK) return st;
  *rollout_ms = roll / reps;
  *combine_ms = comb / reps;
  return MPPI_OK;
}

}  // extern "C"
