/*
 * mppi_amd.h -- C-ABI of the MI355X-native MPPI engine.
 *
 * Every entry point is `extern "C"`, takes plain pointers, sizes and POD parameter blocks, and returns an int status
 * (MPPI_OK == 0, negative MPPI_E_* on failure; the message of the last failure on the calling thread is available
 * from mppi_last_error()). Nothing here ever calls exit().
 *
 * Two levels are exported:
 *   1. kernel-level entry points that work on caller-owned device buffers and a caller-supplied hipStream_t
 *      (passed as void* so this header needs no HIP include). They replace the reference's kernel launchers.
 *   2. a controller object (mppi_controller) that owns its device workspace and runs whole MPPI iterations
 *      (sample -> rollout -> weights -> update), optionally replayed from a captured hipGraph. It replaces
 *      VanillaMPPIController::computeControl and the getters around it.
 *
 * Buffer layouts (all FP32, row-major, C order):
 *   samples / control samples   [K][T][C]   (reference: sampling_distributions/sampling_distribution.cu:176-177)
 *   mean / nominal control      [T][C]      (reference: Eigen col-major C x T, i.e. the same bytes)
 *   costs / weights             [K]
 *   x0                          [S]
 *   cost map                    [height][width]
 *
 * Reference citations are into MPPI-Generic 0.9.0 (include/mppi/...).
 */
#ifndef MPPI_AMD_H
#define MPPI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------------------------------
 * Status codes. The reference exits the process on CUDA errors (utils/gpu_err_chk.cuh:32-40) and on bad launch shapes
 * (core/mppi_common.cu:1302-1313); here every failure is a return code.
 * ---------------------------------------------------------------------------------------------------------------- */
#define MPPI_OK 0
#define MPPI_E_INVALID -1     /* bad argument, shape or parameter block */
#define MPPI_E_HIP -2         /* a HIP runtime call failed */
#define MPPI_E_UNSUPPORTED -3 /* model/cost/sampler combination not compiled into the kernel library */
#define MPPI_E_NOMEM -4       /* device allocation failed */
#define MPPI_E_SIZE -5        /* size exceeds a compiled limit (MPPI_MAX_*) */

#define MPPI_MAX_STATE 8
#define MPPI_MAX_CONTROL 4
#define MPPI_MAX_OUTPUT 8
#define MPPI_MAX_TIMESTEPS 256
#define MPPI_MAX_NN_LAYERS 4

/* Dynamics models (reference: dynamics/cartpole, dynamics/double_integrator, dynamics/racer_dubins,
 * dynamics/autorally/ar_nn_model for the residual-MLP pattern). */
#define MPPI_MODEL_CARTPOLE 1
#define MPPI_MODEL_DOUBLE_INTEGRATOR 2
#define MPPI_MODEL_RACER_DUBINS 3
#define MPPI_MODEL_RACER_DUBINS_NN 4

/* Cost functions (reference: cost_functions/cartpole/cartpole_quadratic_cost, cost_functions/quadratic_cost plus the
 * cost-map and speed terms patterned on cost_functions/autorally/ar_standard_cost). */
#define MPPI_COST_CARTPOLE_QUADRATIC 1
#define MPPI_COST_QUADRATIC_MAP 2

/* Sampling distributions (reference: sampling_distributions/gaussian, sampling_distributions/colored_noise). */
#define MPPI_SAMPLER_GAUSSIAN 1
#define MPPI_SAMPLER_COLORED 2

/* Colored-noise FFT length: the reference uses n = 2T and keeps the first T outputs
 * (sampling_distributions/colored_noise/colored_noise.cu:296); MPPI_FFT_LEN_T uses n = T. */
#define MPPI_FFT_LEN_2T 0
#define MPPI_FFT_LEN_T 1

/* Controller kinds: VanillaMPPIController (controllers/MPPI/mppi_controller.cu:152-237) and ColoredMPPIController
 * (controllers/ColoredMPPI/colored_mppi_controller.cu:145-249: optional state leash, normExp or Tsallis weights,
 * only control 1 clamped after smoothing). */
#define MPPI_CONTROLLER_VANILLA 0
#define MPPI_CONTROLLER_COLORED 1

/* Precision of the residual MLP inside the rollout. */
#define MPPI_NN_FP32 0 /* FP32 weights and activations (parity mode) */
#define MPPI_NN_BF16 1 /* BF16 weights/activations on MFMA, FP32 accumulate */

/* ------------------------------------------------------------------------------------------------------------------
 * Parameter blocks. Fill them with mppi_*_defaults() and then edit fields.
 * ---------------------------------------------------------------------------------------------------------------- */
typedef struct mppi_dynamics_params {
  int32_t model;
  int32_t state_dim;   /* set by mppi_dynamics_defaults */
  int32_t control_dim; /* set by mppi_dynamics_defaults */
  int32_t output_dim;  /* set by mppi_dynamics_defaults */
  /* Dynamics::control_rngs_ / control_deadband_ / zero_control_ (dynamics/dynamics.cuh:130,508-509) */
  float control_min[MPPI_MAX_CONTROL];
  float control_max[MPPI_MAX_CONTROL];
  float control_deadband[MPPI_MAX_CONTROL];
  float zero_control[MPPI_MAX_CONTROL];
  /* CartpoleDynamicsParams (dynamics/cartpole/cartpole_dynamics.cuh:29-31,101) */
  float cart_mass;
  float pole_mass;
  float pole_length;
  float gravity;
  /* RacerDubinsParams (dynamics/racer_dubins/racer_dubins.cuh:81-103) */
  float c_t;
  float c_b;
  float c_v;
  float c_0;
  float steering_constant;
  float steer_command_angle_scale;
  float steer_angle_scale;
  float max_steer_angle;
  float max_steer_rate;
  float brake_delay_constant;
  float max_brake_rate_neg;
  float max_brake_rate_pos;
  float wheel_base;
  int32_t gear_sign;
  /* Residual MLP (FNNHelper layout, utils/nn_helpers/fnn_helper.cu:130-205): per layer l the weights W_l are
   * row-major [out][in] followed by the bias [out]. nn_layers = {in, h1, h2, out}; the output is added to the
   * derivatives of states nn_output_states[0..out) (distinct states for MPPI_NN_BF16). Device pointer to
   * nn_num_params floats. */
  int32_t nn_num_layers;
  int32_t nn_layers[MPPI_MAX_NN_LAYERS];
  int32_t nn_output_states[MPPI_MAX_STATE];
  int32_t nn_precision; /* MPPI_NN_FP32 or MPPI_NN_BF16 */
  int32_t nn_num_params;
  const float* nn_theta;
} mppi_dynamics_params;

typedef struct mppi_cost_params {
  int32_t cost;
  int32_t output_dim;
  /* Quadratic state term sum_i coeff[i] * (y[i] - goal[i])^2. For MPPI_COST_CARTPOLE_QUADRATIC the coefficients are
   * (cart_position, cart_velocity, pole_angle, pole_angular_velocity) and goal is desired_terminal_state
   * (cost_functions/cartpole/cartpole_quadratic_cost.cuh:10-23). */
  float coeff[MPPI_MAX_OUTPUT];
  float goal[MPPI_MAX_OUTPUT];
  /* terminal cost = terminal_cost_coeff * quadratic state term (cartpole) ; 0 -> no terminal cost */
  float terminal_cost_coeff;
  /* Cost-map term map_coeff * bilinear(map, y[map_x_index], y[map_y_index]) with the reference's CPU bilinear
   * semantics (utils/texture_helpers/two_d_texture_helper.cu:151-245, texture_helper.cu:94-134).
   * World -> texel: q = ((R (p - origin)) / resolution) / extent * extent - 0.5, clamped, R = map_rotation (row-major
   * 2 x 2, worldPoseToMapPose's rotation restricted to the plane) when map_rotated, else the identity.
   * map == NULL disables it. At most 2^30 texels (MPPI_E_SIZE otherwise). */
  const float* map; /* device pointer [map_height][map_width] */
  int32_t map_width;
  int32_t map_height;
  float map_origin_x;
  float map_origin_y;
  float map_resolution;
  float map_coeff;
  int32_t map_x_index;
  int32_t map_y_index;
  /* Speed term speed_coeff * (y[speed_index] - speed_target)^2 (cost_functions/autorally/ar_standard_cost.cu:284-297) */
  float speed_coeff;
  float speed_target;
  int32_t speed_index;
  /* Goal trajectory (QuadraticCostTrajectoryParams, cost_functions/quadratic_cost/quadratic_cost.cuh:12-63), for
   * MPPI_COST_QUADRATIC_MAP: goal_traj [goal_traj_len][output_dim] (device pointer for the kernels, host pointer for
   * mppi_host_running_cost); the running cost of step t uses row min(goal_time + t, goal_traj_len - 1) in place of
   * goal (getIndex, :48-57). goal stays the terminal term's goal. goal_traj == NULL: goal at every step. */
  int32_t goal_traj_len;
  const float* goal_traj;
  int32_t goal_time;
  int32_t map_rotated;      /* 0: identity map frame rotation */
  float map_rotation[4];    /* row-major 2 x 2 */
} mppi_cost_params;

typedef struct mppi_sampler_params {
  int32_t kind;
  int32_t control_dim;
  /* GaussianParamsImpl (sampling_distributions/gaussian/gaussian.cuh:21-42) */
  float std_dev[MPPI_MAX_CONTROL];
  float control_cost_coeff[MPPI_MAX_CONTROL];
  float pure_noise_trajectories_percentage;
  float std_dev_decay;
  /* ColoredNoiseParamsImpl (sampling_distributions/colored_noise/colored_noise.cuh:44-73) */
  float exponents[MPPI_MAX_CONTROL];
  float offset_decay_rate;
  float fmin;
  int32_t fft_len_mode; /* MPPI_FFT_LEN_2T (reference) or MPPI_FFT_LEN_T */
  int32_t _pad;
  /* Counter-based RNG (Philox4x32-10): normals are a pure function of (seed, counter, element index). */
  uint64_t seed;
  /* Per-step standard deviations (GaussianTimeVaryingStdDevParams with time_specific_std_dev,
   * sampling_distributions/gaussian/gaussian.cuh:66-80, gaussian.cu:38-43,499-502,659-667): device pointer
   * [num_timesteps][control_dim] used in place of std_dev at step t, in the samples (times std_dev_decay^iteration)
   * and in the likelihood-ratio cost. NULL: std_dev at every step. */
  const float* std_dev_t;
} mppi_sampler_params;

/* Per-instance result of the weight transform (device-resident, 32 bytes). */
typedef struct mppi_weight_stats {
  float baseline;   /* min_k cost_k (core/mppi_common.cu:938-992) */
  int32_t argmin;   /* first strict minimum (core/mppi_common.cu:921-936) */
  double normalizer; /* sum_k exp(-(cost_k - baseline)/lambda), accumulated in FP64 */
  double sum_w2;     /* sum_k w_k^2, FP64 (free-energy variance) */
  float lambda;
  int32_t num_rollouts;
  double weighted_cost; /* sum_k w_k cost_k / normalizer over finite costs, FP64; with a learned proposal the costs
                           averaged are those without the importance-weight correction */
} mppi_weight_stats;

/* Free-energy statistics (core/mppi_common.cu:1101-1117, controllers/controller.cuh:22-30). */
typedef struct mppi_free_energy {
  float free_energy_mean;
  float free_energy_variance;
  float free_energy_modified_variance;
  float normalizer_percent;
  float increase;
  float previous_baseline;
} mppi_free_energy;

const char* mppi_last_error(void);
const char* mppi_status_string(int status);
/* Version of this library, e.g. "0.1.0". */
const char* mppi_version(void);

int mppi_dynamics_defaults(int32_t model, mppi_dynamics_params* out);
int mppi_cost_defaults(int32_t cost, const mppi_dynamics_params* dyn, mppi_cost_params* out);
int mppi_sampler_defaults(int32_t kind, int32_t control_dim, mppi_sampler_params* out);

/* Number of floats of an FNN parameter vector for the given layer sizes ((in+1)*out summed over layers). */
int mppi_nn_num_params(int32_t num_layers, const int32_t* layers, int32_t* out_num_params);

/* First sample index that is drawn as zero-mean "pure noise": the smallest k with (float)k >= (1 - p) * K evaluated in
 * FP32, as in sampling_distributions/gaussian/gaussian.cu:96 and :658. */
int mppi_pure_noise_start(int32_t num_rollouts, float pure_noise_percentage, int32_t* out_index);

/* Host-side plugin methods (the same model code the kernels run), with host pointers:
 * Dynamics::step (dynamics/dynamics.cuh:278-285: computeStateDeriv -> updateState -> stateToOutput) and
 * Dynamics::enforceConstraints (dynamics/dynamics.cuh:247-261). Any of x_next, x_dot, y may be NULL. For
 * MPPI_MODEL_RACER_DUBINS_NN, p->nn_theta must point to host memory here. */
int mppi_host_step(const mppi_dynamics_params* p, const float* x, const float* u, float dt, float* x_next,
                   float* x_dot, float* y);
int mppi_host_enforce_constraints(const mppi_dynamics_params* p, float* u);
/* Weight and map I/O from .npz archives (numpy.savez / savez_compressed; stored or deflate members, zip64) without a
 * third-party loader.
 * mppi_npz_read: array `key` (name without ".npy"; float32/64, int32/64, converted to FP32 in C order); out == NULL
 * returns the shape (ndim <= 4, unused entries 0) and element count only. */
int mppi_npz_read(const char* path, const char* key, float* out, int64_t max_count, int32_t* ndim, int64_t* shape,
                  int64_t* count);
/* FNNHelper::loadParams(prefix, npz) (utils/nn_helpers/fnn_helper.cu:44-127): layer sizes from dynamics_W1 /
 * dynamics_b1 and dynamics_b2, ... (prefix "" or e.g. "model", a '/' appended as add_slash does), and the parameter
 * vector in the FNN layout of mppi_dynamics_params (per layer W [out][in] row-major, then b [out]). theta == NULL
 * returns the layer sizes and the parameter count only. */
int mppi_fnn_load_npz(const char* path, const char* prefix, int32_t max_layers, int32_t* num_layers, int32_t* layers,
                      float* theta, int32_t max_params, int32_t* num_params);
/* ARStandardCost::loadTrackData's map (cost_functions/autorally/ar_standard_cost.cu:85-142): xBounds, yBounds,
 * pixelsPerMeter give width = (x_max - x_min) ppm, height = (y_max - y_min) ppm, origin (x_min, y_min), resolution
 * 1 / ppm; `channel` (NULL: "channel0") gives the [height][width] values in their stored order (the texture upload's
 * rows). values == NULL returns the geometry only. */
int mppi_track_map_load_npz(const char* path, const char* channel, float* values, int64_t max_count, int32_t* width,
                            int32_t* height, float* origin_x, float* origin_y, float* resolution);

/* Dynamics::enforceLeash: per state, the nominal value when it lies within leash_values[i] of the true one, else the
 * true value moved leash_values[i] towards it (dynamics/dynamics.cuh:445-463); the bicycle leashes its position in
 * the body frame of the true yaw and its yaw along the shortest angle (dynamics/racer_dubins/racer_dubins.cu:
 * 176-238). Host pointers; out may alias neither input. */
int mppi_host_enforce_leash(const mppi_dynamics_params* p, const float* state_true, const float* state_nominal,
                            const float* leash_values, float* state_out);
/* Cost::computeRunningCost on the host (the kernels' cost code; c->map and c->goal_traj must point to host memory
 * here), at step 0 and at step t (the step index picks the goal_traj row). */
int mppi_host_running_cost(const mppi_cost_params* c, const float* y, float* out);
int mppi_host_running_cost_at(const mppi_cost_params* c, const float* y, int32_t t, float* out);

/* ------------------------------------------------------------------------------------------------------------------
 * Kernel-level entry points (device pointers, caller stream).
 * ---------------------------------------------------------------------------------------------------------------- */

/* Standard normals eps[i] = N(0,1) for i in [0, n): Philox4x32-10 keyed by `seed`, counter (i/4, offset), Box-Muller.
 * Replaces curandGenerateNormal (sampling_distributions/gaussian/gaussian.cu:380-395). */
int mppi_generate_normals(uint64_t seed, uint64_t offset, float* out_d, int64_t n, void* stream);

/* Gaussian control samples u[k][t][c] (sampling_distributions/gaussian/gaussian.cu:17-277,375-431):
 *   u = mean[t][c]                          if k == 0 or t < optimization_stride
 *   u = s * eps                             if k >= pure_noise_start
 *   u = mean[t][c] + s * eps                otherwise,     s = std_dev[c] * std_dev_decay^iteration_num
 * with eps drawn as by mppi_generate_normals(seed, rng_offset, ...) over the K*T*C buffer.
 * If eps_in_d is non-NULL it is used instead of the generator (a fed noise buffer). */
int mppi_sample_gaussian(const mppi_sampler_params* sampler, int32_t num_rollouts, int32_t num_timesteps,
                         const float* mean_d, float* samples_d, int32_t optimization_stride, int32_t iteration_num,
                         uint64_t rng_offset, const float* eps_in_d, void* stream);

/* Colored (power-law, 1/f^beta per control dimension) samples
 * (sampling_distributions/colored_noise/colored_noise.cu:286-392): complex Gaussian spectra scaled by f^(-beta/2),
 * inverse real DFT of length n (2T or T), first T outputs, offset-decay subtraction x[t] - x[stride]*decay^t,
 * division by sigma_c * n, then the Gaussian transform above. If spectra_in_d is non-NULL it holds the unscaled
 * complex spectra [K][C][n/2+1][2] used instead of the generator (values of magnitude below 255: they are staged in
 * FP16 parts). The transform's basis is built on the first call for a given (n, T) on a device and kept: that first
 * call synchronises the stream and must not be made inside a graph capture (later calls may be captured). */
int mppi_sample_colored(const mppi_sampler_params* sampler, int32_t num_rollouts, int32_t num_timesteps,
                        const float* mean_d, float* samples_d, int32_t optimization_stride, int32_t iteration_num,
                        uint64_t rng_offset, const float* spectra_in_d, void* stream);

/* Fused rollout (core/mppi_common.cu:35-153): per sample, for t < T: read u, enforce constraints, write the clamped u
 * back in place, step the dynamics, add running cost + likelihood-ratio cost; add the terminal cost; divide by T.
 * mean_d is the [T][C] mean the samples were drawn around (used by the likelihood-ratio control cost). */
int mppi_rollout(const mppi_dynamics_params* dyn, const mppi_cost_params* cost, const mppi_sampler_params* sampler,
                 int32_t num_rollouts, int32_t num_timesteps, float dt, float lambda, float alpha, const float* x0_d,
                 const float* mean_d, float* samples_d, float* costs_d, void* stream);

/* Weight transform (core/mppi_common.cu:1067-1089,1483-1506): baseline = min cost (first strict minimum), costs_d is
 * overwritten in place with w_k = exp(-(cost_k - baseline)/lambda), the FP64 normalizer and sum of squares go to
 * stats_d. */
int mppi_compute_weights(float* costs_d, int32_t num_rollouts, float lambda, mppi_weight_stats* stats_d, void* stream);
/* The same with the Tsallis transform (TsallisKernel, core/mppi_common.cu:1004-1021,1469-1481):
 * w_k = (1 - (cost_k - baseline) / gamma)^(1 / (r - 1)) where cost_k - baseline < gamma, else 0; gamma > 0, r != 1.
 * lambda only enters the statistics (free energy). */
int mppi_compute_weights_tsallis(float* costs_d, int32_t num_rollouts, float lambda, float gamma, float r,
                                 mppi_weight_stats* stats_d, void* stream);

/* Weighted control update (core/mppi_common.cu:746-773,1151-1196): mean_out[t][c] = sum_k (w_k / normalizer) *
 * samples[k][t][c], with the normalizer read from stats_d on the device. */
int mppi_weighted_update(const float* weights_d, const mppi_weight_stats* stats_d, const float* samples_d,
                         int32_t num_rollouts, int32_t num_timesteps, int32_t control_dim, float* mean_out_d,
                         void* stream);

/* mppi_weighted_update that allocates nothing. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses that forbid hipMalloc). Its partial rows
 * (ceil(K / 64) * T * C floats) go to a file-scope device array of MPPI_NOALLOC_SCRATCH_BYTES on each GPU, held until
 * the process ends, instead of the workspace mppi_weighted_update grows with hipMalloc; larger sizes return
 * MPPI_E_SIZE. All calls on one GPU share that array: calls that overlap on one GPU (different streams) race. The
 * result equals mppi_weighted_update's bit for bit (the same kernels). */
#define MPPI_NOALLOC_SCRATCH_BYTES (2u << 20)
int mppi_weighted_update_noalloc(const float* weights_d, const mppi_weight_stats* stats_d, const float* samples_d,
                                 int32_t num_rollouts, int32_t num_timesteps, int32_t control_dim, float* mean_out_d,
                                 void* stream);

/* ------------------------------------------------------------------------------------------------------------------
 * Controller object (controllers/MPPI/mppi_controller.cu:152-237, controllers/controller.cuh:70-1058).
 * ---------------------------------------------------------------------------------------------------------------- */
typedef struct mppi_controller mppi_controller;

typedef struct mppi_controller_desc {
  mppi_dynamics_params dynamics;
  mppi_cost_params cost;
  mppi_sampler_params sampler;
  int32_t num_rollouts;
  int32_t num_timesteps;
  float dt;
  float lambda;
  float alpha;
  int32_t num_iters;
  int32_t use_graph;    /* 1: capture one computeControl call into a hipGraph on its second use and replay it;
                           0 (faster on ROCm 7 for host-inclusive calls): direct launches, see
                           mppi_controller_compute_control */
  int32_t keep_samples; /* 1: store the clamped control samples [K][T][C] in the sample buffer every iteration */
  float slide_control_scale[MPPI_MAX_CONTROL]; /* ControllerParams::slide_control_scale_ */
  const float* init_control; /* host [T][C] initial nominal sequence or NULL for zeros */
  const float* cost_map_host; /* optional host [map_height][map_width]; copied into controller-owned memory and
                                 used in place of cost.map */
  const float* nn_theta_host; /* optional host FNN parameters, copied like cost_map_host */
  const float* std_dev_t_host; /* optional host [T][C] per-step standard deviations, copied and used in place of
                                  sampler.std_dev_t */
  const float* goal_traj_host; /* optional host [cost.goal_traj_len][O] goal trajectory, copied and used in place of
                                  cost.goal_traj */
  /* MPPI_CONTROLLER_VANILLA or MPPI_CONTROLLER_COLORED; the fields below apply to the colored controller
   * (ColoredMPPIParams, controllers/ColoredMPPI/colored_mppi_controller.cuh:16-37) */
  int32_t controller_kind;
  float tsallis_gamma; /* gamma: > 0 together with tsallis_r != 0 selects the Tsallis weights; 0: normExp */
  float tsallis_r;     /* r (2 in the reference's defaults) */
  int32_t leash_active; /* leash_active_: leash x0 to the previous solution's state at the last slide's step count */
  float state_leash_dist[MPPI_MAX_STATE]; /* state_leash_dist_ */
} mppi_controller_desc;

int mppi_controller_create(const mppi_controller_desc* desc, mppi_controller** out);
int mppi_controller_destroy(mppi_controller* ctrl);

/* One MPC solve: x0 to the device, num_iters x (sample, rollout, weights, update), the new mean to the host,
 * Savitzky-Golay smoothing keeping u[0] (controllers/controller.cuh:573-607), the state trajectory (:664-684,
 * propagated on the host when first requested) and a control clamp (controllers/MPPI/mppi_controller.cu:224-227).
 * The inputs and results travel through mapped host memory and the host polls a completion word written by the last
 * kernel. Without use_graph, for the built-in models and costs with normExp weights, the calls are direct launches:
 * the iteration arguments are kernel arguments, the rollout reads x0 from host memory, the smoothing and the clamp
 * run in the last kernel (bit-identical to the host functions), and the smoothed sequence stays on the device as the
 * next call's nominal unless the host changes it (set_control_seq, slide_control_seq, set_params). */
int mppi_controller_compute_control(mppi_controller* ctrl, const float* x0_host, int32_t optimization_stride);
/* Enqueue `num_calls` device-resident repetitions of the GPU work of one computeControl call (num_iters x sample,
 * rollout, weights, update) on the controller's stream, replayed from captured hipGraphs of 32 and 8 calls
 * (single-call graphs for the remainder; the same kernels in the same order). x0 is the one the last computeControl
 * uploaded; each call warm-starts from the previous call's raw mean (no smoothing, no slide); stride is 1; iteration
 * it of a call samples with std_dev * std_dev_decay^it; call n, iteration it draws with generator offset
 * base + n num_iters + it, advanced on the device (base: a counter range computeControl does not reach).
 * Returns MPPI_E_INVALID until computeControl has run once. The first call after creation or set_seed uploads the
 * loop's arguments, synchronises the stream and captures the graphs (do not call it inside a stream capture); later
 * calls only enqueue graph launches (no host transfer, no synchronisation). For tight loops and throughput timing. */
int mppi_controller_step_device(mppi_controller* ctrl, int32_t num_calls);
/* getControlSeq(): [T][C] smoothed control sequence. */
int mppi_controller_get_control_seq(const mppi_controller* ctrl, float* out_host);
/* Raw (pre-smoothing) weighted mean of the last iteration [T][C]. */
int mppi_controller_get_raw_mean(const mppi_controller* ctrl, float* out_host);
/* getTargetStateSeq(): [T][S] state trajectory of the last call's smoothed controls from its x0 (computeStateTrajectory,
 * controllers/MPPI/mppi_controller.cu:222), propagated on the host on the first request after each call. */
int mppi_controller_get_state_seq(const mppi_controller* ctrl, float* out_host);
int mppi_controller_get_weight_stats(const mppi_controller* ctrl, mppi_weight_stats* out_host);
int mppi_controller_get_free_energy(const mppi_controller* ctrl, mppi_free_energy* out_host);
/* updateImportanceSampler / setControlSeq: set the nominal [T][C] sequence. */
int mppi_controller_set_control_seq(mppi_controller* ctrl, const float* u_host);
/* Staged cost-map updates (TwoDTextureHelper::updateTexture / updateOrigin / updateResolution / updateRotation and
 * copyToDevice, utils/texture_helpers/two_d_texture_helper.cu:21-60,88-101, texture_helper.cu:147-222,279-338):
 * values [height][width] (or column-major [width][height] when column_major) and the map frame are copied into the
 * controller's staging area at once (any thread) and applied at the start of its next computeControl (or, for the
 * device-resident loop, its next mppi_controller_step_device). New values of the current extent are uploaded in place
 * on the controller's stream, so captured graphs stay valid; a new extent or frame re-captures them. rotation: row-major
 * 2 x 2 or NULL for the identity. Not available for plugin controllers. */
int mppi_controller_update_cost_map(mppi_controller* ctrl, const float* values, int32_t width, int32_t height,
                                    int32_t column_major);
int mppi_controller_update_cost_map_frame(mppi_controller* ctrl, float origin_x, float origin_y, float resolution,
                                          const float* rotation);
/* slideControlSequence(steps) (controllers/controller.cuh:609-621); for the colored controller it also sets the leash
 * index to steps (ColoredMPPIController::slideControlSequence, colored_mppi_controller.cu:260-268). */
int mppi_controller_slide_control_seq(mppi_controller* ctrl, int32_t steps);
/* Device pointers to the controller's buffers for inspection: samples [K][T][C] (clamped samples of the last
 * iteration when keep_samples is set; otherwise unspecified: colored samples are drawn inside the rollout kernel
 * where its workgroup shape allows, else written there pre-clamp by the colored sampler), costs [K] of
 * the last iteration, and the [T][C] mean (after computeControl: the updated mean, which the next iteration samples
 * around, as the reference's control_means_d_). */
int mppi_controller_device_buffers(const mppi_controller* ctrl, float** samples_d, float** costs_d, float** mean_d);
/* Parameter hot-swap (the reference's setParams / paramsToDevice on the plugins and the controller,
 * controllers/controller.cuh:842-876, core/base_plant.hpp:420-456): replaces the dynamics, cost and sampler parameter
 * blocks, dt, lambda, alpha, slide_control_scale and keep_samples of a controller in place. cost_map_host (optional)
 * uploads a new cost map into the controller's copy (reallocated when its size changes); nn_theta_host (optional)
 * new network parameters of the same shape. The number of samples, horizon, iteration count, model, cost and sampler
 * kind must stay the same (MPPI_E_INVALID otherwise: create a new controller). Synchronises the controller's stream
 * and drops its captured graphs (the next calls capture them again); a new seed restarts the generator counter. */
int mppi_controller_set_params(mppi_controller* ctrl, const mppi_controller_desc* desc);
/* Sampled-trajectory export (copySampledControlFromDevice / copyTopControlFromDevice and
 * calculateSampledStateTrajectories, controllers/controller.cu:57-181,287-354, controllers/MPPI/mppi_controller.cuh:
 * 54-59, with the visualisation rollouts of core/mppi_common.cu:371-691). Enabling it (num_random + num_top > 0) makes
 * every computeControl keep the constrained samples of its last iteration (stored from the rollout kernel's LDS tile)
 * and that iteration's sampling mean; 0 and 0 turn it off. 1 + num_random + num_top <= min(K, 4096). Drops the
 * captured graphs. Not available with a learned proposal or a dynamics plugin. */
#define MPPI_MAX_SAMPLED_TRAJECTORIES 4096
int mppi_controller_set_sampled_trajectories(mppi_controller* ctrl, int32_t num_random, int32_t num_top);
/* After a computeControl: count = 1 + num_random + num_top trajectories, re-rolled on the GPU by one gather kernel
 * (one lane per trajectory) from the call's x0: slot 0 is sample 0 (the nominal sequence), then num_random samples
 * drawn without replacement from samples 1 .. 0.98 K - 1 (a generator seeded from the sampler seed and the number of
 * exports so far), then the num_top samples of largest weight (lowest cost; lower index first on ties) by decreasing
 * weight. Per trajectory: indices (sample index), controls [T][C] (as rolled out, constrained), outputs [T][O] (the
 * post-step output y_{t+1} at index t), step_costs [T] ((running cost of y_{t+1} + likelihood-ratio cost of u_t) / T;
 * every step, where the reference's visualisation drops step 0), terminal_costs (terminal cost of y_T / T; the step
 * costs and the terminal cost add up to the sample's rollout cost) and weights (w_k / normalizer of the last
 * iteration). Any output pointer may be NULL; arrays are sized for count trajectories. Synchronises. */
int mppi_controller_get_sampled_trajectories(mppi_controller* ctrl, int32_t* count, int32_t* indices, float* controls,
                                             float* outputs, float* step_costs, float* terminal_costs,
                                             float* weights);
/* Reseed the generator and reset its counter (controllers/controller.cu:203-209). */
int mppi_controller_set_seed(mppi_controller* ctrl, uint64_t seed);
/* The stream all controller work runs on (hipStream_t as void*). */
int mppi_controller_stream(const mppi_controller* ctrl, void** stream);


/* ------------------------------------------------------------------------------------------------------------------
 * User-written device plugins. The reference takes any CRTP Dynamics / Cost class with __device__ methods
 * (dynamics/dynamics.cuh:366-432: computeStateDeriv, updateState, stateToOutput, enforceConstraints;
 * cost_functions/cost.cu:39-53: computeRunningCost, terminalCost) and instantiates its rollout kernel for them. Device
 * code cannot cross a C ABI, so a user's hipcc translation unit instantiates the engine's kernel templates for its
 * own classes (include/mppi_amd/device_plugin.hpp) and hands the controller a table of callbacks: the controller
 * keeps its sampling inputs, graphs, combine kernel, smoothing and getters, and calls the plugin to enqueue the
 * rollout of each iteration.
 * ---------------------------------------------------------------------------------------------------------------- */
typedef struct mppi_plugin_iteration {
  const float* x0_d;        /* [S] initial state */
  const float* mean_d;      /* [T][C] mean the samples are drawn around */
  float* samples_d;         /* [K][T][C]: input when generate == 0, output (clamped samples) when write_samples */
  float* costs_d;           /* [K] costs out (may be NULL) */
  const void* iter_d;       /* device iteration arguments {uint64_t rng_offset; int32_t stride; int32_t iteration} */
  void* partials_d;         /* room for max_partials per-workgroup partials (32 bytes each) */
  float* sums_d;            /* room for max_partials rows of T*C weighted control sums */
  int32_t max_partials;
  int32_t num_rollouts;
  int32_t num_timesteps;
  int32_t generate;         /* 1: draw the Gaussian samples in the kernel; 0: read them from samples_d */
  int32_t write_samples;    /* 1: store the clamped samples to samples_d */
  float dt;
  float lambda;
  float alpha;
  mppi_sampler_params sampler;
  void* stream;             /* hipStream_t */
} mppi_plugin_iteration;

typedef struct mppi_plugin {
  int32_t state_dim;
  int32_t control_dim;
  int32_t output_dim;
  int32_t _pad;
  void* user; /* passed back to every callback; must outlive the controller */
  /* Enqueue one iteration's rollout kernel on it->stream (no synchronisation; it may run inside a graph capture,
   * except on the first call for a device); *num_partials receives the number of partials written (<= max_partials). */
  int (*enqueue_iteration)(void* user, const mppi_plugin_iteration* it, int32_t* num_partials);
  /* Dynamics::step on the host for the state trajectory: x_next from x and a constrained control u. */
  int (*host_step)(void* user, const float* x, const float* u, float dt, float* x_next);
  /* Dynamics::enforceConstraints on the host, in place on one control vector. */
  int (*host_enforce_constraints)(void* user, float* u);
} mppi_plugin;

/* A controller whose rollouts run the plugin's kernels. desc->dynamics supplies only control_dim, the control limits
 * and zero_control (slide); desc->cost is unused; desc->sampler must be Gaussian (drawn in the plugin's kernel) or
 * colored (drawn by the library's sampler into the sample buffer first). Learned proposals are not supported. */
int mppi_controller_create_plugin(const mppi_controller_desc* desc, const mppi_plugin* plugin, mppi_controller** out);

/* ------------------------------------------------------------------------------------------------------------------
 * Batch of independent controllers advanced by one persistent launch per iteration.
 * No single counterpart in the reference: it replaces N VanillaMPPIController objects, each with its own stream
 * and its own sampling/rollout/weights/update launches per computeControl (controllers/MPPI/mppi_controller.cu:
 * 152-237), by one work-queue kernel over all instances' 64-sample tiles plus one combine launch, captured in one
 * hipGraph. Instances may differ in model (double integrator, bicycle, cart-pole), sample count and horizon; each
 * uses Gaussian sampling with its own seed, its own x0, goal and mean, and gets its own FP64 normalizer.
 * ---------------------------------------------------------------------------------------------------------------- */
typedef struct mppi_batch mppi_batch;

typedef struct mppi_instance_desc {
  mppi_dynamics_params dynamics; /* MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_MODEL_RACER_DUBINS or MPPI_MODEL_CARTPOLE */
  mppi_cost_params cost;         /* QUADRATIC_MAP (DI, bicycle) or CARTPOLE_QUADRATIC (cart-pole); cost.map NULL:
                                    the batch's shared map */
  mppi_sampler_params sampler;   /* MPPI_SAMPLER_GAUSSIAN */
  int32_t num_rollouts;          /* multiple of 64, at most 65536 */
  int32_t num_timesteps;
  float dt;
  float lambda;
  float alpha;
  float slide_control_scale[MPPI_MAX_CONTROL]; /* ControllerParams::slide_control_scale_ */
  const float* init_control;     /* host [T][C] or NULL for zeros */
  const float* std_dev_t_host;   /* optional host [T][C] per-step standard deviations (copied; replaces
                                    sampler.std_dev_t) */
  const float* goal_traj_host;   /* optional host [cost.goal_traj_len][O] goal trajectory (copied; replaces
                                    cost.goal_traj) */
} mppi_instance_desc;

/* Launch shape of the persistent kernel: threads per workgroup (256, 512, 1024: one state-chain wavefront, the rest
 * evaluate costs and draw the next chunk's samples), samples per lane of the state-chain wavefront (1, 2 or 4: 64,
 * 128 or 256 samples per tile), horizon chunk (2 to 16 steps; the autotuner tries 3, 4, 6, 8, 12 and 16, and a chunk
 * that the cost wavefronts divide evenly balances them), and the resulting workgroups per CU. iteration_us is the
 * measured time of one batched iteration (0 if not measured). */
typedef struct mppi_batch_config {
  int32_t threads;
  int32_t samples_per_lane;
  int32_t chunk;
  int32_t blocks_per_cu;
  float iteration_us;
} mppi_batch_config;

/* cost_map_host: optional shared map [map_height][map_width] (copied); instances whose cost.map is NULL use it. */
int mppi_batch_create(const mppi_instance_desc* instances, int32_t num_instances, const float* cost_map_host,
                      int32_t map_height, int32_t map_width, mppi_batch** out);
int mppi_batch_destroy(mppi_batch* batch);
/* Times every valid launch shape on the batch (reps iterations each, from the current means, which are restored)
 * and keeps the fastest; `best` (may be NULL) receives it. */
int mppi_batch_autotune(mppi_batch* batch, int32_t reps, mppi_batch_config* best);
int mppi_batch_set_config(mppi_batch* batch, const mppi_batch_config* config);
int mppi_batch_get_config(const mppi_batch* batch, mppi_batch_config* config);
/* One computeControl for every instance: x0_host is [num_instances][MPPI_MAX_STATE]. The inputs are read by the GPU
 * from mapped host memory, the new means are smoothed and constrained on the GPU (the next call's nominal stays in
 * device memory unless the host changes it) and the results written back to mapped host memory; the host polls a
 * completion word written by the last kernel. */
int mppi_batch_compute_control(mppi_batch* batch, const float* x0_host, int32_t optimization_stride);
/* 1: computeControl replays its launches from a captured hipGraph; 0 (default): direct launches, which measured
 * faster on ROCm 7 (a hipGraphLaunch costs more host time than the three launches it replaces). */
int mppi_batch_set_graph_replay(mppi_batch* batch, int32_t enable);
/* num_calls device-resident batched iterations from a captured hipGraph (x0 and means already in device memory,
 * generator counters advanced on the device). */
int mppi_batch_step_device(mppi_batch* batch, int32_t num_calls);
int mppi_batch_get_control_seq(const mppi_batch* batch, int32_t instance, float* out_host);
int mppi_batch_get_raw_mean(const mppi_batch* batch, int32_t instance, float* out_host);
int mppi_batch_get_state_seq(const mppi_batch* batch, int32_t instance, float* out_host);
int mppi_batch_get_weight_stats(const mppi_batch* batch, int32_t instance, mppi_weight_stats* out_host);
int mppi_batch_get_free_energy(const mppi_batch* batch, int32_t instance, mppi_free_energy* out_host);
int mppi_batch_set_control_seq(mppi_batch* batch, int32_t instance, const float* u_host);
int mppi_batch_slide_control_seq(mppi_batch* batch, int32_t steps);
/* Clamped control samples [K][T][C] and costs [K] of an instance's last iteration (copied to the host). */
int mppi_batch_get_samples(const mppi_batch* batch, int32_t instance, float* samples_host, float* costs_host);
int mppi_batch_stream(const mppi_batch* batch, void** stream);
/* Benchmark utility: `reps` device-loop iterations launched eagerly with HIP events around the persistent rollout
 * launch and the combine launch on the batch's stream; average durations of each. Needs a computeControl call first
 * (as mppi_batch_step_device, which also returns MPPI_E_INVALID before it). */
int mppi_batch_time_kernels(mppi_batch* batch, int32_t reps, float* rollout_ms, float* combine_ms);

/* ------------------------------------------------------------------------------------------------------------------
 * Learned proposal network (no counterpart in the reference). Inputs: the state (S values) and the previous control
 * sequence sampled at the 16 knots t_j = j (T - 1) / 15 (16 C values), each normalized as (v - shift) * scale;
 * two hidden layers of 128 GELU units; 2 T C outputs: the mean shift in units of the base sampling std (first T C,
 * [t][c]) and log(sigma_q / sigma_base) (last T C, clamped to [log 0.5, 1] when used). Parameters in the FNN layout
 * (W[out][in] row-major then bias, per layer). Forward, backward and Adam run on BF16 matrix cores with FP32
 * accumulation. Loss per sample: mean over j of (r_j - y_j)^2 exp(-2 s_j) / 2 + s_j with the target residual
 * r = (u* - u_prev) / sigma_base, averaged over the batch.
 * ---------------------------------------------------------------------------------------------------------------- */
typedef struct mppi_proposal mppi_proposal;

int mppi_proposal_num_params(int32_t state_dim, int32_t control_dim, int32_t num_timesteps, int32_t* out);
/* params_host: num_params floats, or NULL for all zeros (the proposal then equals the base distribution). */
int mppi_proposal_create(int32_t state_dim, int32_t control_dim, int32_t num_timesteps, const float* params_host,
                         mppi_proposal** out);
int mppi_proposal_destroy(mppi_proposal* proposal);
int mppi_proposal_get_params(const mppi_proposal* proposal, float* params_host);
/* Sets the parameters and resets the optimizer state. */
int mppi_proposal_set_params(mppi_proposal* proposal, const float* params_host);
int mppi_proposal_set_input_normalization(mppi_proposal* proposal, const float* shift_host, const float* scale_host);
/* outputs_d[b][2 T C] for already normalized inputs_d[b][S + 16 C] (device pointers). */
int mppi_proposal_forward(mppi_proposal* proposal, const float* inputs_d, int32_t batch, float* outputs_d,
                          void* stream);
/* Proposal of one state as the controller uses it: the input from x0_d[S] and the nominal sequence nominal_d[T][C]
 * (normalized with the stored shift/scale), the network, then dmu_d[T][C] = sigma_base * mean shift and
 * sig_d[T][C] = sigma_base * exp(clamp(log-std, log 0.5, 1)); sigma_base_d[C]. fused = 1: one kernel, a workgroup
 * per 16 outputs (the controller's path); 0: the input kernel, the batched forward launches and a head kernel. */
int mppi_proposal_infer(mppi_proposal* proposal, const float* x0_d, const float* nominal_d,
                        const float* sigma_base_d, float* dmu_d, float* sig_d, int32_t fused, void* stream);
/* Loss (optionally copied to loss_host, which synchronizes) and its gradient grad_d[num_params] for targets_d
 * [b][T C]. */
int mppi_proposal_loss_grad(mppi_proposal* proposal, const float* inputs_d, const float* targets_d, int32_t batch,
                            float* grad_d, double* loss_host, void* stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the batch; loss_host as above or NULL (no synchronization). */
int mppi_proposal_train_step(mppi_proposal* proposal, const float* inputs_d, const float* targets_d, int32_t batch,
                             float learning_rate, double* loss_host, void* stream);

/* Learned-proposal mode of a controller (NULL proposal: plain MPPI again). Each computeControl then runs the network
 * on (x0, nominal sequence) and one iteration of num_rollouts samples u = mu + dmu + sigma_q eps (no pure-noise
 * samples), whose costs carry the mean-shift importance correction lambda sum ((u - mu)^2 - (u - mu - dmu)^2) /
 * (2 std_dev^2) (log-density ratio of the shifted to the base Gaussian at the base std; the proposal's spread sets
 * the exploration width without a density ratio). Sample 0 is the warm start mu itself: its (uncorrected) cost is evaluated but it takes no
 * part in the weights. With fallback_tolerance >= 0, a proposal iteration whose weighted mean cost (weighted_cost:
 * the uncorrected costs under the corrected weights) exceeds the warm start's cost c0 by more than
 * fallback_tolerance * |c0| is rejected and the plain iterations (num_rollouts of the controller, num_iters) run
 * instead; a negative tolerance disables the fallback. The proposal must stay alive while attached. */
int mppi_controller_set_proposal(mppi_controller* ctrl, mppi_proposal* proposal, int32_t num_rollouts,
                                 float fallback_tolerance);
int mppi_controller_get_proposal_stats(const mppi_controller* ctrl, int32_t* proposal_calls, int32_t* fallbacks,
                                       mppi_weight_stats* last_proposal_stats, float* last_warm_start_cost);
/* Stream time of the last computeControl from its first upload to its last download (HIP events), i.e. the
 * control latency without the host post-processing (smoothing, state propagation). */
int mppi_controller_get_last_call_ms(const mppi_controller* ctrl, float* ms);
/* Benchmark utility (no counterpart in the reference): `reps` iterations of the device loop (as
 * mppi_controller_step_device, iteration 0 of a call) launched eagerly with HIP events on the controller's stream
 * before the rollout kernel, between it and the combine kernel, and after the combine; the average durations of the
 * two launches are returned, and the number of per-workgroup partials the combine reduces (num_partials may be
 * NULL). Advances the device loop's state like step_device. Needs a computeControl call first. */
int mppi_controller_time_kernels(mppi_controller* ctrl, int32_t reps, float* rollout_ms, float* combine_ms,
                                 int32_t* num_partials);

/* ------------------------------------------------------------------------------------------------------------------
 * Two-system rollouts and the Tube-MPPI / RMPPI iterations: a pair of systems (distributions 0 and 1, each with its
 * own initial state and mean) rolled out together over the same sample index, as the reference's rollout launches
 * with blockDim.z = 2 (controllers/Tube-MPPI/tube_mppi_controller.cu:158-236, core/rmppi_kernels.cu:359-798).
 * ---------------------------------------------------------------------------------------------------------------- */
#define MPPI_TWO_SYSTEM_TUBE 0  /* two independent rollouts (rolloutKernel with z = 2, core/mppi_common.cu:35-153) */
#define MPPI_TWO_SYSTEM_RMPPI 1 /* nominal + feedback-tracking real system (rolloutRMPPIKernel, rmppi_kernels.cu:666) */

typedef struct mppi_two_system_params {
  int32_t mode;                   /* MPPI_TWO_SYSTEM_TUBE or MPPI_TWO_SYSTEM_RMPPI */
  int32_t nominal_index;          /* RMPPI: index (0 or 1) of the nominal system; the other one is the real system */
  float value_function_threshold; /* RMPPI: value_func_threshold of the nominal cost */
  int32_t _pad;
  /* RMPPI feedback gains, device [T][S][C]: the real system's control is u + K_t (x_real - x_nominal) with
   * (K_t e)_j = sum_i gains[(t S + i) C + j] e_i (DDPFeedbackState::fb_gain_traj_ read through its Eigen map,
   * feedback_controllers/DDP/ddp.cuh:30-40,144-151). NULL: no feedback. */
  const float* feedback_gains;
} mppi_two_system_params;

/* Two-system rollout. x0_d [2][S] initial states, mean_d [2][T][C] the means the samples were drawn around,
 * samples_d [2][K][T][C] updated in place with the applied controls (the real system's with its feedback added; all
 * constrained), costs_d [2][K] (system d's sample k at d K + k).
 *   TUBE : cost_d = (sum_t [running(y_d,t) + lr_d(u_d,t)] ) / T + terminal(y_d) / T for each system alone -- equal to
 *          two mppi_rollout calls.
 *   RMPPI (n = nominal_index, r = 1 - n; rmppi_kernels.cu:487-664,666-798; the CPU form of
 *          tests/include/kernel_tests/core/rmppi_kernel_test.cu:7-70):
 *          fb_t = K_t (x_r - x_n); u_r += fb_t; both constrained; both stepped;
 *          cost_r = (sum_t [running(y_r) + lr_r(u_r)] + terminal(y_r)) / T,
 *          track_r = (sum_t [running(y_r) + fb_cost(fb_t)] + terminal(y_r)) / T,
 *          state_n = (sum_t running(y_n) + terminal(y_n)) / T,  lr_n = sum_t lr_n(u_n) / T,
 *          cost_n = 0.5 state_n + 0.5 max(min(track_r, value_function_threshold), state_n) + lr_n,
 *          fb_cost = 0.5 lambda (1 - alpha) sum_c coeff_c fb_c^2 / sigma_c^2 (computeFeedbackCost, gaussian.cu:577-630).
 * lr_d and fb_cost take their standard deviations (std_dev, or std_dev_t at step t) from distribution d's sampler:
 * sampler0 for d = 0, sampler1 for d = 1 (NULL: sampler0); every other sampler field (control_cost_coeff, the
 * pure-noise share) comes from sampler0, the one GaussianParams of both distributions (gaussian.cuh:21-42).
 * Built-in models except the BF16 network. */
int mppi_rollout_two_system(const mppi_dynamics_params* dyn, const mppi_cost_params* cost,
                            const mppi_sampler_params* sampler0, const mppi_sampler_params* sampler1,
                            const mppi_two_system_params* params, int32_t num_rollouts, int32_t num_timesteps, float dt,
                            float lambda, float alpha, const float* x0_d, const float* mean_d, float* samples_d,
                            float* costs_d, void* stream);

/* Two-system controller: TubeMPPIController (mode MPPI_TWO_SYSTEM_TUBE, controllers/Tube-MPPI/tube_mppi_controller.cu)
 * or RobustMPPIController's optimisation (mode MPPI_TWO_SYSTEM_RMPPI, controllers/R-MPPI/robust_mppi_controller.cu:
 * 635-756) over one sample buffer pair [2][K][T][C], the two-system rollout, and per system the weight transform and
 * the weighted update. Both distributions draw the same normals (use_same_noise_for_all_distributions, the default of
 * sampling_distributions/sampling_distribution.cuh:20) with the sampler's std_dev.
 *   Tube : system 0 is the actual system (from the measured state, mean = actual control), system 1 the nominal one
 *          (from the nominal trajectory's first state, mean = nominal control); after each iteration the nominal
 *          trajectory and control are replaced by the actual ones when baseline_actual < baseline_nominal +
 *          nominal_threshold (:261-276); then the nominal control is smoothed (SG) and both state trajectories are
 *          propagated. slide propagates the nominal state one step with the nominal control and slides both sequences
 *          (:318-331).
 *   RMPPI: system 0 is the nominal system (from the nominal state, set with mppi_tube_set_nominal_state), system 1
 *          the real one (from the measured state, with the feedback gains); both sample around the nominal control;
 *          the updates give the nominal control (system 0) and the real control (system 1), both smoothed; the nominal
 *          state trajectory is propagated. The nominal state is the caller's (the reference chooses it by its
 *          candidate evaluation, robust_mppi_controller.cu:548-628, before each call); slide is a no-op as there
 *          (robust_mppi_controller.cuh:190).
 * Getters and setters take `nominal` = 1 for the nominal system, 0 for the actual / real one. */
typedef struct mppi_tube_controller mppi_tube_controller;

typedef struct mppi_tube_desc {
  /* dynamics, cost, sampler, num_rollouts, num_timesteps, dt, lambda, alpha, num_iters, slide_control_scale,
   * init_control (both sequences), cost_map_host, nn_theta_host, std_dev_t_host, goal_traj_host; other fields unused */
  mppi_controller_desc base;
  int32_t mode;                   /* MPPI_TWO_SYSTEM_TUBE or MPPI_TWO_SYSTEM_RMPPI */
  float nominal_threshold;        /* Tube: TubeMPPIParams::nominal_threshold_ (tube_mppi_controller.cuh:17-21) */
  float value_function_threshold; /* RMPPI: value_function_threshold (robust_mppi_controller.cuh:46-53) */
  int32_t num_candidates;   /* RMPPI: num_candidate_nominal_states_, odd and >= 3 (0: 9) */
  int32_t num_eval_samples; /* RMPPI: samples per candidate of the evaluation, eval_dyn_kernel_dim_.x (0: 64) */
  int32_t _pad;
} mppi_tube_desc;

int mppi_tube_create(const mppi_tube_desc* desc, mppi_tube_controller** out);
int mppi_tube_destroy(mppi_tube_controller* ctrl);
int mppi_tube_compute_control(mppi_tube_controller* ctrl, const float* x0_host, int32_t optimization_stride);
int mppi_tube_get_control_seq(const mppi_tube_controller* ctrl, int32_t nominal, float* out_host); /* [T][C] */
int mppi_tube_set_control_seq(mppi_tube_controller* ctrl, int32_t nominal, const float* u_host);
int mppi_tube_get_state_seq(const mppi_tube_controller* ctrl, int32_t nominal, float* out_host);   /* [T][S] */
int mppi_tube_get_weight_stats(const mppi_tube_controller* ctrl, int32_t nominal, mppi_weight_stats* out_host);
int mppi_tube_get_free_energy(const mppi_tube_controller* ctrl, int32_t nominal, mppi_free_energy* out_host);
/* Tube: 1 if the last iteration kept the nominal system, 0 if it was replaced by the actual one (nominal_state_used) */
int mppi_tube_get_nominal_state_used(const mppi_tube_controller* ctrl, int32_t* out);
int mppi_tube_slide_control_seq(mppi_tube_controller* ctrl, int32_t steps);
/* The nominal state: RMPPI's nominal_state_, Tube's nominal trajectory start (set before the first call it replaces
 * the initialisation from the measured state). */
int mppi_tube_set_nominal_state(mppi_tube_controller* ctrl, const float* x_host);
int mppi_tube_get_nominal_state(const mppi_tube_controller* ctrl, float* x_host);
/* RMPPI feedback gains, host [T][S][C] (layout of mppi_two_system_params.feedback_gains); NULL clears them. */
int mppi_tube_set_feedback_gains(mppi_tube_controller* ctrl, const float* gains_host);
/* getFeedbackControl (controllers/controller.cuh:567-571, DDPFeedback::k_): u_fb = K_t (x - x_nominal). */
int mppi_tube_get_feedback_control(const mppi_tube_controller* ctrl, const float* x, const float* x_nominal, int32_t t,
                                   float* u_out);
int mppi_tube_set_thresholds(mppi_tube_controller* ctrl, float nominal_threshold, float value_function_threshold);
/* RMPPI: updateImportanceSamplingControl (robust_mppi_controller.cu:548-628), before computeControl. On the first call
 * the nominal state is the measured state (stride 0). Afterwards N candidate nominal states are formed on the line
 * through the nominal trajectory's first two states and the measured state (line-search weights, :480-498; candidates
 * :351-363), each with the importance-sampler stride round(w1 s + w2 s) (:501-511); samples are drawn around the
 * nominal control with optimization stride s, and every candidate rolls out the first M samples from its state with
 * its shifted controls (initEvalKernel, core/rmppi_kernels.cu:12-358); the chosen candidate is the last whose free
 * energy -lambda log(mean_j exp(-(c_ij - b) / lambda)) + b (b the smallest evaluation cost) is below the
 * value-function threshold (:513-546), else the previous choice. The nominal state becomes that candidate, the nominal
 * control is slid by its stride and the nominal trajectory is propagated. The feedback gains stay the caller's. */
int mppi_tube_update_importance_sampling_control(mppi_tube_controller* ctrl, const float* x_host, int32_t stride);
/* RMPPI: the last evaluation's candidate states [N][S], strides [N], evaluation costs [N][M], candidate free energies
 * [N] and the chosen index (any pointer may be NULL). */
int mppi_tube_get_candidates(const mppi_tube_controller* ctrl, float* states, int32_t* strides, float* costs,
                             float* free_energies, int32_t* best_index);
/* The last iteration's applied controls [2][K][T][C] and costs [2][K] (system index as above; either may be NULL). */
int mppi_tube_get_samples(const mppi_tube_controller* ctrl, float* samples_host, float* costs_host);
/* Benchmark utility (no counterpart in the reference): `reps` repeats of the last iteration run (same states, means,
 * generator offset, stride and iteration; Gaussian sampler only) launched eagerly with HIP events on the controller's
 * stream before the fused two-system rollout kernel, between it and the combine launch (both systems), and after it;
 * the average durations of the rollout launch and of the combine launch are returned, and the number of per-workgroup
 * partials per system (num_partials may be NULL). The repeats rewrite the samples and costs with the same
 * values; the host-side sequences and statistics are unchanged. */
int mppi_tube_time_kernels(mppi_tube_controller* ctrl, int32_t reps, float* rollout_ms, float* combine_ms,
                           int32_t* num_partials);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* MPPI_AMD_H */
