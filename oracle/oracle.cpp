/*
 * oracle.cpp -- CPU restatement of the reference MPPI hot path. TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the HIP engine: only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark load it. The product never links, loads or calls it.
 *
 * It restates, in plain C++ templated on the arithmetic type (float for the reference's FP32 path, double for the
 * FP64 error reports), the reference's host-side loop and plugin formulas:
 *   - rollout loop            tests/include/kernel_tests/core/rollout_kernel_test.cu:504-541
 *                             (with the likelihood-ratio argument order of gaussian.cuh:135-137, and the clamped
 *                             control written back as in core/mppi_common.cu:117-124)
 *   - enforceConstraints      dynamics/dynamics.cuh:247-261 (sign() of utils/math_utils.h:744-747)
 *   - Euler step              dynamics/dynamics.cuh:270-288
 *   - cartpole                dynamics/cartpole/cartpole_dynamics.cu:48-67 (+ normalizeAngle of the device path,
 *                             dynamics/cartpole/cartpole_dynamics.cu:89, utils/angle_utils.cuh:23-29)
 *   - double integrator       dynamics/double_integrator/di_dynamics.cu:14-21
 *   - kinematic bicycle       dynamics/racer_dubins/racer_dubins.cu:5-59 (+ normalizeAngle(yaw) of :157)
 *   - FNN forward             utils/nn_helpers/fnn_helper.cu:130-205 (layout), :354-383 (forward)
 *   - cartpole cost           cost_functions/cartpole/cartpole_quadratic_cost.cu:8-55
 *   - quadratic cost          cost_functions/quadratic_cost/quadratic_cost.cu:10-32
 *   - cost-map bilinear       utils/texture_helpers/two_d_texture_helper.cu:151-245, texture_helper.cu:94-134
 *   - speed term              cost_functions/autorally/ar_standard_cost.cu:284-297
 *   - Gaussian transform      sampling_distributions/gaussian/gaussian.cu:17-277
 *   - likelihood-ratio cost   sampling_distributions/gaussian/gaussian.cu:650-675
 *   - colored noise           sampling_distributions/colored_noise/colored_noise.cu:12-56,286-392,
 *                             colored_noise.cuh:24-34; the inverse real FFT is a direct O(n^2) DFT in FP64
 *   - weights / normalizer    core/mppi_common.cu:894-1002,1091-1117
 *   - weighted update         core/mppi_common.cu:1151-1196
 *   - SG smoothing / slide    controllers/controller.cuh:573-621
 *   - two-system rollouts     Tube-MPPI: controllers/Tube-MPPI/tube_mppi_controller.cu:158-236 (two independent
 *                             rollouts); RMPPI: tests/include/kernel_tests/core/rmppi_kernel_test.cu:7-70,
 *                             core/rmppi_kernels.cu:666-798, feedback cost gaussian.cu:632-647 (rollout_loop.hpp)
 *   - RMPPI candidate eval    core/rmppi_kernels.cu:12-358, rmppi_kernel_test.cu:73-127 (rollout_loop.hpp)
 *   - state trajectory        controllers/controller.cuh:664-684
 *   - Philox4x32-10           Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11
 *   - Box-Muller              Box & Muller 1958
 */
#include "../include/mppi_amd.h"
#include "rollout_loop.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace oracle {

/* ------------------------------------------------------------------------------------------------ Philox4x32-10 */
static inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform in (0, 1]: u = fma((float)x, 2^-32, 2^-33), a single rounding on every IEEE platform. */
static inline float u32_to_unit(uint32_t x)
{
  return std::fma((float)x, 2.3283064365386963e-10f, 1.1641532182693481e-10f);
}

/* Normal number i of the stream (seed, offset): block i/4 of Philox, Box-Muller on lanes (0,1) and (2,3). */
static inline void normals_block(uint64_t seed, uint64_t offset, uint64_t block, double z[4])
{
  const uint32_t ctr[4] = {(uint32_t)block, (uint32_t)(block >> 32), (uint32_t)offset, (uint32_t)(offset >> 32)};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  philox4x32_10(ctr, key, o);
  for (int p = 0; p < 2; ++p) {
    const double u1 = u32_to_unit(o[2 * p]);
    const double u2 = u32_to_unit(o[2 * p + 1]);
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 2.0 * M_PI * u2;
    z[2 * p] = r * std::cos(th);
    z[2 * p + 1] = r * std::sin(th);
  }
}

/* ------------------------------------------------------------------------------------------------------ helpers */
template <class R>
static inline R normalize_angle(R a)
{
  /* utils/angle_utils.cuh:13-29 (double and float overloads) */
  const R pi = (R)M_PI;
  const R res = std::fmod(a + pi, (R)(2.0 * M_PI));
  if (res <= (R)0) return res + pi;
  return res - pi;
}


/* Round to the nearest BF16 (ties to even), the conversion the matrix-core path applies to weights and layer
 * inputs. */
static float bf16_round(float f)
{
  uint32_t b;
  std::memcpy(&b, &f, 4);
  if ((b & 0x7f800000u) == 0x7f800000u) return f;
  b += 0x7fffu + ((b >> 16) & 1u);
  b &= 0xffff0000u;
  std::memcpy(&f, &b, 4);
  return f;
}

/* FNN forward, utils/nn_helpers/fnn_helper.cu:354-383: acts_{l+1} = tanh(W_l acts_l + b_l), last layer linear.
 * With `bf16`, every weight and every layer input is rounded to BF16 before the products (biases and sums stay in
 * R): the arithmetic of the BF16 matrix-core network. */
template <class R>
static void fnn_forward(int num_layers, const int32_t* layers, const float* theta, const R* in, R* out,
                        bool bf16 = false)
{
  R a[256], b[256];
  for (int i = 0; i < layers[0]; ++i) a[i] = in[i];
  int stride = 0;
  for (int l = 0; l + 1 < num_layers; ++l) {
    const int nin = layers[l], nout = layers[l + 1];
    const float* W = theta + stride;
    const float* bias = W + nin * nout;
    stride += nin * nout + nout;
    if (bf16)
      for (int k = 0; k < nin; ++k) a[k] = (R)bf16_round((float)a[k]);
    for (int j = 0; j < nout; ++j) {
      R s = 0;
      for (int k = 0; k < nin; ++k) s += (R)(bf16 ? bf16_round(W[j * nin + k]) : W[j * nin + k]) * a[k];
      s += (R)bias[j];
      b[j] = (l + 2 < num_layers) ? std::tanh(s) : s;
    }
    for (int j = 0; j < nout; ++j) a[j] = b[j];
  }
  for (int j = 0; j < layers[num_layers - 1]; ++j) out[j] = a[j];
}

/* State derivative of every model (host restatements cited in the header). */
template <class R>
static void compute_state_deriv(const mppi_dynamics_params& p, const R* x, const R* u, R* xd)
{
  switch (p.model) {
    case MPPI_MODEL_CARTPOLE: {
      const R th = normalize_angle(x[2]);
      const R s = std::sin(th), c = std::cos(th);
      const R td = x[3], F = u[0];
      const R mc = p.cart_mass, mp = p.pole_mass, l = p.pole_length, g = p.gravity;
      xd[0] = x[1];
      xd[1] = (F + mp * s * (l * td * td + g * c)) / (mc + mp * s * s);
      xd[2] = td;
      xd[3] = (-F * c - mp * l * td * td * c * s - (mc + mp) * g * s) / (l * (mc + mp * s * s));
      break;
    }
    case MPPI_MODEL_DOUBLE_INTEGRATOR:
      xd[0] = x[2];
      xd[1] = x[3];
      xd[2] = u[0];
      xd[3] = u[1];
      break;
    case MPPI_MODEL_RACER_DUBINS:
    case MPPI_MODEL_RACER_DUBINS_NN: {
      /* state: VEL_X, YAW, POS_X, POS_Y, STEER_ANGLE, BRAKE_STATE, STEER_ANGLE_RATE; control: THROTTLE_BRAKE, STEER */
      const bool brake = u[0] < (R)0;
      /* (enable_brake * -u0 - brake_state) * k, clamped to [-max_brake_rate_neg, max_brake_rate_pos] */
      xd[5] = std::min(std::max(((brake ? -u[0] : (R)0) - x[5]) * (R)p.brake_delay_constant, -(R)p.max_brake_rate_neg),
                       (R)p.max_brake_rate_pos);
      xd[0] = (brake ? (R)0 : (R)p.c_t * u[0] * (R)p.gear_sign) + (R)p.c_b * x[5] * (x[0] >= (R)0 ? (R)-1 : (R)1) -
              (R)p.c_v * x[0] + (R)p.c_0;
      xd[1] = (x[0] / (R)p.wheel_base) * std::tan(x[4] / (R)p.steer_angle_scale);
      const R yaw = normalize_angle(x[1]);
      xd[2] = x[0] * std::cos(yaw);
      xd[3] = x[0] * std::sin(yaw);
      xd[4] = std::max(std::min((u[1] * (R)p.steer_command_angle_scale - x[4]) * (R)p.steering_constant,
                                (R)p.max_steer_rate),
                       -(R)p.max_steer_rate);
      xd[6] = 0;
      if (p.model == MPPI_MODEL_RACER_DUBINS_NN) {
        R in[16], out[16];
        for (int i = 0; i < 7; ++i) in[i] = x[i];
        for (int i = 0; i < 2; ++i) in[7 + i] = u[i];
        fnn_forward<R>(p.nn_num_layers, p.nn_layers, p.nn_theta, in, out, p.nn_precision == MPPI_NN_BF16);
        for (int i = 0; i < p.nn_layers[p.nn_num_layers - 1]; ++i) xd[p.nn_output_states[i]] += out[i];
      }
      break;
    }
    default:
      break;
  }
}

template <class R>
static void update_state(const mppi_dynamics_params& p, const R* x, R* xn, const R* xd, R dt)
{
  if (p.model == MPPI_MODEL_RACER_DUBINS || p.model == MPPI_MODEL_RACER_DUBINS_NN) {
    /* dynamics/racer_dubins/racer_dubins.cu:44-59 */
    for (int i = 0; i < 6; ++i) xn[i] = x[i] + xd[i] * dt;
    xn[1] = normalize_angle(xn[1]);
    xn[4] = std::max(std::min(xn[4], (R)p.max_steer_angle), -(R)p.max_steer_angle);
    xn[6] = xd[4];
    xn[5] = std::min(std::max(xn[5], (R)0), -(R)p.control_min[0]);
  } else {
    for (int i = 0; i < p.state_dim; ++i) xn[i] = x[i] + xd[i] * dt;
  }
}

template <class R>
static void step(const mppi_dynamics_params& p, const R* x, R* xn, R* xd, const R* u, R* y, R dt)
{
  compute_state_deriv<R>(p, x, u, xd);
  update_state<R>(p, x, xn, xd, dt);
  for (int i = 0; i < p.output_dim && i < p.state_dim; ++i) y[i] = xn[i];
}

/* Bilinear cost-map query with the reference's CPU semantics (clamp address mode, linear filter). */
template <class R>
static R map_query(const mppi_cost_params& c, const float* map, R px, R py)
{
  const int W = c.map_width, H = c.map_height;
  /* world -> map frame R (p - origin) (worldPoseToMapPose, texture_helper.cu:94-104, in the plane) -> normalised
   * texture coordinate -> texel */
  R mx = px - (R)c.map_origin_x, my = py - (R)c.map_origin_y;
  if (c.map_rotated) {
    const R rx = (R)c.map_rotation[0] * mx + (R)c.map_rotation[1] * my;
    const R ry = (R)c.map_rotation[2] * mx + (R)c.map_rotation[3] * my;
    mx = rx;
    my = ry;
  }
  R nx = (mx / (R)c.map_resolution) / (R)W;
  R ny = (my / (R)c.map_resolution) / (R)H;
  R qx = nx * (R)W - (R)0.5;
  R qy = ny * (R)H - (R)0.5;
  if (qx > (R)(W - 1)) qx = (R)(W - 1);
  else if (qx <= (R)0) qx = 0;
  if (qy > (R)(H - 1)) qy = (R)(H - 1);
  else if (qy <= (R)0) qy = 0;
  const int x0 = std::min((int)std::floor(qx), W - 2), x1 = x0 + 1;
  const int y0 = std::min((int)std::floor(qy), H - 2), y1 = y0 + 1;
  const R q11 = map[y0 * W + x0], q12 = map[y0 * W + x1], q21 = map[y1 * W + x0], q22 = map[y1 * W + x1];
  const R ya = q11 * ((R)x1 - qx) + q12 * (qx - (R)x0);
  const R yb = q21 * ((R)x1 - qx) + q22 * (qx - (R)x0);
  return ya * ((R)y1 - qy) + yb * (qy - (R)y0);
}

template <class R>
static R quad_state(const mppi_cost_params& c, const R* y, const float* goal)
{
  R s = 0;
  for (int i = 0; i < c.output_dim; ++i) {
    const R e = y[i] - (R)goal[i];
    s += e * e * (R)c.coeff[i];
  }
  return s;
}

/* The goal of step t: row getIndex(t) = min(current_time + t, TIME_HORIZON - 1) of the goal trajectory
 * (cost_functions/quadratic_cost/quadratic_cost.cuh:40-57; c.goal_traj is a host pointer here), else the goal. */
static const float* goal_at(const mppi_cost_params& c, int t)
{
  if (c.goal_traj == nullptr) return c.goal;
  const int row = std::min(c.goal_time + t, c.goal_traj_len - 1);
  return c.goal_traj + (size_t)row * c.output_dim;
}

template <class R>
static R running_cost(const mppi_cost_params& c, const float* map, const R* y, int t)
{
  R s = quad_state<R>(c, y, c.cost == MPPI_COST_QUADRATIC_MAP ? goal_at(c, t) : c.goal);
  if (c.cost == MPPI_COST_QUADRATIC_MAP) {
    if (map != nullptr && c.map_coeff != 0.0f) s += (R)c.map_coeff * map_query<R>(c, map, y[c.map_x_index], y[c.map_y_index]);
    if (c.speed_coeff != 0.0f) {
      const R e = y[c.speed_index] - (R)c.speed_target;
      s += (R)c.speed_coeff * e * e;
    }
  }
  return s;
}

template <class R>
static R terminal_cost(const mppi_cost_params& c, const R* y)
{
  if (c.terminal_cost_coeff == 0.0f) return 0;
  return quad_state<R>(c, y, c.goal) * (R)c.terminal_cost_coeff;
}

/* The built-in models and costs as the rollout loop's system (rollout_loop.hpp). */
template <class R>
struct BuiltinSystem {
  const mppi_dynamics_params& dyn;
  const mppi_cost_params& cst;
  const float* map;
  int state_dim() const { return dyn.state_dim; }
  int control_dim() const { return dyn.control_dim; }
  int output_dim() const { return dyn.output_dim; }
  void enforce(R* u) const { enforce_constraints<R>(dyn, u); }
  void step(const R* x, R* xn, R* xd, const R* u, R* y, R dt) const { oracle::step<R>(dyn, x, xn, xd, u, y, dt); }
  R running(const R* y, const R*, int t) const { return running_cost<R>(cst, map, y, t); }
  R terminal(const R* y) const { return terminal_cost<R>(cst, y); }
};

/* The same loop recording, per sample, the post-step outputs y_{t+1} [K][T][O], the per-step costs
 * (running + likelihood-ratio) / T [K][T] and the terminal cost / T [K]. */
template <class R>
struct OutputTrace {
  int T;
  R* outputs;
  R* step_costs;
  R* terminal_costs;
  void step(int k, int t, const R* y, int O, R c) const
  {
    for (int i = 0; i < O; ++i) outputs[((size_t)k * T + t) * O + i] = y[i];
    step_costs[(size_t)k * T + t] = c;
  }
  void terminal(int k, R c) const { terminal_costs[k] = c; }
};

template <class R>
static void rollout_trace(const mppi_dynamics_params* dyn, const mppi_cost_params* cst, const float* map,
                          const mppi_sampler_params* smp, int K, int T, R dt, R lambda, R alpha, const float* x0,
                          const float* mean, float* samples, R* costs, R* outputs, R* step_costs, R* terminal)
{
  const BuiltinSystem<R> sys{*dyn, *cst, map};
  const OutputTrace<R> tr{T, outputs, step_costs, terminal};
  rollout_loop<R>(sys, *smp, K, T, dt, lambda, alpha, x0, mean, samples, costs, 0, K, tr);
}


} /* namespace oracle */

using namespace oracle;

extern "C" {

/* ------------------------------------------------------------------------------------------------ RNG exports */
void oracle_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

float oracle_u32_to_unit(uint32_t x) { return u32_to_unit(x); }

void oracle_normals(uint64_t seed, uint64_t offset, int64_t n, double* out)
{
  for (int64_t b = 0; b * 4 < n; ++b) {
    double z[4];
    normals_block(seed, offset, (uint64_t)b, z);
    for (int j = 0; j < 4 && b * 4 + j < n; ++j) out[b * 4 + j] = z[j];
  }
}

int oracle_pure_noise_start(int K, float pct) { return pure_noise_start(K, pct); }

/* ----------------------------------------------------------------------------------------- Gaussian transform */
void oracle_gaussian_transform(const mppi_sampler_params* s, int K, int T, int C, const float* mean, const float* eps,
                               int stride, int iteration, float* out)
{
  const int pstart = pure_noise_start(K, s->pure_noise_trajectories_percentage);
  const float decay = powf(s->std_dev_decay, (float)iteration);
  for (int k = 0; k < K; ++k)
    for (int t = 0; t < T; ++t)
      for (int c = 0; c < C; ++c) {
        const size_t i = ((size_t)k * T + t) * C + c;
        /* time_specific_std_dev: the step's row of the per-step table (gaussian.cu:38-43,313-318) */
        const float sd = decay * (s->std_dev_t ? s->std_dev_t[t * C + c] : s->std_dev[c]);
        float v;
        if (k == 0 || t < stride) v = mean[t * C + c];
        else if (k >= pstart) v = sd * eps[i];
        else v = mean[t * C + c] + sd * eps[i];
        out[i] = v;
      }
}

/* --------------------------------------------------------------------------------------------- colored noise */
/* Frequency scales [F][C] and sigma[C] exactly as the host code of colored_noise.cu:292-338 computes them (FP32). */
void oracle_colored_scales(const mppi_sampler_params* s, int T, int C, float* scales, float* sigma, int* n_out,
                           int* F_out)
{
  const int n = (s->fft_len_mode == MPPI_FFT_LEN_T) ? T : 2 * T;
  const int F = n / 2 + 1;
  std::vector<float> f(F);
  for (int i = 0; i < F; ++i) f[i] = i / (1.0f * n); /* fftfreq, colored_noise.cuh:24-34 */
  const float cutoff = fmaxf(s->fmin, 1.0f / n);
  int smaller = 0;
  for (int i = 0; i < F; ++i) {
    if (f[i] < cutoff) {
      smaller++;
    } else if (smaller < F) {
      for (int j = 0; j < smaller; ++j) {
        f[j] = f[smaller];
        for (int c = 0; c < C; ++c) scales[j * C + c] = powf(f[smaller], -s->exponents[c] / 2.0f);
      }
    }
    for (int c = 0; c < C; ++c) scales[i * C + c] = powf(f[i], -s->exponents[c] / 2.0f);
  }
  for (int c = 0; c < C; ++c) {
    float sg = 0;
    for (int j = 1; j < F - 1; ++j) sg += scales[j * C + c] * scales[j * C + c];
    const float last = scales[(F - 1) * C + c] * ((1.0f + (n % 2)) / 2.0f);
    sg += last * last;
    sigma[c] = 2.0f * sqrtf(sg) / n;
  }
  *n_out = n;
  *F_out = F;
}

/* Colored samples from unscaled spectra [K][C][F][2]. The inverse real transform is evaluated as a direct DFT in
 * double: x[t] = Re X0 + 2 sum_{0<f<n/2} (Re Xf cos - Im Xf sin)(2 pi f t / n) + [n even] Re X_{n/2} (-1)^t. */
void oracle_colored_samples(const mppi_sampler_params* s, int K, int T, int C, const float* mean, const float* spectra,
                            int stride, int iteration, float* out_eps, float* out_samples)
{
  int n = 0, F = 0;
  std::vector<float> scales((size_t)(2 * T + 2) * C), sigma(C);
  oracle_colored_scales(s, T, C, scales.data(), sigma.data(), &n, &F);
  std::vector<double> ct(n), st(n);
  for (int m = 0; m < n; ++m) {
    ct[m] = std::cos(2.0 * M_PI * m / n);
    st[m] = std::sin(2.0 * M_PI * m / n);
  }
  std::vector<double> x(n);
  std::vector<float> eps((size_t)K * T * C);
  for (int k = 0; k < K; ++k)
    for (int c = 0; c < C; ++c) {
      const float* X = spectra + ((size_t)k * C + c) * F * 2;
      for (int t = 0; t < n; ++t) {
        double acc = (double)(X[0] * scales[0 * C + c]);
        for (int f = 1; f < F; ++f) {
          const double re = (double)(X[2 * f] * scales[f * C + c]);
          /* configureFrequencyNoise zeroes the imaginary part at DC and, when F is odd, at the last bin (:26-33) */
          const double im = (f == F - 1 && (F % 2) == 1) ? 0.0 : (double)(X[2 * f + 1] * scales[f * C + c]);
          const int m = (int)(((long long)f * t) % n);
          if (2 * f == n) acc += re * ct[m];
          else acc += 2.0 * (re * ct[m] - im * st[m]);
        }
        x[t] = acc;
      }
      for (int t = 0; t < T; ++t) {
        const double dec = (s->offset_decay_rate == 0.0f) ? 0.0 : (double)powf(s->offset_decay_rate, (float)t);
        const double v = (x[t] - x[stride] * dec) / ((double)sigma[c] * n);
        eps[((size_t)k * T + t) * C + c] = (float)v;
      }
    }
  if (out_eps) std::memcpy(out_eps, eps.data(), eps.size() * sizeof(float));
  if (out_samples) oracle_gaussian_transform(s, K, T, C, mean, eps.data(), stride, iteration, out_samples);
}

/* ------------------------------------------------------------------------------------------------- dynamics */
void oracle_step_f32(const mppi_dynamics_params* p, const float* x, const float* u_in, float dt, float* xn, float* xd,
                     float* y)
{
  float u[MPPI_MAX_CONTROL];
  for (int i = 0; i < p->control_dim; ++i) u[i] = u_in[i];
  step<float>(*p, x, xn, xd, u, y, dt);
}

void oracle_state_deriv_f64(const mppi_dynamics_params* p, const double* x, const double* u, double* xd)
{
  compute_state_deriv<double>(*p, x, u, xd);
}

void oracle_enforce_constraints_f32(const mppi_dynamics_params* p, float* u) { enforce_constraints<float>(*p, u); }

void oracle_fnn_forward_f32(int num_layers, const int32_t* layers, const float* theta, const float* in, float* out)
{
  fnn_forward<float>(num_layers, layers, theta, in, out);
}

void oracle_fnn_forward_f64(int num_layers, const int32_t* layers, const float* theta, const double* in, double* out)
{
  fnn_forward<double>(num_layers, layers, theta, in, out);
}

void oracle_fnn_forward_bf16_f64(int num_layers, const int32_t* layers, const float* theta, const double* in,
                                 double* out)
{
  fnn_forward<double>(num_layers, layers, theta, in, out, true);
}

float oracle_bf16_round(float v) { return bf16_round(v); }

float oracle_map_query_f32(const mppi_cost_params* c, const float* map, float x, float y)
{
  return map_query<float>(*c, map, x, y);
}

float oracle_running_cost_f32(const mppi_cost_params* c, const float* map, const float* y, int t)
{
  return running_cost<float>(*c, map, y, t);
}

float oracle_terminal_cost_f32(const mppi_cost_params* c, const float* y) { return terminal_cost<float>(*c, y); }

float oracle_lr_cost_f32(const mppi_sampler_params* s, int C, const float* u, const float* mean_t, int pure,
                         float lambda, float alpha, int t)
{
  return lr_cost<float>(*s, C, u, mean_t, pure != 0, lambda, alpha, t);
}

/* Rollout of K samples; `samples` is updated in place with the clamped controls. map: host cost map or NULL.
 * nthreads > 1 splits the samples over std::thread workers (the multi-core CPU baseline). */
void oracle_rollout_f32(const mppi_dynamics_params* dyn, const mppi_cost_params* cst, const float* map,
                        const mppi_sampler_params* smp, int K, int T, float dt, float lambda, float alpha,
                        const float* x0, const float* mean, float* samples, float* costs, int nthreads)
{
  const BuiltinSystem<float> sys{*dyn, *cst, map};
  rollout_loop_threads<float>(sys, *smp, K, T, dt, lambda, alpha, x0, mean, samples, costs, nthreads);
}

void oracle_rollout_f64(const mppi_dynamics_params* dyn, const mppi_cost_params* cst, const float* map,
                        const mppi_sampler_params* smp, int K, int T, double dt, double lambda, double alpha,
                        const float* x0, const float* mean, float* samples, double* costs, int nthreads)
{
  /* FP64 run of the same loop; the clamped controls written back are rounded to float. */
  const BuiltinSystem<double> sys{*dyn, *cst, map};
  rollout_loop_threads<double>(sys, *smp, K, T, dt, lambda, alpha, x0, mean, samples, costs, nthreads);
}

/* Two-system rollout (two_system_loop, rollout_loop.hpp): smp[2] are the two distributions' samplers, nom < 0 Tube,
 * 0 / 1 RMPPI with that nominal index; gains host [T][S][C] or NULL; parts [K][4] or NULL. */
void oracle_rollout_two_system_f32(const mppi_dynamics_params* dyn, const mppi_cost_params* cst, const float* map,
                                   const mppi_sampler_params* smp, int nom, float vft, const float* gains, int K,
                                   int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                                   float* samples, float* costs, float* parts, int nthreads)
{
  const BuiltinSystem<float> sys{*dyn, *cst, map};
  const int nt = std::max(1, nthreads), chunk = (K + nt - 1) / nt;
  std::vector<std::thread> th;
  for (int w = 0; w < nt; ++w) {
    const int b = w * chunk, e = std::min(K, b + chunk);
    if (b >= e) break;
    th.emplace_back([&, b, e]() {
      two_system_loop<float>(sys, smp, nom, vft, gains, K, T, dt, lambda, alpha, x0, mean, samples, costs, parts, b, e);
    });
  }
  for (auto& t : th) t.join();
}

/* RMPPI candidate evaluation (candidate_eval_loop, rollout_loop.hpp). */
void oracle_candidate_eval_f32(const mppi_dynamics_params* dyn, const mppi_cost_params* cst, const float* map,
                               const mppi_sampler_params* smp, int N, int M, int K, int T, float dt, float lambda,
                               float alpha, const float* states, const int* strides, const float* mean,
                               const float* samples, float* costs)
{
  const BuiltinSystem<float> sys{*dyn, *cst, map};
  candidate_eval_loop<float>(sys, *smp, N, M, K, T, dt, lambda, alpha, states, strides, mean, samples, costs);
}

void oracle_rollout_trace_f32(const mppi_dynamics_params* dyn, const mppi_cost_params* cst, const float* map,
                              const mppi_sampler_params* smp, int K, int T, float dt, float lambda, float alpha,
                              const float* x0, const float* mean, float* samples, float* costs, float* outputs,
                              float* step_costs, float* terminal)
{
  rollout_trace<float>(dyn, cst, map, smp, K, T, dt, lambda, alpha, x0, mean, samples, costs, outputs, step_costs,
                       terminal);
}

void oracle_rollout_trace_f64(const mppi_dynamics_params* dyn, const mppi_cost_params* cst, const float* map,
                              const mppi_sampler_params* smp, int K, int T, double dt, double lambda, double alpha,
                              const float* x0, const float* mean, float* samples, double* costs, double* outputs,
                              double* step_costs, double* terminal)
{
  rollout_trace<double>(dyn, cst, map, smp, K, T, dt, lambda, alpha, x0, mean, samples, costs, outputs, step_costs,
                        terminal);
}

/* --------------------------------------------------------------------------------------------- weights/update */
int oracle_best_index(const float* costs, int K)
{
  float best = costs[0];
  int idx = 0;
  for (int i = 1; i < K; ++i)
    if (costs[i] < best) {
      best = costs[i];
      idx = i;
    }
  return idx;
}

/* w_k = exp(-(c_k - b)/lambda) in double; returns the normalizer, writes baseline and argmin. */
double oracle_weights(const float* costs, int K, double lambda, double* w, float* baseline, int* argmin,
                      double* sum_w2)
{
  const int idx = oracle_best_index(costs, K);
  const double b = costs[idx];
  double eta = 0, s2 = 0;
  for (int i = 0; i < K; ++i) {
    w[i] = std::exp(-((double)costs[i] - b) / lambda);
    eta += w[i];
    s2 += w[i] * w[i];
  }
  *baseline = costs[idx];
  *argmin = idx;
  if (sum_w2) *sum_w2 = s2;
  return eta;
}

/* The reference's FP32 device transform (core/mppi_common.cu:994-1002) for bit-level comparison: expf(-inv*(c-b)). */
void oracle_weights_f32(const float* costs, int K, float lambda, float* w)
{
  const float b = costs[oracle_best_index(costs, K)];
  const float inv = 1.0f / lambda;
  for (int i = 0; i < K; ++i) w[i] = expf(-inv * (costs[i] - b));
}

void oracle_weighted_update(const double* w, double eta, const float* samples, int K, int T, int C, double* out)
{
  for (int j = 0; j < T * C; ++j) out[j] = 0;
  for (int k = 0; k < K; ++k) {
    const double wk = w[k] / eta;
    const float* u = samples + (size_t)k * T * C;
    for (int j = 0; j < T * C; ++j) out[j] += wk * (double)u[j];
  }
}

/* core/mppi_common.cu:1101-1117 in double */
void oracle_free_energy(const double* w, int K, double baseline, double lambda, double* mean, double* var, double* mod)
{
  double norm = 0, s2 = 0;
  for (int i = 0; i < K; ++i) {
    norm += w[i];
    s2 += w[i] * w[i];
  }
  norm /= K;
  *mean = -lambda * std::log(norm) + baseline;
  *var = lambda * (s2 / K - norm * norm);
  const double weird = *var / (norm * std::sqrt((double)K));
  *mod = lambda * (weird + 0.5 * weird * weird);
}

/* ColoredMPPIController's per-iteration host path (controllers/ColoredMPPI/colored_mppi_controller.cu:185-221):
 * baseline = computeBaselineCost (first strict minimum, core/mppi_common.cu:894-935), the FP32 transform of the device
 * (normExp, :994-1002, with lambda_inv = (float)(1.0 / lambda); or Tsallis when gamma != 0 and r != 0, :1004-1021),
 * the host normalizer (an FP64 sum of the FP32 weights returned as float, :1091-1099) and computeFreeEnergy with its
 * FP32 accumulators (:1101-1117). w receives the weights. */
float oracle_colored_weights_f32(const float* costs, int K, float lambda, float gamma, float r, float* w,
                                 float* baseline, float* fe_mean, float* fe_var, float* fe_mod)
{
  const float b = costs[oracle_best_index(costs, K)];
  const bool tsallis = gamma != 0.0f && r != 0.0f;
  const float lambda_inv = (float)(1.0 / lambda);
  for (int i = 0; i < K; ++i) {
    const float dif = costs[i] - b;
    if (tsallis) w[i] = dif < gamma ? expf(logf((float)(1.0 - (double)(dif / gamma))) / (r - 1)) : 0.0f;
    else w[i] = expf(-lambda_inv * dif);
  }
  double normalizer = 0.0;
  for (int i = 0; i < K; ++i) normalizer += w[i];
  float var = 0, norm = 0;
  for (int i = 0; i < K; ++i) {
    norm += w[i];
    var += w[i] * w[i];
  }
  norm /= K;
  *fe_mean = -lambda * logf(norm) + b;
  *fe_var = lambda * (var / K - norm * norm);
  const float weird = *fe_var / (norm * sqrtf(1.0f * K));
  *fe_mod = lambda * (weird + 0.5f * weird * weird);
  *baseline = b;
  return (float)normalizer;
}

/* Dynamics::enforceLeash (dynamics/dynamics.cuh:445-463) and the bicycle's (dynamics/racer_dubins/racer_dubins.cu:
 * 176-238: position leashed in the body frame of the true yaw, yaw along shortestAngularDistance and normalized). */
void oracle_enforce_leash(const mppi_dynamics_params* p, const float* xt, const float* xn, const float* lv, float* out)
{
  const int S = p->state_dim;
  const bool bicycle = p->model == MPPI_MODEL_RACER_DUBINS || p->model == MPPI_MODEL_RACER_DUBINS_NN;
  for (int i = 0; i < S; ++i) out[i] = xt[i];
  if (bicycle) {
    const float c = cosf(xt[1]), sn = sinf(xt[1]);
    float dx = xn[2] - xt[2], dy = xn[3] - xt[3];
    const float bx = fminf(fmaxf(dx * c + dy * sn, -lv[2]), lv[2]);
    const float by = fminf(fmaxf(-dx * sn + dy * c, -lv[3]), lv[3]);
    out[2] += bx * c - by * sn;
    out[3] += bx * sn + by * c;
  }
  for (int i = 0; i < S; ++i) {
    if (bicycle && (i == 2 || i == 3)) continue;
    const bool yaw = bicycle && i == 1;
    const float diff = yaw ? normalize_angle<float>(xn[i] - xt[i]) : xn[i] - xt[i];
    if (lv[i] < fabsf(diff)) {
      out[i] = xt[i] + fminf(fmaxf(diff, -lv[i]), lv[i]);
      if (yaw) out[i] = normalize_angle<float>(out[i]);
    } else {
      out[i] = xn[i];
    }
  }
}

/* ------------------------------------------------------------------------------------------ host post-process */
void oracle_smooth(float* u, int T, int C)
{
  const float coef[5] = {-3.0f / 35.0f, 12.0f / 35.0f, 17.0f / 35.0f, 12.0f / 35.0f, -3.0f / 35.0f};
  std::vector<float> buf((size_t)(T + 4) * C);
  for (int c = 0; c < C; ++c) {
    buf[0 * C + c] = u[c];
    buf[1 * C + c] = u[c];
    for (int t = 0; t < T; ++t) buf[(t + 2) * C + c] = u[t * C + c];
    buf[(T + 2) * C + c] = u[(T - 1) * C + c];
    buf[(T + 3) * C + c] = u[(T - 1) * C + c];
  }
  for (int i = 1; i < T; ++i)
    for (int c = 0; c < C; ++c) {
      float s = 0;
      for (int j = 0; j < 5; ++j) s += coef[j] * buf[(i + j) * C + c];
      u[i * C + c] = s;
    }
}

void oracle_slide(float* u, int T, int C, int steps, const float* zero_control, const float* scale)
{
  for (int i = 0; i < T; ++i) {
    const int ind = std::min(i + steps, T - 1);
    for (int c = 0; c < C; ++c) {
      float v = u[ind * C + c];
      if (i + steps > T - 1) v = (v - zero_control[c]) * scale[c] + zero_control[c];
      u[i * C + c] = v;
    }
  }
}

/* The same propagation in FP64 (the controls clamped in FP64), for error reports. */
void oracle_state_trajectory_f64(const mppi_dynamics_params* p, const float* x0, const float* u, int T, double dt,
                                 double* states)
{
  const int S = p->state_dim, C = p->control_dim;
  for (int i = 0; i < S; ++i) states[i] = x0[i];
  double xd[MPPI_MAX_STATE], y[MPPI_MAX_OUTPUT], uc[MPPI_MAX_CONTROL];
  for (int t = 0; t + 1 < T; ++t) {
    for (int c = 0; c < C; ++c) uc[c] = u[t * C + c];
    enforce_constraints<double>(*p, uc);
    step<double>(*p, states + t * S, states + (t + 1) * S, xd, uc, y, dt);
  }
}

/* states [T][S]: state[0] = x0, state[i+1] = step(state[i], clamp(u[i])) */
void oracle_state_trajectory(const mppi_dynamics_params* p, const float* x0, const float* u, int T, float dt,
                             float* states)
{
  const int S = p->state_dim, C = p->control_dim;
  for (int i = 0; i < S; ++i) states[i] = x0[i];
  float xd[MPPI_MAX_STATE], y[MPPI_MAX_OUTPUT], uc[MPPI_MAX_CONTROL];
  for (int t = 0; t + 1 < T; ++t) {
    for (int c = 0; c < C; ++c) uc[c] = u[t * C + c];
    enforce_constraints<float>(*p, uc);
    step<float>(*p, states + t * S, states + (t + 1) * S, xd, uc, y, dt);
  }
}

} /* extern "C" */
