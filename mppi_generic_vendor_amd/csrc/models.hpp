// Dynamics and cost plugins of the engine, as compile-time functors over the POD parameter blocks of mppi_amd.h.
// The same functions run per lane inside the rollout kernels and on the host for the state-trajectory propagation
// that follows an MPC solve. Reference formulas are cited per function.
#pragma once

#include <cstring>

#include "common.hpp"

namespace mppi_amd {

// utils/angle_utils.cuh:23-29 (float version)
// fmodf(a + pi, 2 pi) takes no remainder loop for |a + pi| < 4 pi: there it is a + pi itself (below 2 pi) or
// a + pi -+ 2 pi, which is exact (Sterbenz) -- the value fmodf returns up to the sign of a zero, which the final
// select maps to pi either way.
MPPI_HD float normalize_angle(float a)
{
  const float pi = 3.14159265358979323846f;
  const float two_pi = 6.28318530717958647692f;
  const float s = a + pi;
  const float m = fabsf(s);
  const float r = m < two_pi ? s : (m < 2.0f * two_pi ? s - copysignf(two_pi, s) : fmodf(s, two_pi));
  return (r <= 0.0f) ? r + pi : r - pi;
}

// sin and cos of an angle in radians. On the device: the hardware v_sin_f32 / v_cos_f32, which take revolutions,
// on fract(a / 2 pi) -- the role of normalizeAngle + __sinf/__cosf in the reference's device code
// (dynamics/cartpole/cartpole_dynamics.cu:89-91); on the host: normalizeAngle + sinf/cosf.
MPPI_HD void sincos_angle(float a, float* s, float* c)
{
#ifdef __HIP_DEVICE_COMPILE__
  const float rev = __builtin_amdgcn_fractf(a * 0.15915494309189535f);
  *s = __builtin_amdgcn_sinf(rev);
  *c = __builtin_amdgcn_cosf(rev);
#else
  const float th = normalize_angle(a);
  *s = sinf(th);
  *c = cosf(th);
#endif
}

// 1 / x: v_rcp_f32 on the device (1 ulp), an IEEE division on the host.
MPPI_HD float recip(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// Dynamics::enforceConstraints (dynamics/dynamics.cu:107-126): deadband, then clamp to the control range.
// sign() is utils/math_utils.h:744-747 (sign(0) == 1).
template <int C>
MPPI_HD void enforce_constraints(const mppi_dynamics_params& p, float* u)
{
#pragma unroll
  for (int i = 0; i < C; ++i) {
    float v = u[i];
    if (fabsf(v) < p.control_deadband[i]) v = p.zero_control[i];
    else v += p.control_deadband[i] * -(v >= 0.0f ? 1.0f : -1.0f);
    u[i] = fminf(fmaxf(p.control_min[i], v), p.control_max[i]);
  }
}

// enforce_constraints for control i alone (constraints act on each control separately); `deadband` selects the
// full form (deadband + clamp), as enforce_constraints, or the clamp alone, as clamp_controls, with the same
// arithmetic, so that whole sample tiles can be constrained before the rollout instead of step by step.
MPPI_HD inline float constrain_control(const mppi_dynamics_params& p, int i, float v, bool deadband)
{
  if (deadband) {
    if (fabsf(v) < p.control_deadband[i]) v = p.zero_control[i];
    else v += p.control_deadband[i] * -(v >= 0.0f ? 1.0f : -1.0f);
  }
  return fminf(fmaxf(p.control_min[i], v), p.control_max[i]);
}

// Element transforms for staging a sample tile (f(col, v), col = t C + c): the clamp alone, with the bounds held in
// registers, and the full constraint (deadband + clamp). The kernels pick one for the whole tile.
template <int C>
struct ControlClamp {
  float lo[C], hi[C];
  __device__ explicit ControlClamp(const mppi_dynamics_params& p)
  {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      lo[c] = p.control_min[c];
      hi[c] = p.control_max[c];
    }
  }
  __device__ float operator()(int col, float v) const
  {
    const int c = col % C;
    float l = lo[0], h = hi[0];
#pragma unroll
    for (int k = 1; k < C; ++k) {
      l = (c == k) ? lo[k] : l;
      h = (c == k) ? hi[k] : h;
    }
    return fminf(fmaxf(l, v), h);
  }
};

template <int C>
struct ControlConstrain {
  const mppi_dynamics_params* p;
  __device__ float operator()(int col, float v) const { return constrain_control(*p, col % C, v, true); }
};

// enforce_constraints without a deadband (all control_deadband == 0): the clamp alone.
template <int C>
MPPI_HD void clamp_controls(const mppi_dynamics_params& p, float* u)
{
#pragma unroll
  for (int i = 0; i < C; ++i) u[i] = fminf(fmaxf(p.control_min[i], u[i]), p.control_max[i]);
}

// Savitzky-Golay taps of smooth_control (controllers/controller.cuh:573-607), in the same FP32 order with products and
// sums rounded separately (no FMA contraction), so the device's smoothed sequence is bit-identical to the host's.
MPPI_HD float sg_smooth_at(const float* u, int t, int c, int T, int C)
{
#pragma clang fp contract(off)
  const float coef[5] = {-3.0f / 35.0f, 12.0f / 35.0f, 17.0f / 35.0f, 12.0f / 35.0f, -3.0f / 35.0f};
  if (t == 0) return u[c];
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; ++k) s = s + coef[k] * u[min(max(t + k - 2, 0), T - 1) * C + c];
  return s;
}

// enforce_constraints of control i from copied limits (the same arithmetic), for kernels that constrain a whole
// control sequence without the dynamics parameter block.
struct ControlLimits {
  float lo[MPPI_MAX_CONTROL], hi[MPPI_MAX_CONTROL], dead[MPPI_MAX_CONTROL], zero[MPPI_MAX_CONTROL];
};

inline ControlLimits control_limits(const mppi_dynamics_params& p)
{
  ControlLimits L;
  for (int i = 0; i < MPPI_MAX_CONTROL; ++i) {
    L.lo[i] = p.control_min[i];
    L.hi[i] = p.control_max[i];
    L.dead[i] = p.control_deadband[i];
    L.zero[i] = p.zero_control[i];
  }
  return L;
}

MPPI_HD float constrain_with(const ControlLimits& L, int i, float v)
{
  if (fabsf(v) < L.dead[i]) v = L.zero[i];
  else v += L.dead[i] * -(v >= 0.0f ? 1.0f : -1.0f);
  return fminf(fmaxf(L.lo[i], v), L.hi[i]);
}

template <int MODEL>
struct Model;

// Cart-pole, dynamics/cartpole/cartpole_dynamics.cu:87-105 (with accurate sinf/cosf).
template <>
struct Model<MPPI_MODEL_CARTPOLE> {
  static constexpr int S = 4, C = 1, O = 4;
  MPPI_HD static void deriv(const mppi_dynamics_params& p, const float* x, const float* u, float* xd)
  {
    float s, c;
    sincos_angle(x[2], &s, &c);
    const float td = x[3], F = u[0];
    const float mc = p.cart_mass, mp = p.pole_mass, l = p.pole_length, g = p.gravity;
    const float inv_den = recip(mc + mp * s * s);
    xd[0] = x[1];
    xd[1] = (F + mp * s * (l * td * td + g * c)) * inv_den;
    xd[2] = td;
    xd[3] = (-F * c - mp * l * td * td * c * s - (mc + mp) * g * s) * (inv_den * recip(l));
  }
  MPPI_HD static void update(const mppi_dynamics_params&, const float* x, float* xn, const float* xd, float dt)
  {
#pragma unroll
    for (int i = 0; i < S; ++i) xn[i] = x[i] + xd[i] * dt;
  }
};

// Double integrator, dynamics/double_integrator/di_dynamics.cu:46-53.
template <>
struct Model<MPPI_MODEL_DOUBLE_INTEGRATOR> {
  static constexpr int S = 4, C = 2, O = 4;
  MPPI_HD static void deriv(const mppi_dynamics_params&, const float* x, const float* u, float* xd)
  {
    xd[0] = x[2];
    xd[1] = x[3];
    xd[2] = u[0];
    xd[3] = u[1];
  }
  MPPI_HD static void update(const mppi_dynamics_params&, const float* x, float* xn, const float* xd, float dt)
  {
#pragma unroll
    for (int i = 0; i < S; ++i) xn[i] = x[i] + xd[i] * dt;
  }
};

// Kinematic bicycle (RacerDubins), dynamics/racer_dubins/racer_dubins.cu:5-33,138-166 (deriv) and :44-59 (update).
// State: VEL_X, YAW, POS_X, POS_Y, STEER_ANGLE, BRAKE_STATE, STEER_ANGLE_RATE. Control: THROTTLE_BRAKE, STEER_CMD.
struct Bicycle {
  static constexpr int S = 7, C = 2, O = 7;
  MPPI_HD static void deriv(const mppi_dynamics_params& p, const float* x, const float* u, float* xd)
  {
    const bool brake = u[0] < 0.0f;
    xd[5] = fminf(fmaxf(((brake ? -u[0] : 0.0f) - x[5]) * p.brake_delay_constant, -p.max_brake_rate_neg),
                  p.max_brake_rate_pos);
    xd[0] = (brake ? 0.0f : p.c_t * u[0] * (float)p.gear_sign) + p.c_b * x[5] * (x[0] >= 0.0f ? -1.0f : 1.0f) -
            p.c_v * x[0] + p.c_0;
    float ss, cs;
    sincos_angle(x[4] * recip(p.steer_angle_scale), &ss, &cs);
    xd[1] = (x[0] * recip(p.wheel_base)) * (ss * recip(cs));
    float sy, cy;
    sincos_angle(x[1], &sy, &cy);
    xd[2] = x[0] * cy;
    xd[3] = x[0] * sy;
    xd[4] = fmaxf(fminf((u[1] * p.steer_command_angle_scale - x[4]) * p.steering_constant, p.max_steer_rate),
                  -p.max_steer_rate);
    xd[6] = 0.0f;
  }
  MPPI_HD static void update(const mppi_dynamics_params& p, const float* x, float* xn, const float* xd, float dt)
  {
#pragma unroll
    for (int i = 0; i < 6; ++i) xn[i] = x[i] + xd[i] * dt;
    xn[1] = normalize_angle(xn[1]);
    xn[4] = fmaxf(fminf(xn[4], p.max_steer_angle), -p.max_steer_angle);
    xn[6] = xd[4];
    xn[5] = fminf(fmaxf(xn[5], 0.0f), -p.control_min[0]);
  }
};

template <>
struct Model<MPPI_MODEL_RACER_DUBINS> : Bicycle {};

// Bicycle + residual MLP [x(7), u(2)] -> 64 -> 64 -> 4 (tanh hidden, linear out) added to the derivatives of
// nn_output_states. The per-lane FP32 path below reads the uniform weights through scalar loads; the BF16 MFMA path
// lives in the rollout kernel. Layer layout: FNNHelper (utils/nn_helpers/fnn_helper.cu:130-205).
constexpr int NN_IN = 9, NN_H1 = 64, NN_H2 = 64, NN_OUT = 4;
constexpr int NN_PARAMS = (NN_IN + 1) * NN_H1 + (NN_H1 + 1) * NN_H2 + (NN_H2 + 1) * NN_OUT;

MPPI_HD float tanh_acc(float v) { return tanhf(v); }

// Per-lane evaluation of the BF16 network (nn_precision == MPPI_NN_BF16): weights and layer inputs rounded to BF16
// (ties to even), biases and sums in FP32 -- the arithmetic of the matrix-core path, for the host-side state
// propagation and the per-trajectory re-rolls of the sampled-trajectory export.
MPPI_HD float bf16_round_host(float f)
{
  uint32_t b;
  __builtin_memcpy(&b, &f, 4);
  if ((b & 0x7f800000u) == 0x7f800000u) return f;
  b += 0x7fffu + ((b >> 16) & 1u);
  b &= 0xffff0000u;
  __builtin_memcpy(&f, &b, 4);
  return f;
}

MPPI_HD void nn_forward_f32(const float* __restrict__ th, const float* in, float* out)
{
  float h1[NN_H1];
  const float* W1 = th;
  const float* b1 = W1 + NN_IN * NN_H1;
#pragma unroll 4
  for (int j = 0; j < NN_H1; ++j) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < NN_IN; ++k) s += W1[j * NN_IN + k] * in[k];
    h1[j] = tanh_acc(s + b1[j]);
  }
  const float* W2 = b1 + NN_H1;
  const float* b2 = W2 + NN_H1 * NN_H2;
  float h2[NN_H2];
#pragma unroll 2
  for (int j = 0; j < NN_H2; ++j) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < NN_H1; ++k) s += W2[j * NN_H1 + k] * h1[k];
    h2[j] = tanh_acc(s + b2[j]);
  }
  const float* W3 = b2 + NN_H2;
  const float* b3 = W3 + NN_H2 * NN_OUT;
#pragma unroll
  for (int j = 0; j < NN_OUT; ++j) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < NN_H2; ++k) s += W3[j * NN_H2 + k] * h2[k];
    out[j] = s + b3[j];
  }
}

MPPI_HD void nn_forward_bf16_host(const float* th, const float* in, float* out)
{
  float a[NN_H1], b[NN_H1];
  const int layers[4] = {NN_IN, NN_H1, NN_H2, NN_OUT};
  for (int i = 0; i < NN_IN; ++i) a[i] = in[i];
  int off = 0;
  for (int l = 0; l < 3; ++l) {
    const int nin = layers[l], nout = layers[l + 1];
    const float* W = th + off;
    const float* bias = W + nin * nout;
    off += nin * nout + nout;
    for (int k = 0; k < nin; ++k) a[k] = bf16_round_host(a[k]);
    for (int j = 0; j < nout; ++j) {
      float s = 0.0f;
      for (int k = 0; k < nin; ++k) s += bf16_round_host(W[j * nin + k]) * a[k];
      s += bias[j];
      b[j] = l < 2 ? tanh_acc(s) : s;
    }
    for (int j = 0; j < nout; ++j) a[j] = b[j];
  }
  for (int j = 0; j < NN_OUT; ++j) out[j] = a[j];
}

// The BF16-network bicycle evaluated one sample per lane (the export's re-rolls; the rollout kernels evaluate it on
// the matrix cores).
struct BicycleNNBf16Lane : Bicycle {
  MPPI_HD static void deriv(const mppi_dynamics_params& p, const float* x, const float* u, float* xd)
  {
    Bicycle::deriv(p, x, u, xd);
    float in[NN_IN], out[NN_OUT];
#pragma unroll
    for (int i = 0; i < 7; ++i) in[i] = x[i];
    in[7] = u[0];
    in[8] = u[1];
    nn_forward_bf16_host(p.nn_theta, in, out);
#pragma unroll
    for (int i = 0; i < NN_OUT; ++i) xd[p.nn_output_states[i]] += out[i];
  }
};

template <>
struct Model<MPPI_MODEL_RACER_DUBINS_NN> : Bicycle {
  MPPI_HD static void deriv(const mppi_dynamics_params& p, const float* x, const float* u, float* xd)
  {
    Bicycle::deriv(p, x, u, xd);
    float in[NN_IN], out[NN_OUT];
#pragma unroll
    for (int i = 0; i < 7; ++i) in[i] = x[i];
    in[7] = u[0];
    in[8] = u[1];
#ifdef __HIP_DEVICE_COMPILE__
    nn_forward_f32(p.nn_theta, in, out);
#else
    if (p.nn_precision == MPPI_NN_BF16) nn_forward_bf16_host(p.nn_theta, in, out);
    else nn_forward_f32(p.nn_theta, in, out);
#endif
#pragma unroll
    for (int i = 0; i < NN_OUT; ++i) xd[p.nn_output_states[i]] += out[i];
  }
};

// ---------------------------------------------------------------------------------------------------------- costs
// Bilinear cost-map query with the reference's CPU semantics (utils/texture_helpers/two_d_texture_helper.cu:151-245
// and the world->texture transform of texture_helper.cu:94-134, its rotation restricted to the plane): clamp address
// mode, texel centres at +0.5, x0 = min(floor(q), W - 2). Split into the fetch of the four texels (map_fetch) and their
// blend (map_blend), so that a caller can have the fetches of several queries in flight before it blends them.
struct MapTap {
  float q11, q12, q21, q22;  // texels (x0, y0), (x1, y0), (x0, y1), (x1, y1)
  float qx, qy;              // clamped texel coordinates
  int x0, y0;
};

MPPI_HD MapTap map_fetch(const mppi_cost_params& c, float px, float py)
{
  const int W = c.map_width, H = c.map_height;
  // map frame R (p - origin) (worldPoseToMapPose; R uniform, so the branch is too)
  float dx = px - c.map_origin_x, dy = py - c.map_origin_y;
  if (c.map_rotated) {
    const float rx = c.map_rotation[0] * dx + c.map_rotation[1] * dy;
    const float ry = c.map_rotation[2] * dx + c.map_rotation[3] * dy;
    dx = rx;
    dy = ry;
  }
  // texel coordinate m / resolution - 0.5 (the reference's normalise-by-extent and scale-back cancel); the reciprocal
  // of the uniform resolution is computed once per kernel
  const float inv_res = 1.0f / c.map_resolution;
  const float qx = dx * inv_res - 0.5f, qy = dy * inv_res - 0.5f;
  // clamp to [0, W - 1] as the reference's if/else chain does for every non-NaN value; a NaN query maps to texel 0
  // instead of reaching the index computation (v_med3_f32 with a NaN operand returns the minimum of the others); the
  // clamped coordinate is >= 0, so truncation gives the floor
  MapTap m;
#ifdef __HIP_DEVICE_COMPILE__
  m.qx = __builtin_amdgcn_fmed3f(qx, 0.0f, (float)(W - 1));
  m.qy = __builtin_amdgcn_fmed3f(qy, 0.0f, (float)(H - 1));
#else
  m.qx = fminf(fmaxf(qx, 0.0f), (float)(W - 1));
  m.qy = fminf(fmaxf(qy, 0.0f), (float)(H - 1));
#endif
  m.x0 = min((int)m.qx, W - 2);
  m.y0 = min((int)m.qy, H - 2);
  // 32-bit texel index (maps hold at most 2^30 values); on the device each row's texel pair (x0, x0 + 1) is one
  // 8-byte global load at 4-byte alignment
  const uint32_t i0 = (uint32_t)m.y0 * (uint32_t)W + (uint32_t)m.x0, i1 = i0 + (uint32_t)W;
#ifdef __HIP_DEVICE_COMPILE__
  struct __attribute__((aligned(4))) TexelPair {
    float a, b;
  };
  using GlobalPair = const __attribute__((address_space(1))) TexelPair;
  using GlobalFloat = const __attribute__((address_space(1))) float;
  GlobalFloat* map = (GlobalFloat*)c.map;
  GlobalPair* r0 = (GlobalPair*)(map + i0);
  GlobalPair* r1 = (GlobalPair*)(map + i1);
  m.q11 = r0->a;
  m.q12 = r0->b;
  m.q21 = r1->a;
  m.q22 = r1->b;
#else
  const float* map = c.map;
  m.q11 = map[i0];
  m.q12 = map[i0 + 1];
  m.q21 = map[i1];
  m.q22 = map[i1 + 1];
#endif
  return m;
}

MPPI_HD float map_blend(const MapTap& m)
{
  const int x1 = m.x0 + 1, y1 = m.y0 + 1;
  const float ya = m.q11 * ((float)x1 - m.qx) + m.q12 * (m.qx - (float)m.x0);
  const float yb = m.q21 * ((float)x1 - m.qx) + m.q22 * (m.qx - (float)m.x0);
  return ya * ((float)y1 - m.qy) + yb * (m.qy - (float)m.y0);
}

MPPI_HD float map_query(const mppi_cost_params& c, float px, float py) { return map_blend(map_fetch(c, px, py)); }

template <int O>
MPPI_HD float quad_state(const mppi_cost_params& c, const float* y)
{
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < O; ++i) {
    const float e = y[i] - c.goal[i];
    s += e * e * c.coeff[i];
  }
  return s;
}

// The same term against the goal of step t: row min(goal_time + t, goal_traj_len - 1) of the goal trajectory
// (QuadraticCostTrajectoryParams::getIndex, cost_functions/quadratic_cost/quadratic_cost.cuh:48-57), or the single
// goal without one. t is uniform across the lanes that evaluate one step, so the row is read with scalar loads.
template <int O>
MPPI_HD float quad_state_at(const mppi_cost_params& c, const float* y, int t)
{
  if (c.goal_traj == nullptr) return quad_state<O>(c, y);
  const int row = c.goal_time + t < c.goal_traj_len ? c.goal_time + t : c.goal_traj_len - 1;
  const float* g = c.goal_traj + (size_t)row * O;
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < O; ++i) {
    const float e = y[i] - g[i];
    s += e * e * c.coeff[i];
  }
  return s;
}

template <int COST, int O>
struct Cost;

// cost_functions/cartpole/cartpole_quadratic_cost.cu:23-55
template <int O>
struct Cost<MPPI_COST_CARTPOLE_QUADRATIC, O> {
  MPPI_HD static float running(const mppi_cost_params& c, const float* y, int) { return quad_state<O>(c, y); }
  MPPI_HD static float terminal(const mppi_cost_params& c, const float* y)
  {
    return quad_state<O>(c, y) * c.terminal_cost_coeff;
  }
};

// Quadratic goal term of step t (cost_functions/quadratic_cost/quadratic_cost.cu:35-54, the goal trajectory's row
// for t when one is given) + cost-map term + speed term (cost_functions/autorally/ar_standard_cost.cu:284-297);
// terminal = terminal_cost_coeff * quadratic term against the single goal (0 by default, as the reference's
// QuadraticCost::terminalCost).
template <int O>
struct Cost<MPPI_COST_QUADRATIC_MAP, O> {
  MPPI_HD static bool uses_map(const mppi_cost_params& c) { return c.map != nullptr && c.map_coeff != 0.0f; }
  // the running cost with the map term's texels already fetched (tap, read only when uses_map)
  MPPI_HD static float running_tap(const mppi_cost_params& c, const float* y, int t, const MapTap& tap)
  {
    float s = quad_state_at<O>(c, y, t);
    if (uses_map(c)) s += c.map_coeff * map_blend(tap);
    if (c.speed_coeff != 0.0f) {
      const float e = y[c.speed_index] - c.speed_target;
      s += c.speed_coeff * e * e;
    }
    return s;
  }
  MPPI_HD static float running(const mppi_cost_params& c, const float* y, int t)
  {
    MapTap tap{};
    if (uses_map(c)) tap = map_fetch(c, y[c.map_x_index], y[c.map_y_index]);
    return running_tap(c, y, t, tap);
  }
  MPPI_HD static float terminal(const mppi_cost_params& c, const float* y)
  {
    return (c.terminal_cost_coeff == 0.0f) ? 0.0f : quad_state<O>(c, y) * c.terminal_cost_coeff;
  }
};

// Likelihood-ratio control cost of step t (sampling_distributions/gaussian/gaussian.cu:491-574,650-675):
// 0.5 * lambda * (1 - alpha) * sum_c coeff_c * mu_c * (mu_c - 2 u_c) / sigma_tc^2, mu = 0 for pure-noise samples;
// sigma_tc = std_dev_t[t C + c] with per-step standard deviations (gaussian.cu:665-668), std_dev[c] otherwise.
template <int C>
__device__ __forceinline__ float lr_cost(const mppi_sampler_params& s, const float* u, const float* mu, bool pure,
                                         float lambda, float alpha, int t)
{
  float cost = 0.0f;
  if (s.std_dev_t == nullptr) {
#pragma unroll
    for (int i = 0; i < C; ++i) {
      const float m = pure ? 0.0f : mu[i];
      // coeff / sigma^2 is uniform and hoisted out of the horizon loop
      cost += m * (m - 2.0f * u[i]) * (s.control_cost_coeff[i] / (s.std_dev[i] * s.std_dev[i]));
    }
  } else {
    // t is uniform across the lanes that evaluate one step: scalar loads of the step's row
    const float* sd = s.std_dev_t + t * C;
#pragma unroll
    for (int i = 0; i < C; ++i) {
      const float m = pure ? 0.0f : mu[i];
      cost += m * (m - 2.0f * u[i]) * (s.control_cost_coeff[i] / (sd[i] * sd[i]));
    }
  }
  return 0.5f * lambda * (1.0f - alpha) * cost;
}

// Standard deviation of control c at step t before the iteration's decay: the per-step table or std_dev.
MPPI_HD float step_std_dev(const mppi_sampler_params& s, int t, int c, int C)
{
  return s.std_dev_t ? s.std_dev_t[t * C + c] : s.std_dev[c];
}

}  // namespace mppi_amd
