// A model and cost that the engine does not ship, written the way a user of the plugin API writes them: a unicycle
// (state x, y, heading; controls speed, turn rate) whose output is (x, y, cos heading, sin heading), and a cost with
// a goal, a heading term, a time-weighted obstacle bump, a control term and a terminal term. The formulas are
// templates over the arithmetic type so that the GPU plugin (float) and the test's CPU reference (float and double)
// evaluate the same expressions; what the test checks is the engine's sampling, rollout pipeline, weights and
// combine around them.
#pragma once

#ifndef UNI_HD
#define UNI_HD
#endif

#include <cmath>

struct UnicycleParams {
  float speed_scale;
  float turn_scale;
};

struct UnicycleCostParams {
  float goal_x, goal_y, goal_heading;
  float q_pos, q_heading;
  float obs_x, obs_y, obs_radius, obs_weight, obs_time_gain;
  float control_weight;
  float terminal_weight;
};

// the test's control limits: speed in [-0.5, 2], turn rate in [-1.5, 1.5], a speed deadband of 0.05 snapping to 0
struct UnicycleSetup {
  UnicycleParams dyn;
  UnicycleCostParams cost;
  float control_min[2], control_max[2], control_deadband[2], zero_control[2];
};

UNI_HD inline float uni_cos(float a) { return cosf(a); }
UNI_HD inline float uni_sin(float a) { return sinf(a); }
UNI_HD inline float uni_exp(float a) { return expf(a); }
inline double uni_cos(double a) { return std::cos(a); }
inline double uni_sin(double a) { return std::sin(a); }
inline double uni_exp(double a) { return std::exp(a); }

template <class R>
UNI_HD inline void unicycle_deriv(const UnicycleParams& p, const R* x, const R* u, R* xd)
{
  const R v = u[0] * (R)p.speed_scale;
  xd[0] = v * uni_cos(x[2]);
  xd[1] = v * uni_sin(x[2]);
  xd[2] = u[1] * (R)p.turn_scale;
}

// Euler with the heading kept in [-pi, pi)
template <class R>
UNI_HD inline void unicycle_update(const R* x, R* xn, const R* xd, R dt)
{
  for (int i = 0; i < 3; ++i) xn[i] = x[i] + xd[i] * dt;
  const R pi = (R)3.14159265358979323846;
  if (xn[2] >= pi) xn[2] -= (R)2 * pi;
  else if (xn[2] < -pi) xn[2] += (R)2 * pi;
}

template <class R>
UNI_HD inline void unicycle_output(const R* x, R* y)
{
  y[0] = x[0];
  y[1] = x[1];
  y[2] = uni_cos(x[2]);
  y[3] = uni_sin(x[2]);
}

template <class R>
UNI_HD inline R unicycle_state_cost(const UnicycleCostParams& c, const R* y, int t)
{
  const R dx = y[0] - (R)c.goal_x, dy = y[1] - (R)c.goal_y;
  const R align = y[2] * uni_cos((R)c.goal_heading) + y[3] * uni_sin((R)c.goal_heading);
  const R ox = y[0] - (R)c.obs_x, oy = y[1] - (R)c.obs_y;
  const R bump = (R)c.obs_weight * uni_exp(-(ox * ox + oy * oy) / ((R)c.obs_radius * (R)c.obs_radius));
  return (R)c.q_pos * (dx * dx + dy * dy) + (R)c.q_heading * ((R)1 - align) +
         bump * ((R)1 + (R)c.obs_time_gain * (R)t);
}

template <class R>
UNI_HD inline R unicycle_control_cost(const UnicycleCostParams& c, const R* u, int)
{
  return (R)0.5 * (R)c.control_weight * (u[0] * u[0] + u[1] * u[1]);
}

template <class R>
UNI_HD inline R unicycle_terminal_cost(const UnicycleCostParams& c, const R* y)
{
  const R dx = y[0] - (R)c.goal_x, dy = y[1] - (R)c.goal_y;
  return (R)c.terminal_weight * (dx * dx + dy * dy);
}
