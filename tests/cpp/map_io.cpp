// Weight and map I/O through the C++ API: RacerDubinsNN::loadParams from an .npz network (argv[1]: a 9-64-64-4
// network whose dynamics_W / dynamics_b arrays hold consecutive ranges, so theta[i] must be i), QuadraticMapCost::
// loadTrackData from an .npz track map (argv[2]: the 10 x 20 m, 2 pixels-per-metre map of the reference's test maps),
// and updateTexture with the controller running: values of the same extent, uploaded in place, that add a constant c
// to the whole map move the baseline by map_coeff c and leave the control sequence unchanged (the weights do not
// move); the column-major form gives the same result bit for bit. Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mppi_amd/mppi.hpp"

using namespace mppi_amd;
constexpr int T_STEPS = 60;
using DYN = DoubleIntegratorDynamics;
using COST = QuadraticMapCost<DYN>;
using SAMPLER = sampling_distributions::GaussianDistributionC<DYN::CONTROL_DIM>;
using FB = NoFeedback<DYN, T_STEPS>;
using CTRL = VanillaMPPIController<DYN, COST, FB, T_STEPS, 2048, SAMPLER>;

static std::vector<float> blob_map(int w, int h)
{
  std::vector<float> m((size_t)w * h);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x)
      m[(size_t)y * w + x] = 30.0f * std::exp(-((x - 90.0f) * (x - 90.0f) + (y - 140.0f) * (y - 140.0f)) / 800.0f) +
                             20.0f * std::exp(-((x - 170.0f) * (x - 170.0f) + (y - 60.0f) * (y - 60.0f)) / 1500.0f);
  return m;
}

static void setup_cost(COST& cost, const std::vector<float>& map)
{
  COST::COST_PARAMS_T p;
  const float q[4] = {2.0f, 2.0f, 0.5f, 0.5f};
  for (int i = 0; i < 4; ++i) p.s_coeffs[i] = q[i];
  p.s_goal[0] = p.s_goal[1] = 2.0f;
  p.map_resolution = 0.05f;
  p.map_origin[0] = p.map_origin[1] = -6.4f;
  cost.setParams(p);
  cost.setCostMap(map, 256, 256);
}

int main(int argc, char** argv)
{
  int failures = 0;
  auto check = [&](bool ok, const char* what) {
    std::printf("[map_io] %s: %s\n", what, ok ? "ok" : "FAILED");
    if (!ok) failures++;
  };
  if (argc < 3) {
    std::printf("usage: map_io <network.npz> <track_map.npz>\n");
    return 2;
  }
  {
    RacerDubinsNN nn;
    nn.loadParams(argv[1]);
    const auto& th = nn.getTheta();
    bool ok = th.size() == 5060;
    for (size_t i = 0; ok && i < th.size(); ++i) ok = th[i] == (float)i;
    check(ok, "RacerDubinsNN::loadParams: theta[i] = i over the 5060 parameters");
  }
  {
    COST cost;
    cost.loadTrackData(argv[2]);
    const auto p = cost.getParams();
    std::printf("[map_io] track map %d x %d texels, origin (%.2f, %.2f), resolution %.3f\n", cost.getWidth(),
                cost.getHeight(), p.map_origin[0], p.map_origin[1], p.map_resolution);
    check(cost.getWidth() == 20 && cost.getHeight() == 40 && p.map_origin[0] == -5.0f && p.map_origin[1] == -10.0f &&
              p.map_resolution == 0.5f,
          "loadTrackData geometry");
    check(cost.costMapHost()[0] == 1.0f && cost.costMapHost()[799] == 800.0f, "loadTrackData values");
  }
  {
    const float dt = 0.02f, shift = 7.0f;
    const std::vector<float> map = blob_map(256, 256);
    std::vector<float> shifted(map), shifted_cm(map.size());
    for (float& v : shifted) v += shift;
    for (int x = 0; x < 256; ++x)
      for (int y = 0; y < 256; ++y) shifted_cm[(size_t)x * 256 + y] = shifted[(size_t)y * 256 + x];
    DYN model;
    SAMPLER sampler;
    FB fb(&model, dt);
    COST ca, cb, cc;
    setup_cost(ca, map);
    setup_cost(cb, map);
    setup_cost(cc, map);
    CTRL a(&model, &ca, &fb, &sampler, dt, 1, 1.0f, 0.0f), b(&model, &cb, &fb, &sampler, dt, 1, 1.0f, 0.0f),
        c(&model, &cc, &fb, &sampler, dt, 1, 1.0f, 0.0f);
    DYN::state_array x = DYN::state_array::Zero();
    for (int k = 0; k < 2; ++k) {
      a.computeControl(x, 1);
      b.computeControl(x, 1);
      c.computeControl(x, 1);
    }
    ca.updateTexture(shifted);
    cc.updateTexture(shifted_cm, true);
    a.computeControl(x, 1);
    b.computeControl(x, 1);
    c.computeControl(x, 1);
    const auto ua = a.getControlSeq(), ub = b.getControlSeq(), uc = c.getControlSeq();
    double du = 0.0;
    for (int i = 0; i < DYN::CONTROL_DIM * T_STEPS; ++i) du = std::fmax(du, std::fabs(ua.data()[i] - ub.data()[i]));
    std::printf("[map_io] shifted map: baseline %.6f vs %.6f + %.1f, control max |diff| %.3g\n", a.getBaselineCost(),
                b.getBaselineCost(), shift, du);
    check(std::fabs(a.getBaselineCost() - (b.getBaselineCost() + shift)) < 1e-4f * std::fabs(a.getBaselineCost()),
          "baseline moved by the shift");
    check(du < 1e-4, "control sequence unchanged");
    check(std::memcmp(ua.data(), uc.data(), sizeof(float) * DYN::CONTROL_DIM * T_STEPS) == 0,
          "column-major update equals the row-major one");
  }
  std::printf("[map_io] %s\n", failures == 0 ? "all checks passed" : "some checks failed");
  return failures == 0 ? 0 : 1;
}
