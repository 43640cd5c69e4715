// Two-system rollouts (Tube-MPPI, RMPPI): for every sample index k, system 0 and system 1 are rolled out together
// over the horizon by one workgroup lane, as the reference's rollout kernels with blockDim.z = 2 do
// (core/mppi_common.cu:35-153 for Tube-MPPI, controllers/Tube-MPPI/tube_mppi_controller.cu:158-236;
// rolloutRMPPIKernel / rolloutRMPPIDynamicsKernel + rolloutRMPPICostKernel, core/rmppi_kernels.cu:359-798).
//
// One workgroup of TS_THREADS threads takes R samples (64, 32 or 16) of both distributions:
//   - both control tiles [R][T C] and both means are staged in LDS (rows padded to T C + 1 floats), and for RMPPI the
//     feedback gains [T][S][C];
//   - the producer wavefront (one sample per lane) carries the two state chains: per step it forms the real system's
//     feedback fb_t = K_t (x_r - x_n) (RMPPI), constrains both controls, writes them back into the tiles (the applied
//     controls the weighted update averages), steps both systems and hands their outputs (and fb_t) to the consumers
//     through a double-buffered LDS chunk;
//   - the 15 consumer wavefronts evaluate the running, likelihood-ratio and feedback costs of the previous chunk,
//     thread (w, lane) taking sample `lane` and the steps t0 + w, t0 + w + 15, ...;
//   - wavefront 0 adds the consumer sums in a fixed order (deterministic) and forms the two costs; the tiles go back to
//     the sample buffer with 16-byte stores.
// Per sample-step the kernel reads and writes 2 C floats of each system (16 C bytes) -- the HBM traffic of the
// reference's two rollouts -- and its VALU work is two dynamics steps and two cost evaluations.
//
// The fused form (FUSED, Gaussian samples) draws both tiles in LDS from the same normals around each system's mean
// (mppi_sample_gaussian's numbering and transform with distribution 0's parameters, as the two sampler launches it
// replaces) and ends like the single-system fused iteration: per system, the tile's local baseline, weights and FP64
// sums and the weighted column sums of the applied controls as a Partial and a partial row, which combine_kernel
// merges into the new mean. The applied controls and the costs are still written out.
#include <algorithm>
#include <type_traits>

#include "iteration_core.hpp"
#include "kernels.hpp"
#include "models.hpp"
#include "rollout_core.hpp"
#include "tile.hpp"

namespace mppi_amd {

constexpr int TS_THREADS = RC_THREADS;
constexpr int TS_CONSUMERS = TS_THREADS / 64 - 1;
constexpr int TS_ACC = 4;  // consumer accumulators per lane (Tube uses 2)

struct TwoSystemArgs {
  float vft;  // RMPPI: value-function threshold
  const float* gains;
};

// LDS floats for R rows and `chunk` steps per hand-off: tiles, means, gains (RMPPI), output chunks [2][chunk][64][2][O|1],
// feedback chunks [2][chunk][64][C] (RMPPI), consumer partial sums [TS_ACC][15][64], terminal costs [2][64].
template <int S, int C, int O>
static size_t ts_lds_floats(int T, int R, int chunk, bool rmppi)
{
  const size_t TC = (size_t)T * C;
  const size_t tiles = (2 * (size_t)R * (TC + 1) + 3) & ~(size_t)3;
  const size_t means = 2 * ((TC + 3) & ~(size_t)3);
  const size_t gains = rmppi ? (((size_t)T * S * C + 3) & ~(size_t)3) : 0;
  const size_t outs = 2 * (size_t)chunk * 64 * 2 * (O | 1);
  const size_t fbs = rmppi ? 2 * (size_t)chunk * 64 * C : 0;
  return tiles + means + gains + outs + fbs + (size_t)TS_ACC * TS_CONSUMERS * 64 + 2 * 64 + 3 * 64;
}

// Both systems' sample rows k0 .. k0 + nrows - 1 from the same normals, around each system's mean mu_d (LDS, [TC]):
// tile_d[r ROW + col] = mu_t if k == 0 or t < stride, s_tc eps if k >= pstart, mu_t + s_tc eps otherwise -- the
// transform of gaussian_tile_fill (unconstrained: the producer constrains after the feedback), evaluated once per
// normal for both tiles.
template <int C, int NT>
__device__ __forceinline__ void two_tile_fill(const mppi_sampler_params& sp, const IterArgs a, int k0, int nrows,
                                              int TC, int ROW, int pstart, const float* mu0, const float* mu1,
                                              float* tile0, float* tile1)
{
  const int tid = threadIdx.x;
  const float decay = powf(sp.std_dev_decay, (float)a.iteration);
  const int64_t base = ((int64_t)k0 * TC) >> 2;
  const int n = nrows * TC;
  for (int e0 = tid * 4; e0 < n; e0 += NT * 4) {
    const Normal4 z = normals4(sp.seed, a.rng_offset, (uint64_t)(base + (e0 >> 2)));
    int r = e0 / TC;
    int col = e0 - r * TC;
    int t = col / C;
    int c = col - t * C;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (e0 + j < n) {
        const int k = k0 + r;
        const float sd = decay * step_std_dev(sp, t, c, C);
        const float m0 = mu0[col], m1 = mu1[col];
        float v0, v1;
        if (k == 0 || t < a.stride) {
          v0 = m0;
          v1 = m1;
        } else if (k >= pstart) {
          v0 = sd * z.z[j];
          v1 = v0;
        } else {
          v0 = m0 + sd * z.z[j];
          v1 = m1 + sd * z.z[j];
        }
        tile0[r * ROW + col] = v0;
        tile1[r * ROW + col] = v1;
      }
      ++col;
      if (++c == C) {
        c = 0;
        ++t;
      }
      if (col == TC) {
        col = 0;
        t = 0;
        c = 0;
        ++r;
      }
    }
  }
}

// computeFeedbackCost (sampling_distributions/gaussian/gaussian.cu:577-630): 0.5 lambda (1 - alpha) sum_c coeff_c
// fb_c^2 / sigma_c^2 with the distribution's standard deviations (std_dev, or the step's row of std_dev_t).
template <int C>
MPPI_HD float feedback_cost(const mppi_sampler_params& s, const float* fb, float lambda, float alpha, int t)
{
  float cost = 0.0f;
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const float sd = s.std_dev_t ? s.std_dev_t[t * C + i] : s.std_dev[i];
    cost += s.control_cost_coeff[i] * (fb[i] * fb[i]) / (sd * sd);
  }
  return 0.5f * lambda * (1.0f - alpha) * cost;
}

// NOM = -1: Tube (two independent systems); NOM = 0 or 1: RMPPI with system NOM nominal and 1 - NOM real (compile-time,
// so that the two state chains stay in registers).
template <int MODEL, int COST, int NOM, bool FUSED>
__global__ void __launch_bounds__(TS_THREADS) two_system_rollout_kernel(
    mppi_dynamics_params dyn, mppi_cost_params cost, mppi_sampler_params sp0, mppi_sampler_params sp1,
    TwoSystemArgs ta, int K, int T, float dt, float lambda, float alpha, const float* __restrict__ x0,
    const float* __restrict__ mean, float* __restrict__ samples, float* __restrict__ costs, int pstart, int R,
    int chunk, IterRef itr, Partial* __restrict__ parts, float* __restrict__ psums, int gstride)
{
  using M = BuiltinModel<MODEL>;
  using CF = BuiltinCost<COST, M::O>;
  constexpr int S = M::S, C = M::C, O = M::O, OP = O | 1;
  constexpr bool RMPPI = NOM >= 0;
  constexpr int nom = RMPPI ? NOM : 1, real = 1 - nom;
  extern __shared__ float lds[];
  const int TC = T * C, ROW = TC + 1, TC4 = (TC + 3) & ~3;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int k0 = blockIdx.x * R;
  const int nrows = min(R, K - k0);
  float* tile = lds;                                             // [2][R][ROW]
  float* mu = tile + ((2 * R * ROW + 3) & ~3);                   // [2][TC4]
  float* gl = mu + 2 * TC4;                                      // [T][S][C]
  float* ybuf = gl + (RMPPI ? ((T * S * C + 3) & ~3) : 0);       // [2][chunk][64][2][OP]
  float* fbuf = ybuf + 2 * chunk * 64 * 2 * OP;                  // [2][chunk][64][C]
  float* part = fbuf + (RMPPI ? 2 * chunk * 64 * C : 0);         // [TS_ACC][15][64]
  float* term = part + TS_ACC * TS_CONSUMERS * 64;               // [2][64]
  float* cl = term + 2 * 64;                                     // [2][64] costs (FUSED)
  float* wl = cl + 2 * 64;                                       // [64] local weights (FUSED)

  if constexpr (!FUSED) {
    for (int d = 0; d < 2; ++d)
      tile_load<TS_THREADS>(samples + ((size_t)d * K + k0) * TC, tile + d * R * ROW, nrows, TC, ROW, tid);
  }
  for (int j = tid; j < 2 * TC; j += TS_THREADS) {
    const int d = j >= TC ? 1 : 0;
    mu[d * TC4 + (j - d * TC)] = mean[j];
  }
  if constexpr (RMPPI) {
    for (int j = tid; j < T * S * C; j += TS_THREADS) gl[j] = ta.gains ? ta.gains[j] : 0.0f;
  }
  if constexpr (FUSED) {
    __syncthreads();
    two_tile_fill<C, TS_THREADS>(sp0, load_iter_args(itr), k0, nrows, TC, ROW, pstart, mu, mu + TC4, tile,
                                 tile + R * ROW);
  }
  bool deadband = false;
#pragma unroll
  for (int c = 0; c < C; ++c) deadband |= dyn.control_deadband[c] != 0.0f;
  __syncthreads();

  // lanes beyond the tile's rows repeat its last row: they compute the same values as its lane and store them to the
  // same words
  const int lrow = min(lane, R - 1);
  float* row0 = tile + lrow * ROW;
  float* row1 = tile + R * ROW + lrow * ROW;
  const int nch = (T + chunk - 1) / chunk;
  // Tube with at most 32 samples per workgroup: the producer splits its wavefront by system, lane (sd, srow) =
  // (lane / 32, lane % 32) carrying system sd of sample srow alone in x[0] and storing its outputs to the slots of
  // sample srow (half the instructions per step). Otherwise each lane carries both systems of its sample (RMPPI's
  // coupled feedback is cheaper within one lane than through the cross-half exchange of the split form). Either way
  // slot r of the chunk buffers and the terminal costs belongs to sample row r.
  const bool split = !RMPPI && R <= 32;
  const int sd = lane >> 5;
  const int srow = min(lane & 31, R - 1);
  float* rows = tile + sd * R * ROW + srow * ROW;

  float x[2][S], y[2][O];
  if (wave == 0) {
    if (split) {
#pragma unroll
      for (int i = 0; i < S; ++i) x[0][i] = x0[sd * S + i];
      M::output(dyn, x[0], y[0]);
    } else {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
#pragma unroll
        for (int i = 0; i < S; ++i) x[d][i] = x0[d * S + i];
        M::output(dyn, x[d], y[d]);
      }
    }
  }
  float acc[TS_ACC];
#pragma unroll
  for (int q = 0; q < TS_ACC; ++q) acc[q] = 0.0f;
  const bool pure = (k0 + lrow) >= pstart;
  // the producer's operands of the next step (controls of both systems, RMPPI gains), loaded one step ahead so that
  // their LDS latency passes while the current step is computed; step t + 1's words are not written before step t + 1
  constexpr int GN = RMPPI ? S * C : 1;
  float un[2][C], gn[GN];
  auto load_gains = [&](int t) {
    if constexpr (RMPPI) {
#pragma unroll
      for (int j = 0; j < GN; ++j) gn[j] = gl[t * S * C + j];
    }
  };
  auto load_step = [&](int t) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      un[0][c] = row0[t * C + c];
      un[1][c] = row1[t * C + c];
    }
    load_gains(t);
  };
  auto load_step_split = [&](int t) {
#pragma unroll
    for (int c = 0; c < C; ++c) un[0][c] = rows[t * C + c];
  };
  if (wave == 0) {
    if (split) load_step_split(0);
    else load_step(0);
  }
  // Before a chunk's steps every load still in flight is waited for (the chunk's first prefetch among them): the
  // parameter loads issued ahead of the chunk loop are then known to be complete inside the step loop, which otherwise
  // waits for them -- and with them for the step's own prefetch -- at every step.
  auto settle_loads = [] { __builtin_amdgcn_s_waitcnt(0); };
  // the producer's steps t0 .. t1 - 1, with the constraint form (deadband or plain clamp) fixed outside the loop
  auto produce = [&](auto deadband_tag, int t0, int t1, float* yb, float* fbb) {
    settle_loads();
    for (int t = t0; t < t1; ++t) {
      float u[2][C], g[GN], fbs[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        u[0][c] = un[0][c];
        u[1][c] = un[1][c];
      }
#pragma unroll
      for (int j = 0; j < GN; ++j) g[j] = gn[j];
      load_step(min(t + 1, T - 1));  // at the last step a harmless reload
      if constexpr (RMPPI) {
        // fb_t = K_t (x_r - x_n) with the gains of step t
        float fb[C];
#pragma unroll
        for (int c = 0; c < C; ++c) fb[c] = 0.0f;
#pragma unroll
        for (int i = 0; i < S; ++i) {
          const float e = x[real][i] - x[nom][i];
#pragma unroll
          for (int c = 0; c < C; ++c) fb[c] += g[i * C + c] * e;
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
          u[real][c] += fb[c];
          fbs[c] = fb[c];
        }
      }
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        if constexpr (decltype(deadband_tag)::value) enforce_constraints<C>(dyn, u[d]);
        else clamp_controls<C>(dyn, u[d]);
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        row0[t * C + c] = u[0][c];
        row1[t * C + c] = u[1][c];
      }
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        float xd[S], xn[S];
        M::deriv(dyn, x[d], u[d], xd);
        M::update(dyn, x[d], xn, xd, dt);
        M::output(dyn, xn, y[d]);
#pragma unroll
        for (int i = 0; i < O; ++i) yb[(t - t0) * 64 * 2 * OP + d * OP + i] = y[d][i];
#pragma unroll
        for (int i = 0; i < S; ++i) x[d][i] = xn[i];
      }
      // the feedback's hand-off last: its registers are then not reused while the store is in flight
      if constexpr (RMPPI) {
#pragma unroll
        for (int c = 0; c < C; ++c) fbb[(t - t0) * 64 * C + c] = fbs[c];
      }
    }
  };
  // the split form (Tube): system sd of sample srow per lane, yb at the sample's slot
  auto produce_split = [&](auto deadband_tag, int t0, int t1, float* yb) {
    settle_loads();
    for (int t = t0; t < t1; ++t) {
      float u[C];
#pragma unroll
      for (int c = 0; c < C; ++c) u[c] = un[0][c];
      load_step_split(min(t + 1, T - 1));  // at the last step a harmless reload
      if constexpr (decltype(deadband_tag)::value) enforce_constraints<C>(dyn, u);
      else clamp_controls<C>(dyn, u);
#pragma unroll
      for (int c = 0; c < C; ++c) rows[t * C + c] = u[c];
      float xd[S], xn[S];
      M::deriv(dyn, x[0], u, xd);
      M::update(dyn, x[0], xn, xd, dt);
      M::output(dyn, xn, y[0]);
#pragma unroll
      for (int i = 0; i < O; ++i) yb[(t - t0) * 64 * 2 * OP + sd * OP + i] = y[0][i];
#pragma unroll
      for (int i = 0; i < S; ++i) x[0][i] = xn[i];
    }
  };

  for (int it = 0; it <= nch; ++it) {
    if (wave == 0 && it < nch) {
      const int t0 = it * chunk, t1 = min(T, t0 + chunk);
      float* ybs = ybuf + (it & 1) * chunk * 64 * 2 * OP;
      float* fbs = fbuf + (it & 1) * chunk * 64 * C;
      if (split) {
        if (deadband) produce_split(std::true_type{}, t0, t1, ybs + srow * 2 * OP);
        else produce_split(std::false_type{}, t0, t1, ybs + srow * 2 * OP);
      } else {
        if (deadband) produce(std::true_type{}, t0, t1, ybs + lane * 2 * OP, fbs + lane * C);
        else produce(std::false_type{}, t0, t1, ybs + lane * 2 * OP, fbs + lane * C);
      }
      if (it == nch - 1) {
        if (split) {
          term[sd * 64 + srow] = CF::terminal(cost, y[0]);
        } else {
          term[lane] = CF::terminal(cost, y[0]);
          term[64 + lane] = CF::terminal(cost, y[1]);
        }
      }
    } else if (wave > 0 && it > 0) {
      const int t0 = (it - 1) * chunk, t1 = min(T, t0 + chunk);
      const float* yb = ybuf + ((it - 1) & 1) * chunk * 64 * 2 * OP + lrow * 2 * OP;
      const float* fbb = fbuf + ((it - 1) & 1) * chunk * 64 * C + lrow * C;
      for (int t = t0 + wave - 1; t < t1; t += TS_CONSUMERS) {
        const float* y0 = yb + (t - t0) * 64 * 2 * OP;
        const float* u0 = row0 + t * C;
        const float* u1 = row1 + t * C;
        if constexpr (!RMPPI) {
          acc[0] += CF::running(cost, y0, u0, t) + lr_cost<C>(sp0, u0, mu + t * C, pure, lambda, alpha, t);
          acc[1] += CF::running(cost, y0 + OP, u1, t) + lr_cost<C>(sp1, u1, mu + TC4 + t * C, pure, lambda, alpha, t);
        } else {
          const float* yr = y0 + real * OP;
          const float* yn = y0 + nom * OP;
          const float* ur = real == 0 ? u0 : u1;
          const float* un = real == 0 ? u1 : u0;
          const mppi_sampler_params& spr = real == 0 ? sp0 : sp1;
          const mppi_sampler_params& spn = real == 0 ? sp1 : sp0;  // compile-time selections
          const float rc_r = CF::running(cost, yr, ur, t);
          acc[0] += rc_r + lr_cost<C>(spr, ur, mu + real * TC4 + t * C, pure, lambda, alpha, t);
          acc[1] += rc_r + feedback_cost<C>(spr, fbb + (t - t0) * 64 * C, lambda, alpha, t);
          acc[2] += CF::running(cost, yn, un, t);
          acc[3] += lr_cost<C>(spn, un, mu + nom * TC4 + t * C, pure, lambda, alpha, t);
        }
      }
    }
    __syncthreads();
  }
  if (wave > 0) {
#pragma unroll
    for (int q = 0; q < TS_ACC; ++q) part[(q * TS_CONSUMERS + wave - 1) * 64 + lane] = acc[q];
  }
  if constexpr (!FUSED) {
    for (int d = 0; d < 2; ++d)
      tile_store<TS_THREADS>(samples + ((size_t)d * K + k0) * TC, tile + d * R * ROW, nrows, TC, ROW, tid);
  }
  __syncthreads();
  if (wave == 0) {
    float s[TS_ACC];
#pragma unroll
    for (int q = 0; q < TS_ACC; ++q) {
      s[q] = 0.0f;
#pragma unroll
      for (int w = 0; w < TS_CONSUMERS; ++w) s[q] += part[(q * TS_CONSUMERS + w) * 64 + lane];
    }
    const float inv_t = 1.0f / (float)T;
    float c0, c1;
    if constexpr (!RMPPI) {
      c0 = s[0] / (float)T + term[lrow] / (float)T;
      c1 = s[1] / (float)T + term[64 + lrow] / (float)T;
    } else {
      const float tr = term[real * 64 + lrow], tn = term[nom * 64 + lrow];
      const float cost_r = (s[0] + tr) * inv_t;
      const float track_r = (s[1] + tr) * inv_t;
      const float state_n = (s[2] + tn) * inv_t;
      const float lr_n = s[3] * inv_t;
      const float cost_n = 0.5f * state_n + 0.5f * fmaxf(fminf(track_r, ta.vft), state_n) + lr_n;
      c0 = real == 0 ? cost_r : cost_n;
      c1 = real == 0 ? cost_n : cost_r;
    }
    if constexpr (FUSED) {
      cl[lane] = c0;
      cl[64 + lane] = c1;
    } else if (lane < nrows) {
      costs[k0 + lane] = c0;
      costs[K + k0 + lane] = c1;
    }
  }
  if constexpr (FUSED) {
    // per system: costs and applied controls out, the tile's partial of the weighted update
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      tile_partials<TS_THREADS, true, false>(cl + 64 * d, cl + 64 * d, wl, tile + d * R * ROW, ROW, TC, k0, nrows,
                                             blockIdx.x, lambda, costs + (size_t)d * K, parts + (size_t)d * gstride,
                                             psums + (size_t)d * gstride * TC, samples + (size_t)d * K * TC);
      __syncthreads();
    }
  }
}

// Samples per workgroup: the smallest of 64, 32, 16 whose grid still fits the CUs once (a small K spreads over more
// CUs), then halved until the tiles and buffers fit the LDS; the chunk is the first of 15 (one step per cost
// wavefront), 8 and 4 that fits.
template <int MODEL, int COST, int NOM, bool FUSED>
static int launch_two_system_t(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                               const mppi_sampler_params& sp0, const mppi_sampler_params& sp1, const TwoSystemArgs& ta,
                               int K, int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                               float* samples, float* costs, hipStream_t s, IterRef itr = IterRef{nullptr, {0, 0, 0}},
                               Partial* parts = nullptr, float* psums = nullptr, int gstride = 0, int* G_out = nullptr)
{
  using M = BuiltinModel<MODEL>;
  constexpr bool RMPPI = NOM >= 0;
  int cus = 0, dev = 0;
  MPPI_HIP_CHECK(hipGetDevice(&dev));
  MPPI_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int R = fused_rows_for(K, cus > 0 ? cus : 256), chunk = 0;
  const size_t limit = LDS_LIMIT / sizeof(float);
  for (;;) {
    for (int ch : {TS_CONSUMERS, 8, 4})
      if (chunk == 0 && ts_lds_floats<M::S, M::C, M::O>(T, R, ch, RMPPI) <= limit) chunk = ch;
    if (chunk != 0 || R == 16) break;
    R /= 2;
  }
  if (chunk == 0) return fail(MPPI_E_SIZE, "two-system rollout: horizon too long for the LDS control tiles");
  const size_t lds = sizeof(float) * ts_lds_floats<M::S, M::C, M::O>(T, R, chunk, RMPPI);
  MPPI_CHECK(allow_max_dynamic_lds((const void*)two_system_rollout_kernel<MODEL, COST, NOM, FUSED>));
  const int pstart = pure_noise_start(K, sp0.pure_noise_trajectories_percentage);
  const int grid = (K + R - 1) / R;
  if (FUSED && grid > gstride) return fail(MPPI_E_SIZE, "two-system iteration: partial workspace too small");
  if (G_out != nullptr) *G_out = grid;
  two_system_rollout_kernel<MODEL, COST, NOM, FUSED><<<grid, TS_THREADS, lds, s>>>(
      dyn, cost, sp0, sp1, ta, K, T, dt, lambda, alpha, x0, mean, samples, costs, pstart, R, chunk, itr, parts, psums,
      gstride);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// ====================================================================================== RMPPI candidate evaluation
// initEvalKernel (core/rmppi_kernels.cu:12-358, the fused form of initEvalDynKernel + initEvalCostKernel): candidate i
// (workgroup row blockIdx.y) rolls out samples j < M of the sample buffer from its state states[i], reading sample j's
// controls shifted by the candidate's stride, u_t = samples[j][min(t + stride_i, T - 1)], constrained; the running cost
// plus the likelihood-ratio cost against the mean at t (the sample index for the pure-noise rule is i M + j);
// cost = sum / T + terminal / T into costs[i M + j]. The shifted tile is staged in LDS and rolled out by the pipelined
// rollout of rollout_core.hpp.
template <int MODEL, int COST>
__global__ void __launch_bounds__(rollout_threads<MODEL>()) candidate_eval_kernel(
    mppi_dynamics_params dyn, mppi_cost_params cost, mppi_sampler_params sp, int M, int T, float dt, float lambda,
    float alpha, const float* __restrict__ states, const int* __restrict__ strides, const float* __restrict__ mean,
    const float* __restrict__ samples, float* __restrict__ costs, int pstart, int chunk)
{
  constexpr int C = Model<MODEL>::C, S = Model<MODEL>::S;
  constexpr int NT = rollout_threads<MODEL>();
  extern __shared__ float lds[];
  const int TC = T * C, ROW = TC + 1;
  float* tile = lds;
  float* mu = tile + RC_LANES * ROW;
  float* x0_l = mu + ((TC + 3) & ~3);
  float* scratch = x0_l + ((S + 3) & ~3);
  const int tid = threadIdx.x;
  const int cand = blockIdx.y;
  const int k0 = blockIdx.x * RC_LANES;
  const int nrows = min(RC_LANES, M - k0);
  const int stride = strides[cand];
  bool deadband = false;
#pragma unroll
  for (int c = 0; c < C; ++c) deadband |= dyn.control_deadband[c] != 0.0f;
  for (int e = tid; e < nrows * TC; e += NT) {
    const int r = e / TC, col = e - r * TC;
    const int tt = col / C, c = col - tt * C;
    const float v = samples[(size_t)(k0 + r) * TC + min(tt + stride, T - 1) * C + c];
    tile[r * ROW + col] = constrain_control(dyn, c, v, deadband);
  }
  for (int j = tid; j < TC; j += NT) mu[j] = mean[j];
  if (tid < S) x0_l[tid] = states[cand * S + tid];
  __syncthreads();
  float* cost_l;
  float* raw_l;
  rollout_tile<MODEL, COST>(dyn, cost, sp, T, dt, lambda, alpha, x0_l, tile, ROW, mu, pstart, cand * M + k0, scratch,
                            chunk, cost_l, raw_l, nullptr, RC_LANES);
  if (tid < nrows) costs[cand * M + k0 + tid] = cost_l[tid];
}

template <int MODEL, int COST>
static int launch_candidate_eval_t(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                                   const mppi_sampler_params& sp, int N, int M, int K, int T, float dt, float lambda,
                                   float alpha, const float* states, const int* strides, const float* mean,
                                   const float* samples, float* costs, hipStream_t s)
{
  constexpr int C = Model<MODEL>::C, S = Model<MODEL>::S, O = Model<MODEL>::O;
  const size_t base = (size_t)RC_LANES * (T * C + 1) + (size_t)((T * C + 3) & ~3) + (size_t)((S + 3) & ~3);
  const size_t limit = LDS_LIMIT / sizeof(float);
  const int chunk = base < limit ? rollout_chunk_for<MODEL, O>(limit - base) : 0;
  if (chunk == 0) return fail(MPPI_E_SIZE, "candidate evaluation: horizon too long for the LDS control tile");
  const size_t lds = sizeof(float) * (base + rollout_scratch_for<MODEL, O>(chunk));
  MPPI_CHECK(allow_max_dynamic_lds((const void*)candidate_eval_kernel<MODEL, COST>));
  const int pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  const dim3 grid((M + RC_LANES - 1) / RC_LANES, N);
  candidate_eval_kernel<MODEL, COST><<<grid, rollout_threads<MODEL>(), lds, s>>>(
      dyn, cost, sp, M, T, dt, lambda, alpha, states, strides, mean, samples, costs, pstart, chunk);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

int launch_candidate_eval(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                          const mppi_sampler_params& sp, int N, int M, int K, int T, float dt, float lambda,
                          float alpha, const float* states, const int* strides, const float* mean,
                          const float* samples, float* costs, hipStream_t s)
{
  if (N < 1 || M < 1 || M > K) return fail(MPPI_E_INVALID, "candidate evaluation: counts out of range");
  if (dyn.model == MPPI_MODEL_RACER_DUBINS_NN && dyn.nn_precision == MPPI_NN_BF16)
    return fail(MPPI_E_UNSUPPORTED, "candidate evaluation: the BF16 network model is not supported");
#define MPPI_CE_CASE(MODEL, COST)                                                                                  \
  if (dyn.model == MODEL && cost.cost == COST)                                                                    \
    return launch_candidate_eval_t<MODEL, COST>(dyn, cost, sp, N, M, K, T, dt, lambda, alpha, states, strides,   \
                                                mean, samples, costs, s);
  MPPI_CE_CASE(MPPI_MODEL_CARTPOLE, MPPI_COST_CARTPOLE_QUADRATIC)
  MPPI_CE_CASE(MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_QUADRATIC_MAP)
  MPPI_CE_CASE(MPPI_MODEL_RACER_DUBINS, MPPI_COST_QUADRATIC_MAP)
  MPPI_CE_CASE(MPPI_MODEL_RACER_DUBINS_NN, MPPI_COST_QUADRATIC_MAP)
#undef MPPI_CE_CASE
  return fail(MPPI_E_UNSUPPORTED, "candidate evaluation: model/cost combination is not compiled into the library");
}

int launch_two_system(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp0,
                      const mppi_sampler_params& sp1, const mppi_two_system_params& tp, int K, int T, float dt,
                      float lambda, float alpha, const float* x0, const float* mean, float* samples, float* costs,
                      hipStream_t s)
{
  if (tp.mode != MPPI_TWO_SYSTEM_TUBE && tp.mode != MPPI_TWO_SYSTEM_RMPPI)
    return fail(MPPI_E_INVALID, "two-system rollout: mode must be MPPI_TWO_SYSTEM_TUBE or MPPI_TWO_SYSTEM_RMPPI");
  if (tp.nominal_index != 0 && tp.nominal_index != 1)
    return fail(MPPI_E_INVALID, "two-system rollout: nominal_index must be 0 or 1");
  if (dyn.model == MPPI_MODEL_RACER_DUBINS_NN && dyn.nn_precision == MPPI_NN_BF16)
    return fail(MPPI_E_UNSUPPORTED, "two-system rollout: the BF16 network model is not supported (use MPPI_NN_FP32)");
  const TwoSystemArgs ta{tp.value_function_threshold, tp.feedback_gains};
  const int nom = tp.mode == MPPI_TWO_SYSTEM_RMPPI ? tp.nominal_index : -1;
#define MPPI_TS_CASE(MODEL, COST)                                                                                   \
  if (dyn.model == MODEL && cost.cost == COST) {                                                                   \
    if (nom < 0)                                                                                                   \
      return launch_two_system_t<MODEL, COST, -1, false>(dyn, cost, sp0, sp1, ta, K, T, dt, lambda, alpha, x0,     \
                                                         mean, samples, costs, s);                                 \
    if (nom == 0)                                                                                                  \
      return launch_two_system_t<MODEL, COST, 0, false>(dyn, cost, sp0, sp1, ta, K, T, dt, lambda, alpha, x0,      \
                                                        mean, samples, costs, s);                                  \
    return launch_two_system_t<MODEL, COST, 1, false>(dyn, cost, sp0, sp1, ta, K, T, dt, lambda, alpha, x0, mean,  \
                                                      samples, costs, s);                                          \
  }
  MPPI_TS_CASE(MPPI_MODEL_CARTPOLE, MPPI_COST_CARTPOLE_QUADRATIC)
  MPPI_TS_CASE(MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_QUADRATIC_MAP)
  MPPI_TS_CASE(MPPI_MODEL_RACER_DUBINS, MPPI_COST_QUADRATIC_MAP)
  MPPI_TS_CASE(MPPI_MODEL_RACER_DUBINS_NN, MPPI_COST_QUADRATIC_MAP)
#undef MPPI_TS_CASE
  return fail(MPPI_E_UNSUPPORTED, "two-system rollout: model/cost combination is not compiled into the kernel library");
}

int launch_two_system_fused(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                            const mppi_sampler_params& sp, const mppi_two_system_params& tp, int K, int T, float dt,
                            float lambda, float alpha, const float* x0, const float* mean, float* samples,
                            float* costs, IterRef itr, Partial* parts, float* psums, int gstride, int* G,
                            hipStream_t s)
{
  if (tp.mode != MPPI_TWO_SYSTEM_TUBE && tp.mode != MPPI_TWO_SYSTEM_RMPPI)
    return fail(MPPI_E_INVALID, "two-system iteration: mode must be MPPI_TWO_SYSTEM_TUBE or MPPI_TWO_SYSTEM_RMPPI");
  if (tp.nominal_index != 0 && tp.nominal_index != 1)
    return fail(MPPI_E_INVALID, "two-system iteration: nominal_index must be 0 or 1");
  if (sp.kind != MPPI_SAMPLER_GAUSSIAN) return fail(MPPI_E_INVALID, "two-system iteration: Gaussian samples only");
  if (dyn.model == MPPI_MODEL_RACER_DUBINS_NN && dyn.nn_precision == MPPI_NN_BF16)
    return fail(MPPI_E_UNSUPPORTED, "two-system iteration: the BF16 network model is not supported (use MPPI_NN_FP32)");
  const TwoSystemArgs ta{tp.value_function_threshold, tp.feedback_gains};
  const int nom = tp.mode == MPPI_TWO_SYSTEM_RMPPI ? tp.nominal_index : -1;
#define MPPI_TSF_CASE(MODEL, COST)                                                                                  \
  if (dyn.model == MODEL && cost.cost == COST) {                                                                   \
    if (nom < 0)                                                                                                   \
      return launch_two_system_t<MODEL, COST, -1, true>(dyn, cost, sp, sp, ta, K, T, dt, lambda, alpha, x0, mean,  \
                                                        samples, costs, s, itr, parts, psums, gstride, G);         \
    if (nom == 0)                                                                                                  \
      return launch_two_system_t<MODEL, COST, 0, true>(dyn, cost, sp, sp, ta, K, T, dt, lambda, alpha, x0, mean,   \
                                                       samples, costs, s, itr, parts, psums, gstride, G);          \
    return launch_two_system_t<MODEL, COST, 1, true>(dyn, cost, sp, sp, ta, K, T, dt, lambda, alpha, x0, mean,     \
                                                     samples, costs, s, itr, parts, psums, gstride, G);            \
  }
  MPPI_TSF_CASE(MPPI_MODEL_CARTPOLE, MPPI_COST_CARTPOLE_QUADRATIC)
  MPPI_TSF_CASE(MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_QUADRATIC_MAP)
  MPPI_TSF_CASE(MPPI_MODEL_RACER_DUBINS, MPPI_COST_QUADRATIC_MAP)
  MPPI_TSF_CASE(MPPI_MODEL_RACER_DUBINS_NN, MPPI_COST_QUADRATIC_MAP)
#undef MPPI_TSF_CASE
  return fail(MPPI_E_UNSUPPORTED, "two-system iteration: model/cost combination is not compiled into the library");
}

}  // namespace mppi_amd
