// Workgroup pieces of one fused MPPI iteration shared by the built-in controller kernel (fused.hip) and the kernels
// a user's translation unit instantiates for its own dynamics and cost (plugin_kernels.hpp):
//   gaussian_tile_fill  -- the workgroup's rows of Gaussian control samples (setGaussianControls,
//                          sampling_distributions/gaussian/gaussian.cu:17-277) drawn from Philox straight into the LDS
//                          control tile, constrained on the way in;
//   tile_partials       -- local baseline and exp-weights of the tile's costs (core/mppi_common.cu:938-1002 against
//                          the workgroup's own minimum), FP64 sums, and the weighted column sums S_g[T C]
//                          (core/mppi_common.cu:746-773 over the tile's rows), which combine_kernel merges.
#pragma once

#include <climits>

#include "fused.hpp"
#include "rollout_core.hpp"
#include "tile.hpp"

namespace mppi_amd {

__device__ __forceinline__ IterArgs load_iter_args(const IterRef& r) { return r.ptr ? *r.ptr : r.val; }

// Rows k0 .. k0 + nrows - 1 of the Gaussian samples into tile[r ROW + col] (col = t C + c):
//   u = mu_t                         if k == 0 or t < stride
//   u = mu_t + dmu_t + sig_t eps     with a proposal (prop_l: dmu [TC4], sig [TC4] in LDS)
//   u = s_tc eps                     if k >= pstart      (s_tc = std_dev_c decay^iteration, or the per-step
//                                                         std_dev_t[t][c] decay^iteration)
//   u = mu_t + s_tc eps              otherwise,
// then f(col, u) (the constraint). eps is normal e of the stream (seed, rng_offset) for element e = k T C + col, the
// numbering of mppi_sample_gaussian: the tile is a contiguous span starting at element k0 T C (a multiple of 4), so
// Philox block b covers tile elements 4 (b - base) .. 4 (b - base) + 3.
template <int C, int NT, bool PROP, class F>
__device__ __forceinline__ void gaussian_tile_fill(const mppi_sampler_params& sp, const IterArgs a, int k0, int nrows,
                                                   int TC, int ROW, int pstart, const float* mu, const float* prop_l,
                                                   int TC4, float* tile, F f)
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
        const float m = mu[col];
        float v;
        if (k == 0 || t < a.stride) v = m;
        else if (PROP) v = m + prop_l[col] + prop_l[TC4 + col] * z.z[j];
        else if (k >= pstart) v = decay * step_std_dev(sp, t, c, C) * z.z[j];
        else v = m + decay * step_std_dev(sp, t, c, C) * z.z[j];
        tile[r * ROW + col] = f(col, v);
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

// After the rollout (cost_l / raw_l from rollout_tile*): wavefront 0 forms the tile's local baseline (first strict
// minimum), the weights w_r = exp(-(c_r - b_g) / lambda) into wl[0..64), and the FP64 sums eta_g, sum w^2, sum w c
// (raw costs) into parts[g]; costs_out (optional) receives the costs. Then, after a barrier, the clamped tile is
// stored to `samples` when WRITE, and every column j of the tile is reduced to sums[g TC + j] = sum_r w_r u_rj in row
// order. With PROP, sample 0 (the warm start) is a probe: its cost is reported but it takes no part in the weights.
template <int NT, bool WRITE, bool PROP>
__device__ __forceinline__ void tile_partials(const float* cost_l, const float* raw_l, float* wl, const float* tile,
                                              int ROW, int TC, int k0, int nrows, int g, float lambda,
                                              float* __restrict__ costs_out, Partial* __restrict__ parts,
                                              float* __restrict__ sums, float* __restrict__ samples)
{
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  if (tid < RC_LANES) {
    const int k = k0 + tid;
    const float my_cost = cost_l[lane];
    if (lane < nrows && costs_out != nullptr) costs_out[k] = my_cost;
    const bool valid = lane < nrows && !(PROP && k == 0);
    // local weights against the workgroup's own baseline
    const float c = valid ? sanitize_cost(my_cost) : INFINITY;
    float b = c;
    int bi = valid ? k : INT_MAX;
    wave_argmin(b, bi);
    const bool fin = valid && c <= FLT_MAX;
    const float w = fin ? expf(-(1.0f / lambda) * (c - b)) : 0.0f;
    wl[lane] = w;
    const double e = wave_sum((double)w);
    const double e2 = wave_sum((double)w * (double)w);
    // weighted mean of the uncorrected costs (the proposal's achieved cost)
    const float raw = raw_l[lane];
    const double ec = wave_sum((fin && raw <= FLT_MAX) ? (double)w * (double)raw : 0.0);
    if (lane == 0) {
      Partial p;
      p.baseline = b;
      p.argmin = bi;
      p.eta = e;
      p.s2 = e2;
      p.wc = ec;
      parts[g] = p;
    }
  }
  __syncthreads();
  if (WRITE) tile_store<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid);
  for (int j = tid; j < TC; j += NT) {
    float acc = 0.0f;
    // in row order; unrolled so that the LDS reads of 8 rows are in flight together
#pragma unroll 8
    for (int r = 0; r < nrows; ++r) acc += wl[r] * tile[r * ROW + j];
    sums[(size_t)g * TC + j] = acc;
  }
}

// Samples per workgroup of a 1024-thread fused iteration: the smallest of 64, 32, 16 with ceil(K / rows) <= CU count
// (the state chain takes as long for 16 live lanes as for 64, while drawing the tile and reducing it scale with the
// samples); 64 when K exceeds 32 samples per CU.
inline int fused_rows_for(int K, int cus)
{
  int rows = RC_LANES;
  while (rows > 16 && (K + rows / 2 - 1) / (rows / 2) <= cus) rows /= 2;
  return rows;
}

}  // namespace mppi_amd
