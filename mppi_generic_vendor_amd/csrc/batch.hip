// Many independent MPPI instances advanced by one persistent launch per iteration.
//
// The reference runs one VanillaMPPIController per system, each computeControl issuing its own sampling, rollout,
// weight and update launches (controllers/MPPI/mppi_controller.cu:152-237). Here all instances
// tiles form one work list, ordered longest horizon first, that a persistent kernel drains through an atomic
// counter. Per tile, one workgroup:
//   - keeps the instance's mean in LDS and draws the Gaussian samples chunk by chunk (Philox, the same element
//     numbering as mppi_sample_gaussian, so results equal the single-instance path) into a three-slot LDS ring;
//   - runs the state chain on wavefront 0 (samples_per_lane samples per lane) one chunk behind the generator while
//     the other wavefronts evaluate the running + likelihood-ratio costs of the previous chunk and write its clamped
//     samples to the instance's sample region ([T][R][C] per tile, coalesced);
//   - reduces the tile against its own baseline into (b_g, argmin_g, eta_g, sum w^2_g) and S_g[t][c].
// A combine launch (one workgroup per instance) merges the partials with the log-sum-exp rule of the controller path
// and writes each instance's new mean in place and its FP64 statistics. Both launches are captured in a hipGraph.
#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>
#include <numeric>
#include <vector>

#include "fused.hpp"
#include "host.hpp"
#include "kernels.hpp"
#include "models.hpp"

namespace mppi_amd {

constexpr int BATCH_MAX_K = 65536;
constexpr int BATCH_MAX_STATE = MPPI_MAX_STATE;

struct BatchInstance {
  mppi_dynamics_params dyn;
  mppi_cost_params cost;
  mppi_sampler_params sp;
  int32_t K, T, C, pstart;
  int32_t tiles;      // ceil(K / R) under the current launch shape
  int32_t part_base;  // index of the first Partial of the instance
  int64_t sum_base;   // float offset of S_g rows [tiles][T*C]
  int64_t u_base;     // float offset of the clamped samples [tiles][T][R][C]
  int32_t mean_base;  // float offset of the mean [T][C]
  int32_t cost_base;  // float offset of the costs [K]
  float dt, lambda, alpha;
  int32_t _pad;
};

// Cost-map queries of one consumer wavefront whose texel loads are issued together (per sample of a lane).
#ifndef BATCH_MAP_PREFETCH
#define BATCH_MAP_PREFETCH 1
#endif

// The final column sums run on up to SUM_WAVES wavefronts, each with a [64][8 C + 1] block of LDS.
constexpr int SUM_WAVES = 4;
__host__ __device__ constexpr int sum_block_floats(int C) { return 64 * (8 * C + 1); }

// LDS floats of one workgroup: the work-item slot (4 floats), then per tile the mean, the sample ring
// [4][chunk][R][C] and the outputs [2][chunk][R][O] (their room, at least the column sums' blocks, is reused by those
// blocks after the rollout), the consumer partial sums, terminal costs, costs and weights.
static size_t batch_lds_floats(int TCmax, int Cmax, int Omax, int NT, int R, int ch)
{
  const int ncons = NT / 64 - 1;
  const size_t rollout = 4ull * ch * R * Cmax + 2ull * ch * R * Omax;
  const size_t sums = (size_t)std::min(NT / 64, SUM_WAVES) * sum_block_floats(Cmax);
  return 4 + (size_t)((TCmax + 3) & ~3) + std::max(rollout, sums) + (size_t)ncons * R + 3ull * R;
}

// v from lane (lane ^ MASK) for MASK = 8 (DPP row rotation by 8 inside each row of 16), 16 and 32 (row-swap
// permutes); every lane of the wavefront must be active.
template <int MASK>
__device__ __forceinline__ float xor_partner(float v)
{
  static_assert(MASK == 8 || MASK == 16 || MASK == 32, "lane distance");
  if constexpr (MASK == 8)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xF, 0xF, false));
  else
    return partner<MASK == 16 ? 4 : 5>(v);
}

template <int MODEL, int COST, int NT, int SPL>
__device__ __forceinline__ void batch_tile(const BatchInstance& I, const IterArgs a, int g,
                                           const float* __restrict__ x0, const float* __restrict__ means,
                                           float* __restrict__ ubuf, Partial* __restrict__ parts,
                                           float* __restrict__ sums, float* __restrict__ costs, float* lds, int ch)
{
  using M = Model<MODEL>;
  using CF = Cost<COST, M::O>;
  constexpr int S = M::S, C = M::C, O = M::O;
  static_assert(O == S, "outputs are the states for the batched models");
  constexpr int R = 64 * SPL;
  constexpr int NW = NT / 64, NCONS = NW - 1;
  const int T = I.T, TC = T * C, K = I.K;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int k0 = g * R;
  const int nrows = min(R, K - k0);
  float* mu = lds;
  float* ring = mu + ((TC + 3) & ~3);     // [4][ch][R][C]
  float* ybuf = ring + 4 * ch * R * C;    // [2][ch][R][O]
  // [NCONS][R], after the larger of the rollout buffers and the column sums' blocks (batch_lds_floats)
  float* part = ring + max(4 * ch * R * C + 2 * ch * R * O, min(NW, SUM_WAVES) * sum_block_floats(C));
  float* term = part + NCONS * R;         // [R]
  float* cst = term + R;                  // [R]
  float* wl = cst + R;                    // [R]
  float* utile = ubuf + I.u_base + (int64_t)g * T * R * C;  // [T][R][C]
  const mppi_sampler_params& sp = I.sp;
  for (int j = tid; j < TC; j += NT) mu[j] = means[I.mean_base + j];
  const float decay = powf(sp.std_dev_decay, (float)a.iteration);
  const int nch = (T + ch - 1) / ch;
  __syncthreads();

  // Samples of chunk j into ring slot j % 4, by threads t_off .. t_off + nth - 1. Element (k, t, c) is normal
  // k*TC + t*C + c of the instance's stream. C divides 4, so a Philox block (4 normals) holds SPB = 4 / C whole steps of
  // one row: thread item (qi, r) evaluates a block of row r's range, whose first step lies m = (e mod 4) / C steps
  // before t0 when the row's first element e is not on a block boundary. A block that runs past the chunk's end (its
  // last steps belong to chunk j + 1) also stores those steps, into slot (j + 1) % 4, which holds no live chunk then
  // (the producer works on j - 1 and the cost wavefronts on j - 2); chunk j + 1 then skips that block. So every chunk
  // after the first needs ceil(n / SPB) blocks per row whatever the alignment (a chunk of at least SPB steps keeps
  // the stored steps inside chunk j + 1). Consecutive items are consecutive rows of one block index, so a wavefront's
  // LDS stores are row-contiguous and its mean reads are broadcasts. Element indices stay below K T C <= 2^26
  // (BATCH_MAX_K, MPPI_MAX_TIMESTEPS, MPPI_MAX_CONTROL), so 32-bit arithmetic holds them and the Philox counter's high
  // word is 0.
  static_assert(4 % C == 0, "a Philox block holds whole steps");
  constexpr int SPB = 4 / C;
  float sdc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) sdc[c] = decay * sp.std_dev[c];
  const float* sdt = sp.std_dev_t;
  auto gen = [&](int j, int t_off, int nth) {
    const int t0 = j * ch, n = min(T, t0 + ch) - t0;
    const bool aligned = (TC % 4 == 0) && ((t0 * C) % 4 == 0);
    const int nb = (n + SPB - 1) / SPB + ((j == 0 && !aligned) ? 1 : 0);
    float* slot = ring + (j % 4) * ch * R * C;
    float* next = ring + ((j + 1) % 4) * ch * R * C;
    for (int item = tid - t_off; item < R * nb; item += nth) {
      const int r = item & (R - 1), qi = item / R;
      const int kk = k0 + r;
      const int es = kk * TC + t0 * C;
      const int m = (es & 3) / C;                     // steps of the row's first block before the chunk
      const int q = qi + ((j > 0 && m > 0) ? 1 : 0);  // block 0 of a misaligned row came with chunk j - 1
      const int s0 = q * SPB - m;                     // chunk step of the block's first normal
      if (s0 >= n) continue;
      const Normal4 z = normals4(sp.seed, a.rng_offset, (uint64_t)(uint32_t)((es >> 2) + q));
      const bool keep = kk == 0, pure_row = kk >= I.pstart;
#pragma unroll
      for (int h = 0; h < SPB; ++h) {
        const int s = s0 + h;
        const int t = t0 + s;
        if (s < 0 || t >= T) continue;
        float v[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const float mv = mu[t * C + c];
          const float sd = sdt ? decay * sdt[t * C + c] : sdc[c];
          if (keep || t < a.stride) v[c] = mv;
          else if (pure_row) v[c] = sd * z.z[h * C + c];
          else v[c] = mv + sd * z.z[h * C + c];
        }
        float* dst = s < n ? slot + (s * R + r) * C : next + ((s - n) * R + r) * C;
        if constexpr (C == 2) *reinterpret_cast<float2*>(dst) = make_float2(v[0], v[1]);
        else if constexpr (C == 4) *reinterpret_cast<float4*>(dst) = make_float4(v[0], v[1], v[2], v[3]);
        else dst[0] = v[0];
      }
    }
  };

  gen(0, 0, NT);
  __syncthreads();

  float x[SPL][S];
  if (wave == 0) {
#pragma unroll
    for (int s = 0; s < SPL; ++s)
#pragma unroll
      for (int i = 0; i < S; ++i) x[s][i] = x0[i];
  }
  bool deadband = false;
#pragma unroll
  for (int c = 0; c < C; ++c) deadband |= I.dyn.control_deadband[c] != 0.0f;
  float acc[SPL];
  bool pure[SPL];
#pragma unroll
  for (int s = 0; s < SPL; ++s) {
    acc[s] = 0.0f;
    pure[s] = (k0 + lane + 64 * s) >= I.pstart;
  }

  for (int it = 0; it <= nch; ++it) {
    if (wave == 0 && it < nch) {
      const int t0 = it * ch, n = min(T, t0 + ch) - t0;
      float* rs = ring + (it % 4) * ch * R * C;
      float* yb = ybuf + (it & 1) * ch * R * O;
      float un[SPL][C];
#pragma unroll
      for (int s = 0; s < SPL; ++s)
#pragma unroll
        for (int c = 0; c < C; ++c) un[s][c] = rs[(lane + 64 * s) * C + c];
      for (int tt = 0; tt < n; ++tt) {
        const int tn = min(tt + 1, n - 1);
#pragma unroll
        for (int s = 0; s < SPL; ++s) {
          const int r = lane + 64 * s;
          float u[C];
#pragma unroll
          for (int c = 0; c < C; ++c) {
            u[c] = un[s][c];
            un[s][c] = rs[(tn * R + r) * C + c];
          }
          if (deadband) enforce_constraints<C>(I.dyn, u);
          else clamp_controls<C>(I.dyn, u);
#pragma unroll
          for (int c = 0; c < C; ++c) rs[(tt * R + r) * C + c] = u[c];
          float xd[S], xn[S];
          M::deriv(I.dyn, x[s], u, xd);
          M::update(I.dyn, x[s], xn, xd, I.dt);
#pragma unroll
          for (int i = 0; i < S; ++i) {
            yb[(tt * R + r) * O + i] = xn[i];
            x[s][i] = xn[i];
          }
        }
      }
      if (it == nch - 1) {
#pragma unroll
        for (int s = 0; s < SPL; ++s) term[lane + 64 * s] = CF::terminal(I.cost, x[s]);
      }
    } else if (wave > 0) {
      if (it > 0) {
        const int t0 = (it - 1) * ch, t1 = min(T, t0 + ch);
        const float* rs = ring + ((it - 1) % 4) * ch * R * C;
        const float* yb = ybuf + ((it - 1) & 1) * ch * R * O;
        // the wavefront's steps t0 + wave - 1 + j NCONS in groups of PF: the cost-map texels of a group's steps are
        // fetched first, so that their loads are in flight together, then the group's costs are added in step order
        constexpr int PF = BATCH_MAP_PREFETCH / SPL > 0 ? BATCH_MAP_PREFETCH / SPL : 1;
        for (int tb = t0 + wave - 1; tb < t1; tb += NCONS * PF) {
          MapTap tap[PF][SPL];
          if constexpr (COST == MPPI_COST_QUADRATIC_MAP) {
            if (CF::uses_map(I.cost)) {
#pragma unroll
              for (int g = 0; g < PF; ++g) {
                const int t = tb + g * NCONS;
                if (t < t1) {
#pragma unroll
                  for (int s = 0; s < SPL; ++s) {
                    const float* yy = yb + ((t - t0) * R + lane + 64 * s) * O;
                    tap[g][s] = map_fetch(I.cost, yy[I.cost.map_x_index], yy[I.cost.map_y_index]);
                  }
                }
              }
            }
          }
#pragma unroll
          for (int g = 0; g < PF; ++g) {
            const int t = tb + g * NCONS;
            if (t < t1) {
#pragma unroll
              for (int s = 0; s < SPL; ++s) {
                const int r = lane + 64 * s;
                // y is read in place from LDS: the cost's runtime-indexed terms (map position, speed) become LDS
                // addresses instead of register selects
                const float* yy = yb + ((t - t0) * R + r) * O;
                const float* u = rs + ((t - t0) * R + r) * C;
                float run;
                if constexpr (COST == MPPI_COST_QUADRATIC_MAP) run = CF::running_tap(I.cost, yy, t, tap[g][s]);
                else run = CF::running(I.cost, yy, t);
                acc[s] += run + lr_cost<C>(sp, u, mu + t * C, pure[s], I.lambda, I.alpha, t);
              }
            }
          }
        }
        // clamped samples of the chunk to the tile's [T][R][C] region (contiguous, 16-byte stores)
        const int n4 = (t1 - t0) * R * C / 4;
        const float4* src = reinterpret_cast<const float4*>(rs);
        float4* dst = reinterpret_cast<float4*>(utile + (int64_t)t0 * R * C);
        for (int i = tid - 64; i < n4; i += NT - 64) dst[i] = src[i];
      }
      if (it + 1 < nch) gen(it + 1, 64, NT - 64);
    }
    __syncthreads();
  }
  if (wave > 0) {
#pragma unroll
    for (int s = 0; s < SPL; ++s) part[(wave - 1) * R + lane + 64 * s] = acc[s];
  }
  __syncthreads();
  if (wave == 0) {
    float v = INFINITY;
    int vi = INT_MAX;
    float cs[SPL];
#pragma unroll
    for (int s = 0; s < SPL; ++s) {
      const int r = lane + 64 * s;
      float run = 0.0f;
#pragma unroll 4
      for (int w = 0; w < NCONS; ++w) run += part[w * R + r];
      const float cost_r = run / (float)T + term[r] / (float)T;
      const bool valid = r < nrows;
      if (valid) costs[I.cost_base + k0 + r] = cost_r;
      cs[s] = valid ? sanitize_cost(cost_r) : INFINITY;
      argmin_combine(v, vi, cs[s], valid ? k0 + r : INT_MAX);
    }
    wave_argmin(v, vi);
    double e = 0.0, e2 = 0.0, ec = 0.0;
#pragma unroll
    for (int s = 0; s < SPL; ++s) {
      const bool fin = cs[s] <= FLT_MAX;
      const float w = fin ? expf(-(1.0f / I.lambda) * (cs[s] - v)) : 0.0f;
      wl[lane + 64 * s] = w;
      e += (double)w;
      e2 += (double)w * (double)w;
      if (fin) ec += (double)w * (double)cs[s];
    }
    e = wave_sum(e);
    e2 = wave_sum(e2);
    ec = wave_sum(ec);
    if (lane == 0) {
      Partial p;
      p.baseline = v;
      p.argmin = vi;
      p.eta = e;
      p.s2 = e2;
      p.wc = ec;
      parts[I.part_base + g] = p;
    }
  }
  __syncthreads();
  // S_g[t][c] = sum_r w_r u[t][r][c], P = 8 C outputs (8 steps) per wavefront pass: every lane issues the loads of its
  // rows for the 8 steps at once (one memory latency per pass instead of one per output) and scales them by the rows'
  // weights; the 64 x P block goes through LDS (row stride P + 1) so that lane (q, h) adds rows h P .. h P + P - 1 of
  // output q, and lane exchanges add the 64 / P row groups. The rows of a tile are summed in a fixed order
  // (deterministic).
  constexpr int P = 8 * C, NSUM = NW < SUM_WAVES ? NW : SUM_WAVES;
  if (wave < NSUM) {
    float* blk = ring + wave * sum_block_floats(C);
    const int q = lane % P, h = lane / P;
    for (int t0 = wave * 8; t0 < T; t0 += NSUM * 8) {
      float v[P];
#pragma unroll
      for (int i = 0; i < P; ++i) v[i] = 0.0f;
#pragma unroll
      for (int s = 0; s < SPL; ++s) {
        const int r = lane + 64 * s;
        const float wr = wl[r];
        const float* src = utile + ((int64_t)t0 * R + r) * C;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if (t0 + j < T) {
#pragma unroll
            for (int c = 0; c < C; ++c) v[j * C + c] += wr * src[(int64_t)j * R * C + c];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < P; ++i) blk[lane * (P + 1) + i] = v[i];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      float acc_q = 0.0f;
#pragma unroll
      for (int i = 0; i < P; ++i) acc_q += blk[(h * P + i) * (P + 1) + q];
      if constexpr (P <= 8) acc_q += xor_partner<8>(acc_q);
      if constexpr (P <= 16) acc_q += xor_partner<16>(acc_q);
      acc_q += xor_partner<32>(acc_q);
      if (h == 0 && t0 * C + q < TC) sums[I.sum_base + (int64_t)g * TC + t0 * C + q] = acc_q;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
  }
}

template <int NT, int SPL>
__global__ void __launch_bounds__(NT) batch_rollout_kernel(const BatchInstance* __restrict__ inst,
                                                           const int2* __restrict__ work, int nwork,
                                                           int* __restrict__ queue,
                                                           const IterArgs* __restrict__ iters,
                                                           const float* __restrict__ x0s,
                                                           const float* __restrict__ means,
                                                           float* __restrict__ ubuf, Partial* __restrict__ parts,
                                                           float* __restrict__ sums, float* __restrict__ costs, int ch)
{
  extern __shared__ float lds[];
  int* s_item = reinterpret_cast<int*>(lds);  // the tile's buffers start 4 floats in
  for (;;) {
    if (threadIdx.x == 0) *s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(*s_item);
    if (w >= nwork) break;
    const int2 wi = work[w];
    const int i = __builtin_amdgcn_readfirstlane(wi.x);
    const int g = __builtin_amdgcn_readfirstlane(wi.y);
    const BatchInstance& I = inst[i];
    const IterArgs a = iters[i];
    const float* x0 = x0s + (int64_t)i * BATCH_MAX_STATE;
    switch (I.dyn.model) {
      case MPPI_MODEL_DOUBLE_INTEGRATOR:
        batch_tile<MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_QUADRATIC_MAP, NT, SPL>(I, a, g, x0, means, ubuf, parts,
                                                                                   sums, costs, lds + 4, ch);
        break;
      case MPPI_MODEL_RACER_DUBINS:
        batch_tile<MPPI_MODEL_RACER_DUBINS, MPPI_COST_QUADRATIC_MAP, NT, SPL>(I, a, g, x0, means, ubuf, parts, sums,
                                                                              costs, lds + 4, ch);
        break;
      default:
        batch_tile<MPPI_MODEL_CARTPOLE, MPPI_COST_CARTPOLE_QUADRATIC, NT, SPL>(I, a, g, x0, means, ubuf, parts,
                                                                               sums, costs, lds + 4, ch);
        break;
    }
    __syncthreads();
  }
}

// Where the combine also writes a call's results (device-side aliases of mapped host memory; null members are
// skipped): raw means and constrained controls [total_mean], statistics [n], then the host completion signal.
struct BatchCopies {
  float* mean = nullptr;
  float* control = nullptr;
  mppi_weight_stats* stats = nullptr;
  HostSignal signal;
};

// One workgroup per instance: global baseline (first strict minimum over the tiles' minima), FP64 normalizer and
// sum of squares, new mean S / eta, statistics; optionally advances the instance's generator counter. The new mean is
// then Savitzky-Golay smoothed and constrained (computeControl's post-processing, controllers/controller.cuh:573-607
// and controllers/MPPI/mppi_controller.cu:224-227) into the instance's mean in place, the next call's nominal.
// Workgroup 0 resets the work-queue counter for the next launch.
constexpr int BATCH_COMBINE_THREADS = 256;
constexpr int BATCH_MAX_TILES = BATCH_MAX_K / 64;

__global__ void __launch_bounds__(BATCH_COMBINE_THREADS) batch_combine_kernel(
    const BatchInstance* __restrict__ inst, const Partial* __restrict__ parts, const float* __restrict__ sums,
    float* __restrict__ means, mppi_weight_stats* __restrict__ stats, IterArgs* iters, int advance,
    int* __restrict__ queue, BatchCopies out)
{
  __shared__ float escale[BATCH_MAX_TILES];
  __shared__ float raw[MPPI_MAX_TIMESTEPS * MPPI_MAX_CONTROL];
  __shared__ float sv[BATCH_COMBINE_THREADS / 64];
  __shared__ int si[BATCH_COMBINE_THREADS / 64];
  __shared__ double se[BATCH_COMBINE_THREADS / 64], se2[BATCH_COMBINE_THREADS / 64], sec[BATCH_COMBINE_THREADS / 64];
  const int i = blockIdx.x;
  const BatchInstance& I = inst[i];
  const int G = I.tiles, TC = I.T * I.C;
  const Partial* P = parts + I.part_base;
  const float* Sg = sums + I.sum_base;
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  float v = INFINITY;
  int vi = INT_MAX;
  for (int q = tid; q < G; q += BATCH_COMBINE_THREADS) argmin_combine(v, vi, P[q].baseline, P[q].argmin);
  wave_argmin(v, vi);
  if (lane == 0) {
    sv[wid] = v;
    si[wid] = vi;
  }
  __syncthreads();
  float b = sv[0];
  int bi = si[0];
  for (int q = 1; q < BATCH_COMBINE_THREADS / 64; ++q) argmin_combine(b, bi, sv[q], si[q]);
  if (bi == INT_MAX) bi = 0;
  const double inv = 1.0 / (double)I.lambda;
  double e = 0.0, e2 = 0.0, ec = 0.0;
  for (int q = tid; q < G; q += BATCH_COMBINE_THREADS) {
    const Partial p = P[q];
    const double sc = (p.baseline <= FLT_MAX) ? exp(-inv * ((double)p.baseline - (double)b)) : 0.0;
    escale[q] = (float)sc;
    e += p.eta * sc;
    e2 += p.s2 * sc * sc;
    ec += p.wc * sc;
  }
  e = wave_sum(e);
  e2 = wave_sum(e2);
  ec = wave_sum(ec);
  if (lane == 0) {
    se[wid] = e;
    se2[wid] = e2;
    sec[wid] = ec;
  }
  __syncthreads();
  double eta = 0.0, s2 = 0.0, wcost = 0.0;
  for (int q = 0; q < BATCH_COMBINE_THREADS / 64; ++q) {
    eta += se[q];
    s2 += se2[q];
    wcost += sec[q];
  }
  for (int j = tid; j < TC; j += BATCH_COMBINE_THREADS) {
    float acc = 0.0f;
    for (int q = 0; q < G; ++q) acc += escale[q] * Sg[(int64_t)q * TC + j];
    const float m = (float)((double)acc / eta);
    raw[j] = m;
    if (out.mean != nullptr) out.mean[I.mean_base + j] = m;
  }
  __syncthreads();
  const int T = I.T, C = I.C;
  for (int j = tid; j < TC; j += BATCH_COMBINE_THREADS) {
    const int t = j / C, c = j - t * C;
    const float u = constrain_control(I.dyn, c, sg_smooth_at(raw, t, c, T, C), true);
    means[I.mean_base + j] = u;
    if (out.control != nullptr) out.control[I.mean_base + j] = u;
  }
  if (tid == 0) {
    mppi_weight_stats st;
    st.baseline = b;
    st.argmin = bi;
    st.normalizer = eta;
    st.sum_w2 = s2;
    st.lambda = I.lambda;
    st.num_rollouts = I.K;
    st.weighted_cost = wcost / eta;
    stats[i] = st;
    if (out.stats != nullptr) out.stats[i] = st;
    if (advance) iters[i].rng_offset += 1;
    if (i == 0) *queue = 0;
  }
  signal_host_done(out.signal);
}

// The per-call inputs from mapped host memory, one workgroup per instance: iteration arguments, x0 and, where the
// host changed the control sequence since the last call (upload[i]), the nominal sequence. Otherwise the mean the
// last combine left in place (the same values the host holds) stays.
__global__ void __launch_bounds__(BATCH_COMBINE_THREADS) batch_stage_kernel(
    const BatchInstance* __restrict__ inst, const IterArgs* __restrict__ iters_h, IterArgs* __restrict__ iters_d,
    const float* __restrict__ x0_h, float* __restrict__ x0_d, const float* __restrict__ nominal_h,
    const int* __restrict__ upload_h, float* __restrict__ means)
{
  const int i = blockIdx.x, tid = threadIdx.x;
  constexpr int W = (int)(sizeof(IterArgs) / 4);
  if (tid < W)
    reinterpret_cast<uint32_t*>(iters_d + i)[tid] = reinterpret_cast<const uint32_t*>(iters_h + i)[tid];
  else if (tid >= 64 && tid < 64 + BATCH_MAX_STATE)
    x0_d[i * BATCH_MAX_STATE + tid - 64] = x0_h[i * BATCH_MAX_STATE + tid - 64];
  const int up = __builtin_amdgcn_readfirstlane(upload_h[i]);
  if (up != 0) {
    const BatchInstance& I = inst[i];
    const int TC = I.T * I.C;
    for (int j = tid; j < TC; j += BATCH_COMBINE_THREADS) means[I.mean_base + j] = nominal_h[I.mean_base + j];
  }
}

template <int NT, int SPL>
static int launch_batch_t(const BatchInstance* inst, const int2* work, int nwork, int* queue, const IterArgs* iters,
                          const float* x0s, const float* means, float* ubuf, Partial* parts, float* sums, float* costs,
                          int ch, size_t lds_bytes, int blocks, hipStream_t s)
{
  MPPI_CHECK(allow_max_dynamic_lds((const void*)batch_rollout_kernel<NT, SPL>));
  batch_rollout_kernel<NT, SPL><<<blocks, NT, lds_bytes, s>>>(inst, work, nwork, queue, iters, x0s, means, ubuf,
                                                              parts, sums, costs, ch);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

template <int NT, int SPL>
static int occupancy_t(size_t lds_bytes, int* out)
{
  MPPI_CHECK(allow_max_dynamic_lds((const void*)batch_rollout_kernel<NT, SPL>));
  MPPI_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(out, (const void*)batch_rollout_kernel<NT, SPL>, NT,
                                                              lds_bytes));
  return MPPI_OK;
}

#define MPPI_BATCH_SHAPES(X) \
  X(256, 1) X(512, 1) X(1024, 1) X(256, 2) X(512, 2) X(1024, 2) X(256, 4) X(512, 4) X(1024, 4)

static int batch_occupancy(int NT, int SPL, size_t lds_bytes, int* out)
{
#define MPPI_OCC(nt, spl) \
  if (NT == nt && SPL == spl) return occupancy_t<nt, spl>(lds_bytes, out);
  MPPI_BATCH_SHAPES(MPPI_OCC)
#undef MPPI_OCC
  return fail(MPPI_E_INVALID, "batch: unsupported launch shape");
}

static int launch_batch(int NT, int SPL, const BatchInstance* inst, const int2* work, int nwork, int* queue,
                        const IterArgs* iters, const float* x0s, const float* means, float* ubuf, Partial* parts,
                        float* sums, float* costs, int ch, size_t lds_bytes, int blocks, hipStream_t s)
{
#define MPPI_LAUNCH(nt, spl)                                                                                    \
  if (NT == nt && SPL == spl)                                                                                   \
    return launch_batch_t<nt, spl>(inst, work, nwork, queue, iters, x0s, means, ubuf, parts, sums, costs, ch,   \
                                   lds_bytes, blocks, s);
  MPPI_BATCH_SHAPES(MPPI_LAUNCH)
#undef MPPI_LAUNCH
  return fail(MPPI_E_INVALID, "batch: unsupported launch shape");
}

}  // namespace mppi_amd

using namespace mppi_amd;

// ======================================================================================================= C-ABI
struct mppi_batch {
  int n = 0;
  std::vector<mppi_instance_desc> desc;
  std::vector<int> K, T, C, S;
  int TCmax = 0, Cmax = 1, Omax = 1, cus = 0;
  hipStream_t stream = nullptr;
  float* map_d = nullptr;
  float* tables_d = nullptr;  // the instances' per-step standard deviations and goal trajectories, back to back
  BatchInstance* inst_d = nullptr;
  std::vector<BatchInstance> inst_h;
  int2* work_d = nullptr;
  int nwork = 0;
  int* queue_d = nullptr;
  IterArgs* iters_d = nullptr;
  IterArgs* iters_h = nullptr;
  IterArgs* loop_iters_d = nullptr;
  float* x0_d = nullptr;
  float* x0_h = nullptr;
  float* means_d = nullptr;
  float* means_backup_d = nullptr;
  float* nominal_h = nullptr;
  float* control_out_h = nullptr;
  int* upload_h = nullptr;
  // device-side aliases of the mapped host buffers above
  IterArgs* iters_hd = nullptr;
  float* x0_hd = nullptr;
  float* nominal_hd = nullptr;
  float* control_out_hd = nullptr;
  int* upload_hd = nullptr;
  mppi_weight_stats* stats_hd = nullptr;
  // completion signal: device workgroup count and epoch, the mapped host word, signals enqueued
  uint32_t* sig_d = nullptr;
  uint32_t* done_h = nullptr;
  uint32_t* done_hd = nullptr;
  uint32_t signals = 0;
  hipEvent_t ev_end = nullptr;
  bool use_graph = false;
  std::vector<char> dirty;  // the host's control sequence differs from the device's nominal
  // The last call's results are read where the combine left them: the controls and statistics in mapped host memory,
  // the raw means in raw_d (copied to the host on the first request). ctl_host[i]: control[i] (host) replaces the
  // mapped controls after a setControlSeq / slide.
  float* raw_d = nullptr;
  std::vector<char> ctl_host;
  mutable std::vector<char> raw_fresh;
  std::vector<float> prev_baseline;
  std::vector<int> mean_base;
  int total_mean = 0;
  mppi_weight_stats* stats_d = nullptr;
  mppi_weight_stats* stats_h = nullptr;
  float* ubuf_d = nullptr;
  Partial* parts_d = nullptr;
  float* sums_d = nullptr;
  float* costs_d = nullptr;
  std::vector<int> cost_base;
  mppi_batch_config cfg{};
  size_t lds_bytes = 0;
  int blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t loop_graph = nullptr;
  hipGraphExec_t loop_exec = nullptr;
  int eager_runs = 0;
  bool computed = false;    // computeControl has run once (the device loop starts from its x0 and means)
  bool loop_ready = false;  // loop_iters_d holds the device loop's arguments
  uint64_t rng_counter = 0;
  std::vector<std::vector<float>> control;
  mutable std::vector<std::vector<float>> raw_mean;
  // per instance, propagated on the first request after a call from that call's x0 and smoothed controls
  mutable std::vector<std::vector<float>> state_seq;
  mutable std::vector<char> state_stale;
};

static void batch_drop_graphs(mppi_batch* b)
{
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->loop_exec) (void)hipGraphExecDestroy(b->loop_exec);
  if (b->loop_graph) (void)hipGraphDestroy(b->loop_graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  b->loop_exec = nullptr;
  b->loop_graph = nullptr;
  b->eager_runs = 0;
}

static int batch_free(mppi_batch* b)
{
  if (!b) return MPPI_OK;
  batch_drop_graphs(b);
  for (void* p : {(void*)b->map_d, (void*)b->tables_d, (void*)b->inst_d, (void*)b->work_d, (void*)b->queue_d, (void*)b->iters_d,
                  (void*)b->loop_iters_d, (void*)b->x0_d, (void*)b->means_d, (void*)b->means_backup_d,
                  (void*)b->stats_d, (void*)b->ubuf_d, (void*)b->parts_d, (void*)b->sums_d, (void*)b->costs_d,
                  (void*)b->sig_d, (void*)b->raw_d})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)b->iters_h, (void*)b->x0_h, (void*)b->nominal_h, (void*)b->stats_h, (void*)b->control_out_h,
                  (void*)b->upload_h, (void*)b->done_h})
    if (p) (void)hipHostFree(p);
  if (b->ev_end) (void)hipEventDestroy(b->ev_end);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MPPI_OK;
}

// Instance table, work list (tiles ordered by decreasing horizon, then instance, then tile) and LDS size for the
// launch shape `c`; uploads them. Returns MPPI_E_SIZE when the shape does not fit the LDS.
static int batch_apply_config(mppi_batch* b, const mppi_batch_config& c)
{
  const int R = 64 * c.samples_per_lane;
  // the sample generation stores a Philox block's steps past a chunk into the next chunk only: chunk >= 4 / C
  for (int Ci : b->C)
    if (c.chunk * Ci < 4) return fail(MPPI_E_INVALID, "batch: chunk must be at least 4 / control_dim steps");
  const size_t lds = sizeof(float) * batch_lds_floats(b->TCmax, b->Cmax, b->Omax, c.threads, R, c.chunk);
  if (lds > LDS_LIMIT) return fail(MPPI_E_SIZE, "batch: launch shape exceeds the LDS");
  int occ = 0;
  MPPI_CHECK(batch_occupancy(c.threads, c.samples_per_lane, lds, &occ));
  if (occ < 1) return fail(MPPI_E_SIZE, "batch: launch shape does not fit on a CU");
  int part = 0;
  int64_t sum = 0, u = 0;
  std::vector<int2> work;
  for (int i = 0; i < b->n; ++i) {
    BatchInstance& I = b->inst_h[i];
    I.tiles = (I.K + R - 1) / R;
    I.part_base = part;
    I.sum_base = sum;
    I.u_base = u;
    part += I.tiles;
    sum += (int64_t)I.tiles * I.T * I.C;
    u += (int64_t)I.tiles * I.T * R * I.C;
    for (int g = 0; g < I.tiles; ++g) work.push_back(int2{i, g});
  }
  std::stable_sort(work.begin(), work.end(),
                   [&](const int2& x, const int2& y) { return b->inst_h[x.x].T > b->inst_h[y.x].T; });
  MPPI_HIP_CHECK(hipMemcpy(b->inst_d, b->inst_h.data(), sizeof(BatchInstance) * b->n, hipMemcpyHostToDevice));
  MPPI_HIP_CHECK(hipMemcpy(b->work_d, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice));
  b->nwork = (int)work.size();
  b->lds_bytes = lds;
  b->blocks = std::max(1, std::min(b->nwork, occ * b->cus));
  b->cfg = c;
  b->cfg.blocks_per_cu = occ;
  batch_drop_graphs(b);
  return MPPI_OK;
}

// One batched iteration: persistent rollout over every tile, then the per-instance combine.
// `between` (optional) is recorded between the persistent rollout launch and the combine launch.
// `out` (optional): the call's host copies and completion signal.
static int batch_enqueue_iteration(mppi_batch* b, IterArgs* iters, int advance, hipEvent_t between = nullptr,
                                   const BatchCopies* out = nullptr)
{
  MPPI_CHECK(launch_batch(b->cfg.threads, b->cfg.samples_per_lane, b->inst_d, b->work_d, b->nwork, b->queue_d, iters,
                          b->x0_d, b->means_d, b->ubuf_d, b->parts_d, b->sums_d, b->costs_d, b->cfg.chunk,
                          b->lds_bytes, b->blocks, b->stream));
  if (between != nullptr) MPPI_HIP_CHECK(hipEventRecord(between, b->stream));
  batch_combine_kernel<<<b->n, BATCH_COMBINE_THREADS, 0, b->stream>>>(b->inst_d, b->parts_d, b->sums_d, b->means_d,
                                                                      b->stats_d, iters, advance, b->queue_d,
                                                                      out != nullptr ? *out : BatchCopies{});
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// One computeControl of the whole batch: the inputs staged from mapped host memory, the iteration, the results and
// the completion signal written to mapped host memory by the combine.
static int batch_enqueue_call(mppi_batch* b)
{
  batch_stage_kernel<<<b->n, BATCH_COMBINE_THREADS, 0, b->stream>>>(b->inst_d, b->iters_hd, b->iters_d, b->x0_hd,
                                                                    b->x0_d, b->nominal_hd, b->upload_hd, b->means_d);
  MPPI_HIP_CHECK(hipGetLastError());
  BatchCopies out;
  out.mean = b->raw_d;
  out.control = b->control_out_hd;
  out.stats = b->stats_hd;
  out.signal.count = b->sig_d;
  out.signal.epoch = b->sig_d + 1;
  out.signal.flag = b->done_hd;
  return batch_enqueue_iteration(b, b->iters_d, 0, nullptr, &out);
}

extern "C" {

int mppi_batch_create(const mppi_instance_desc* instances, int32_t n, const float* cost_map_host, int32_t map_h,
                      int32_t map_w, mppi_batch** out)
{
  if (!instances || !out || n < 1) return fail(MPPI_E_INVALID, "batch: null argument or no instances");
  *out = nullptr;
  mppi_batch* b = new (std::nothrow) mppi_batch();
  if (!b) return fail(MPPI_E_NOMEM, "batch: out of host memory");
  auto bail = [&](int st) {
    batch_free(b);
    return st;
  };
  b->n = n;
  b->desc.assign(instances, instances + n);
  if (cost_map_host) {
    if (map_h < 2 || map_w < 2) return bail(fail(MPPI_E_INVALID, "batch: cost map must be at least 2x2"));
    const int st = dev_alloc(&b->map_d, sizeof(float) * (size_t)map_h * map_w);
    if (st != MPPI_OK) return bail(st);
    const hipError_t e =
        hipMemcpy(b->map_d, cost_map_host, sizeof(float) * (size_t)map_h * map_w, hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMemcpy"));
  }
  b->inst_h.resize(n);
  struct Table {
    int inst, kind;  // kind 0: per-step std_dev [T][C], 1: goal trajectory [H][O]
    const float* host;
    size_t count;
  };
  std::vector<Table> tables;
  int64_t ubuf = 0, sums = 0, costs = 0;
  int parts = 0;
  for (int i = 0; i < n; ++i) {
    mppi_instance_desc& d = b->desc[i];
    const int m = d.dynamics.model;
    if (m != MPPI_MODEL_DOUBLE_INTEGRATOR && m != MPPI_MODEL_RACER_DUBINS && m != MPPI_MODEL_CARTPOLE)
      return bail(fail(MPPI_E_UNSUPPORTED, "batch: instance " + std::to_string(i) + ": model not batched"));
    const int want_cost = (m == MPPI_MODEL_CARTPOLE) ? MPPI_COST_CARTPOLE_QUADRATIC : MPPI_COST_QUADRATIC_MAP;
    if (d.cost.cost != want_cost)
      return bail(fail(MPPI_E_UNSUPPORTED, "batch: instance " + std::to_string(i) + ": model/cost pair not batched"));
    if (d.sampler.kind != MPPI_SAMPLER_GAUSSIAN)
      return bail(fail(MPPI_E_UNSUPPORTED, "batch: instance " + std::to_string(i) + ": Gaussian sampling only"));
    if (d.cost.cost == MPPI_COST_QUADRATIC_MAP && d.cost.map == nullptr && d.cost.map_coeff != 0.0f) {
      if (!b->map_d) return bail(fail(MPPI_E_INVALID, "batch: instance " + std::to_string(i) + " needs a map"));
      d.cost.map = b->map_d;
      d.cost.map_height = map_h;
      d.cost.map_width = map_w;
    }
    // host tables stand in for their device copies while the blocks are validated
    if (d.goal_traj_host != nullptr) d.cost.goal_traj = d.goal_traj_host;
    int st = validate_sizes(d.num_rollouts, d.num_timesteps);
    if (st == MPPI_OK) st = validate_dynamics(d.dynamics);
    if (st == MPPI_OK) st = validate_cost(d.cost, d.dynamics, true);
    if (st == MPPI_OK) st = validate_sampler(d.sampler, d.dynamics);
    if (st != MPPI_OK) return bail(fail(st, "batch: instance " + std::to_string(i) + ": " + mppi_last_error()));
    if (d.num_rollouts % 64 != 0 || d.num_rollouts > BATCH_MAX_K)
      return bail(fail(MPPI_E_INVALID, "batch: num_rollouts must be a multiple of 64 and at most 65536"));
    if (!(d.lambda > 0.0f)) return bail(fail(MPPI_E_INVALID, "batch: lambda must be > 0"));
    int S, C, O;
    model_dims(m, &S, &C, &O);
    const int K = d.num_rollouts, T = d.num_timesteps;
    b->K.push_back(K);
    b->T.push_back(T);
    b->C.push_back(C);
    b->S.push_back(S);
    b->TCmax = std::max(b->TCmax, T * C);
    b->Cmax = std::max(b->Cmax, C);
    b->Omax = std::max(b->Omax, O);
    BatchInstance& I = b->inst_h[i];
    std::memset(&I, 0, sizeof(I));
    I.dyn = d.dynamics;
    I.cost = d.cost;
    I.sp = d.sampler;
    I.K = K;
    I.T = T;
    I.C = C;
    I.pstart = pure_noise_start(K, d.sampler.pure_noise_trajectories_percentage);
    I.mean_base = b->total_mean;
    I.cost_base = (int)costs;
    I.dt = d.dt;
    I.lambda = d.lambda;
    I.alpha = d.alpha;
    b->mean_base.push_back(b->total_mean);
    b->cost_base.push_back((int)costs);
    b->total_mean += T * C;
    const int t64 = (K + 63) / 64;
    parts += t64;
    sums += (int64_t)t64 * T * C;
    ubuf += (int64_t)((K + 255) / 256) * 256 * T * C;
    costs += K;
    if (d.std_dev_t_host != nullptr) {
      for (int j = 0; j < T * C; ++j)
        if (!(d.std_dev_t_host[j] > 0.0f))
          return bail(fail(MPPI_E_INVALID, "batch: instance " + std::to_string(i) + ": per-step std_dev must be > 0"));
      tables.push_back({i, 0, d.std_dev_t_host, (size_t)T * C});
    }
    if (d.goal_traj_host != nullptr) tables.push_back({i, 1, d.goal_traj_host, (size_t)d.cost.goal_traj_len * O});
    b->control.emplace_back((size_t)T * C, 0.0f);
    if (d.init_control) std::copy(d.init_control, d.init_control + (size_t)T * C, b->control.back().begin());
    b->raw_mean.emplace_back((size_t)T * C, 0.0f);
    b->state_seq.emplace_back((size_t)T * S, 0.0f);
  }
  b->state_stale.assign(n, 0);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(hip_fail(e, "batch: device setup"));
  int st = MPPI_OK;
  auto A = [&](int s2) {
    if (st == MPPI_OK) st = s2;
  };
  A(dev_alloc(&b->inst_d, sizeof(BatchInstance) * n));
  A(dev_alloc(&b->work_d, sizeof(int2) * (size_t)parts));
  A(dev_alloc(&b->queue_d, sizeof(int)));
  A(dev_alloc(&b->iters_d, sizeof(IterArgs) * n));
  A(dev_alloc(&b->loop_iters_d, sizeof(IterArgs) * n));
  A(dev_alloc(&b->x0_d, sizeof(float) * BATCH_MAX_STATE * n));
  A(dev_alloc(&b->means_d, sizeof(float) * b->total_mean));
  A(dev_alloc(&b->means_backup_d, sizeof(float) * b->total_mean));
  A(dev_alloc(&b->stats_d, sizeof(mppi_weight_stats) * n));
  A(dev_alloc(&b->ubuf_d, sizeof(float) * (size_t)ubuf));
  A(dev_alloc(&b->parts_d, sizeof(Partial) * (size_t)parts));
  A(dev_alloc(&b->sums_d, sizeof(float) * (size_t)sums));
  A(dev_alloc(&b->costs_d, sizeof(float) * (size_t)costs));
  A(host_alloc_mapped(&b->iters_h, &b->iters_hd, sizeof(IterArgs) * n));
  A(host_alloc_mapped(&b->x0_h, &b->x0_hd, sizeof(float) * BATCH_MAX_STATE * n));
  A(host_alloc_mapped(&b->nominal_h, &b->nominal_hd, sizeof(float) * b->total_mean));
  A(dev_alloc(&b->raw_d, sizeof(float) * b->total_mean));
  A(host_alloc_mapped(&b->control_out_h, &b->control_out_hd, sizeof(float) * b->total_mean));
  A(host_alloc_mapped(&b->upload_h, &b->upload_hd, sizeof(int) * n));
  A(host_alloc_mapped(&b->stats_h, &b->stats_hd, sizeof(mppi_weight_stats) * n));
  A(host_alloc_mapped(&b->done_h, &b->done_hd, sizeof(uint32_t)));
  A(dev_alloc(&b->sig_d, 2 * sizeof(uint32_t)));
  b->dirty.assign(n, 1);
  b->ctl_host.assign(n, 1);
  b->raw_fresh.assign(n, 1);
  b->prev_baseline.assign(n, 0.0f);
  size_t table_floats = 0;
  for (const Table& t : tables) table_floats += t.count;
  if (table_floats > 0) A(dev_alloc(&b->tables_d, sizeof(float) * table_floats));
  if (st != MPPI_OK) return bail(st);
  size_t off = 0;
  for (const Table& t : tables) {
    e = hipMemcpy(b->tables_d + off, t.host, sizeof(float) * t.count, hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMemcpy"));
    if (t.kind == 0) b->inst_h[t.inst].sp.std_dev_t = b->desc[t.inst].sampler.std_dev_t = b->tables_d + off;
    else b->inst_h[t.inst].cost.goal_traj = b->desc[t.inst].cost.goal_traj = b->tables_d + off;
    off += t.count;
  }
  for (mppi_instance_desc& d : b->desc) d.std_dev_t_host = d.goal_traj_host = nullptr;
  e = hipMemset(b->queue_d, 0, sizeof(int));
  if (e == hipSuccess) e = hipMemset(b->sig_d, 0, 2 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_end, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMemset(b->x0_d, 0, sizeof(float) * BATCH_MAX_STATE * n);
  if (e == hipSuccess) e = hipMemset(b->means_d, 0, sizeof(float) * b->total_mean);
  if (e != hipSuccess) return bail(hip_fail(e, "hipMemset"));
  // default launch shape: 1024 threads, one sample per lane, the largest chunk that fits
  mppi_batch_config c{1024, 1, 16, 0, 0.0f};
  st = MPPI_E_SIZE;
  for (int ch : {16, 8, 4}) {
    c.chunk = ch;
    st = batch_apply_config(b, c);
    if (st == MPPI_OK) break;
  }
  if (st != MPPI_OK) return bail(st);
  *out = b;
  return MPPI_OK;
}

int mppi_batch_destroy(mppi_batch* b)
{
  if (b && b->stream) (void)hipStreamSynchronize(b->stream);
  return batch_free(b);
}

int mppi_batch_set_config(mppi_batch* b, const mppi_batch_config* c)
{
  if (!b || !c) return fail(MPPI_E_INVALID, "null argument");
  if (c->samples_per_lane != 1 && c->samples_per_lane != 2 && c->samples_per_lane != 4)
    return fail(MPPI_E_INVALID, "batch: samples_per_lane must be 1, 2 or 4");
  if (c->threads != 256 && c->threads != 512 && c->threads != 1024)
    return fail(MPPI_E_INVALID, "batch: threads must be 256, 512 or 1024");
  if (c->chunk < 2 || c->chunk > 16) return fail(MPPI_E_INVALID, "batch: chunk must be in [2, 16]");
  MPPI_HIP_CHECK(hipStreamSynchronize(b->stream));
  return batch_apply_config(b, *c);
}

int mppi_batch_get_config(const mppi_batch* b, mppi_batch_config* c)
{
  if (!b || !c) return fail(MPPI_E_INVALID, "null argument");
  *c = b->cfg;
  return MPPI_OK;
}

int mppi_batch_autotune(mppi_batch* b, int32_t reps, mppi_batch_config* best_out)
{
  if (!b) return fail(MPPI_E_INVALID, "null argument");
  if (reps < 1) return fail(MPPI_E_INVALID, "batch: reps must be >= 1");
  MPPI_HIP_CHECK(hipStreamSynchronize(b->stream));
  b->dirty.assign(b->n, 1);
  MPPI_HIP_CHECK(hipMemcpy(b->means_backup_d, b->means_d, sizeof(float) * b->total_mean, hipMemcpyDeviceToDevice));
  for (int i = 0; i < b->n; ++i) b->iters_h[i] = IterArgs{b->rng_counter, 1, 0};
  MPPI_HIP_CHECK(hipMemcpy(b->iters_d, b->iters_h, sizeof(IterArgs) * b->n, hipMemcpyHostToDevice));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  MPPI_HIP_CHECK(hipEventCreate(&e0));
  MPPI_HIP_CHECK(hipEventCreate(&e1));
  // every valid shape is timed in two rounds and keeps its faster round, so that a shape timed while the clocks
  // were still ramping up is not penalised
  std::vector<mppi_batch_config> shapes;
  for (int nt : {256, 512, 1024})
    for (int spl : {1, 2, 4})
      for (int ch : {3, 4, 6, 8, 12, 16}) shapes.push_back(mppi_batch_config{nt, spl, ch, 0, 0.0f});
  std::vector<float> best_us(shapes.size(), -1.0f);
  int st = MPPI_OK;
  for (int round = 0; round < 2 && st == MPPI_OK; ++round)
    for (size_t k = 0; k < shapes.size() && st == MPPI_OK; ++k) {
      if (batch_apply_config(b, shapes[k]) != MPPI_OK) continue;
      shapes[k].blocks_per_cu = b->cfg.blocks_per_cu;
      float ms = 0.0f;
      st = batch_enqueue_iteration(b, b->iters_d, 0);  // warm-up
      if (st == MPPI_OK && hipEventRecord(e0, b->stream) != hipSuccess) st = MPPI_E_HIP;
      for (int r = 0; r < reps && st == MPPI_OK; ++r) st = batch_enqueue_iteration(b, b->iters_d, 0);
      if (st == MPPI_OK && hipEventRecord(e1, b->stream) != hipSuccess) st = MPPI_E_HIP;
      if (st == MPPI_OK && hipEventSynchronize(e1) != hipSuccess) st = MPPI_E_HIP;
      if (st == MPPI_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = MPPI_E_HIP;
      if (hipMemcpy(b->means_d, b->means_backup_d, sizeof(float) * b->total_mean, hipMemcpyDeviceToDevice) !=
              hipSuccess &&
          st == MPPI_OK)
        st = MPPI_E_HIP;
      const float us = 1000.0f * ms / reps;
      if (st == MPPI_OK && (best_us[k] < 0.0f || us < best_us[k])) best_us[k] = us;
    }
  mppi_batch_config best{0, 0, 0, 0, 0.0f};
  for (size_t k = 0; k < shapes.size(); ++k)
    if (best_us[k] >= 0.0f && (best.threads == 0 || best_us[k] < best.iteration_us)) {
      best = shapes[k];
      best.iteration_us = best_us[k];
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != MPPI_OK) return st == MPPI_E_HIP ? fail(st, "batch autotune: HIP error") : st;
  if (best.threads == 0) return fail(MPPI_E_SIZE, "batch autotune: no launch shape fits");
  const float us = best.iteration_us;
  MPPI_CHECK(batch_apply_config(b, best));
  b->cfg.iteration_us = us;
  if (best_out) *best_out = b->cfg;
  return MPPI_OK;
}

int mppi_batch_compute_control(mppi_batch* b, const float* x0_host, int32_t stride)
{
  if (!b || !x0_host) return fail(MPPI_E_INVALID, "null argument");
  for (int i = 0; i < b->n; ++i)
    if (stride < 0 || stride > b->T[i]) return fail(MPPI_E_INVALID, "optimization_stride out of range");
  std::memcpy(b->x0_h, x0_host, sizeof(float) * BATCH_MAX_STATE * b->n);
  for (int i = 0; i < b->n; ++i) {
    b->prev_baseline[i] = b->stats_h[i].baseline;
    // the nominal is uploaded only where the device's differs from the host's sequence
    if (b->dirty[i]) {
      const size_t base = (size_t)b->mean_base[i], tc = (size_t)b->T[i] * b->C[i];
      const float* src = b->ctl_host[i] ? b->control[i].data() : b->control_out_h + base;
      std::memmove(b->nominal_h + base, src, sizeof(float) * tc);
    }
    b->upload_h[i] = b->dirty[i] ? 1 : 0;
    b->iters_h[i] = IterArgs{b->rng_counter, stride, 0};
  }
  b->rng_counter++;
  if (b->use_graph && b->eager_runs > 0) {
    if (!b->graph_exec) {
      MPPI_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int st = batch_enqueue_call(b);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (st != MPPI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      b->graph = g;
      MPPI_HIP_CHECK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    }
    MPPI_HIP_CHECK(hipGraphLaunch(b->graph_exec, b->stream));
  } else {
    MPPI_CHECK(batch_enqueue_call(b));
    b->eager_runs++;
  }
  const uint32_t want = ++b->signals;
  MPPI_HIP_CHECK(hipEventRecord(b->ev_end, b->stream));
  const int st = wait_host_signal(b->done_h, want, b->ev_end);
  if (st != MPPI_OK) {
    b->signals = *b->done_h;
    return st;
  }
  // the combine smoothed and constrained the new means on the device (into the next call's nominal as well) and
  // wrote them and the statistics to mapped host memory, where the getters read them
  std::fill(b->ctl_host.begin(), b->ctl_host.end(), 0);
  std::fill(b->raw_fresh.begin(), b->raw_fresh.end(), 0);
  std::fill(b->state_stale.begin(), b->state_stale.end(), 1);
  std::fill(b->dirty.begin(), b->dirty.end(), 0);
  b->computed = true;
  return MPPI_OK;
}

int mppi_batch_set_graph_replay(mppi_batch* b, int32_t enable)
{
  if (!b) return fail(MPPI_E_INVALID, "null argument");
  MPPI_HIP_CHECK(hipStreamSynchronize(b->stream));
  b->use_graph = enable != 0;
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  return MPPI_OK;
}

// Arguments of the device loop (once): stride 1, iteration 0, rng offsets from a range computeControl does not reach.
static int batch_loop_args_init(mppi_batch* b)
{
  if (b->loop_ready) return MPPI_OK;
  for (int i = 0; i < b->n; ++i) b->iters_h[i] = IterArgs{b->rng_counter + ((uint64_t)1 << 40), 1, 0};
  MPPI_HIP_CHECK(hipMemcpyAsync(b->loop_iters_d, b->iters_h, sizeof(IterArgs) * b->n, hipMemcpyHostToDevice,
                                b->stream));
  MPPI_HIP_CHECK(hipStreamSynchronize(b->stream));
  b->loop_ready = true;
  return MPPI_OK;
}

int mppi_batch_step_device(mppi_batch* b, int32_t num_calls)
{
  if (!b) return fail(MPPI_E_INVALID, "null argument");
  if (num_calls < 0) return fail(MPPI_E_INVALID, "num_calls must be >= 0");
  if (!b->computed)
    return fail(MPPI_E_INVALID, "batch step_device: call computeControl once first (the loop starts from its state)");
  if (num_calls == 0) return MPPI_OK;
  b->dirty.assign(b->n, 1);  // the loop moves the device means away from the host's sequences
  if (!b->loop_exec) {
    MPPI_CHECK(batch_loop_args_init(b));
    MPPI_CHECK(batch_enqueue_iteration(b, b->loop_iters_d, 1));  // eager pass before capture
    MPPI_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int st = batch_enqueue_iteration(b, b->loop_iters_d, 1);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != MPPI_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->loop_graph = g;
    MPPI_HIP_CHECK(hipGraphInstantiate(&b->loop_exec, b->loop_graph, nullptr, nullptr, 0));
    --num_calls;
  }
  for (int i = 0; i < num_calls; ++i) MPPI_HIP_CHECK(hipGraphLaunch(b->loop_exec, b->stream));
  return MPPI_OK;
}

int mppi_batch_time_kernels(mppi_batch* b, int32_t reps, float* rollout_ms, float* combine_ms)
{
  if (!b || !rollout_ms || !combine_ms || reps < 1) return fail(MPPI_E_INVALID, "bad argument");
  if (!b->computed) return fail(MPPI_E_INVALID, "batch time_kernels: call computeControl once first");
  MPPI_CHECK(batch_loop_args_init(b));
  b->dirty.assign(b->n, 1);
  std::vector<hipEvent_t> ev((size_t)(2 * reps + 1), nullptr);
  int st = MPPI_OK;
  for (auto& e : ev)
    if (st == MPPI_OK && hipEventCreate(&e) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventCreate");
  if (st == MPPI_OK && hipEventRecord(ev[0], b->stream) != hipSuccess) st = fail(MPPI_E_HIP, "hipEventRecord");
  // device-loop iterations, launch after launch: rollout between ev[2r] and ev[2r+1], combine to ev[2r+2]
  for (int r = 0; r < reps && st == MPPI_OK; ++r) {
    st = batch_enqueue_iteration(b, b->loop_iters_d, 1, ev[2 * r + 1]);
    if (st == MPPI_OK && hipEventRecord(ev[2 * r + 2], b->stream) != hipSuccess)
      st = fail(MPPI_E_HIP, "hipEventRecord");
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
  (void)hipStreamSynchronize(b->stream);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (st != MPPI_OK) return st;
  *rollout_ms = roll / reps;
  *combine_ms = comb / reps;
  return MPPI_OK;
}

static int batch_index(const mppi_batch* b, int32_t i)
{
  if (!b) return fail(MPPI_E_INVALID, "null argument");
  if (i < 0 || i >= b->n) return fail(MPPI_E_INVALID, "batch: instance index out of range");
  return MPPI_OK;
}

// The raw mean of instance i's last call, copied from the device on its first request after the call.
static int batch_raw_mean(const mppi_batch* b, int i)
{
  if (b->raw_fresh[i]) return MPPI_OK;
  MPPI_HIP_CHECK(hipStreamSynchronize(b->stream));
  MPPI_HIP_CHECK(hipMemcpy(b->raw_mean[i].data(), b->raw_d + b->mean_base[i], sizeof(float) * b->raw_mean[i].size(),
                           hipMemcpyDeviceToHost));
  b->raw_fresh[i] = 1;
  return MPPI_OK;
}

int mppi_batch_get_control_seq(const mppi_batch* b, int32_t i, float* out)
{
  MPPI_CHECK(batch_index(b, i));
  if (!out) return fail(MPPI_E_INVALID, "null argument");
  const size_t tc = b->control[i].size();
  std::memcpy(out, b->ctl_host[i] ? b->control[i].data() : b->control_out_h + b->mean_base[i], sizeof(float) * tc);
  return MPPI_OK;
}

int mppi_batch_get_raw_mean(const mppi_batch* b, int32_t i, float* out)
{
  MPPI_CHECK(batch_index(b, i));
  if (!out) return fail(MPPI_E_INVALID, "null argument");
  MPPI_CHECK(batch_raw_mean(b, i));
  std::copy(b->raw_mean[i].begin(), b->raw_mean[i].end(), out);
  return MPPI_OK;
}

int mppi_batch_get_state_seq(const mppi_batch* b, int32_t i, float* out)
{
  MPPI_CHECK(batch_index(b, i));
  if (!out) return fail(MPPI_E_INVALID, "null argument");
  if (b->state_stale[i]) {
    // computeStateTrajectory of the smoothed, not yet constrained controls (the device smoothing is the host
    // function's arithmetic) from the call's x0
    MPPI_CHECK(batch_raw_mean(b, i));
    std::vector<float> u(b->raw_mean[i]);
    smooth_control(u, b->T[i], b->C[i]);
    propagate(b->desc[i].dynamics, b->x0_h + (size_t)i * BATCH_MAX_STATE, u.data(), b->T[i], b->desc[i].dt,
              b->state_seq[i].data());
    b->state_stale[i] = 0;
  }
  std::copy(b->state_seq[i].begin(), b->state_seq[i].end(), out);
  return MPPI_OK;
}

int mppi_batch_get_weight_stats(const mppi_batch* b, int32_t i, mppi_weight_stats* out)
{
  MPPI_CHECK(batch_index(b, i));
  if (!out) return fail(MPPI_E_INVALID, "null argument");
  *out = b->stats_h[i];
  return MPPI_OK;
}

int mppi_batch_get_free_energy(const mppi_batch* b, int32_t i, mppi_free_energy* out)
{
  MPPI_CHECK(batch_index(b, i));
  if (!out) return fail(MPPI_E_INVALID, "null argument");
  mppi_free_energy fe{};
  fe.previous_baseline = b->prev_baseline[i];
  free_energy_from_stats(b->stats_h[i], b->desc[i].lambda, b->K[i], &fe);
  *out = fe;
  return MPPI_OK;
}

int mppi_batch_set_control_seq(mppi_batch* b, int32_t i, const float* u)
{
  MPPI_CHECK(batch_index(b, i));
  if (!u) return fail(MPPI_E_INVALID, "null argument");
  std::copy(u, u + b->control[i].size(), b->control[i].begin());
  b->ctl_host[i] = 1;
  b->dirty[i] = 1;
  return MPPI_OK;
}

int mppi_batch_slide_control_seq(mppi_batch* b, int32_t steps)
{
  if (!b) return fail(MPPI_E_INVALID, "null argument");
  if (steps < 0) return fail(MPPI_E_INVALID, "steps must be >= 0");
  // slideControlSequenceHelper (controllers/controller.cuh:609-621) per instance
  for (int k = 0; k < b->n; ++k) {
    if (!b->ctl_host[k])
      std::memcpy(b->control[k].data(), b->control_out_h + b->mean_base[k], sizeof(float) * b->control[k].size());
    b->ctl_host[k] = 1;
    b->dirty[k] = 1;
    const int T = b->T[k], C = b->C[k];
    const float* zero = b->desc[k].dynamics.zero_control;
    const float* scale = b->desc[k].slide_control_scale;
    std::vector<float>& u = b->control[k];
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

int mppi_batch_get_samples(const mppi_batch* b, int32_t i, float* samples_host, float* costs_host)
{
  MPPI_CHECK(batch_index(b, i));
  MPPI_HIP_CHECK(hipStreamSynchronize(b->stream));
  const BatchInstance& I = b->inst_h[i];
  const int R = 64 * b->cfg.samples_per_lane, K = I.K, T = I.T, C = I.C;
  if (samples_host) {
    std::vector<float> tmp((size_t)I.tiles * T * R * C);
    MPPI_HIP_CHECK(hipMemcpy(tmp.data(), b->ubuf_d + I.u_base, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < K; ++k) {
      const int g = k / R, r = k % R;
      for (int t = 0; t < T; ++t)
        for (int c = 0; c < C; ++c)
          samples_host[((size_t)k * T + t) * C + c] = tmp[(((size_t)g * T + t) * R + r) * C + c];
    }
  }
  if (costs_host)
    MPPI_HIP_CHECK(hipMemcpy(costs_host, b->costs_d + I.cost_base, sizeof(float) * K, hipMemcpyDeviceToHost));
  return MPPI_OK;
}

int mppi_batch_stream(const mppi_batch* b, void** stream)
{
  if (!b || !stream) return fail(MPPI_E_INVALID, "null argument");
  *stream = (void*)b->stream;
  return MPPI_OK;
}

}  // extern "C"
