// Rollout of a workgroup's samples (up to 64) over the horizon (rolloutKernel, core/mppi_common.cu:105-152): for
// t < T read u_t from the LDS control row, enforce constraints, write the clamped u_t back, step the dynamics,
// accumulate the running cost of the post-step output plus the likelihood-ratio cost; add the terminal cost; divide
// by T.
//
// Constraints act on each control alone, so the kernels apply them to the whole tile while staging it in LDS
// (constrain_control, the same arithmetic): the tile holds the constrained controls on entry, which are also the
// values written back. The work is then split across the wavefronts of the workgroup so that the serial chain only
// carries the dynamics:
//   producer  (wavefront 0, one sample per lane) -- the state chain: u_t (read one group of steps ahead), dynamics
//             step, and the output y_t stored to a double-buffered LDS chunk ybuf[t - t0][lane][O | 1];
//   consumers (wavefronts 1..15) -- the running + likelihood-ratio cost of every (sample, step) of the previous
//             chunk: thread (w, lane) takes sample `lane` and the steps t0 + w, t0 + w + 15, ...; the cost-map
//             gathers of these independent items overlap in flight.
// One barrier per chunk hands a chunk from the producer to the consumers. The per-sample sums of the consumer
// wavefronts are added in a fixed order, so the result is deterministic.
#pragma once

#include "models.hpp"
#include "nn_bf16.hpp"

namespace mppi_amd {

constexpr int RC_LANES = 64;      // samples per workgroup (lanes of the producer wavefront)
constexpr int RC_THREADS = 1024;  // producer + 15 consumer wavefronts (all of them stage the tile)
constexpr int RC_CONSUMERS = RC_THREADS / 64 - 1;

// LDS floats the pipelined rollout needs besides the control tile: the two output chunks, the consumer partial sums
// (running costs and importance-weight corrections), the terminal costs, the costs and the uncorrected costs. The
// output chunks are [step][lane][O | 1]: an odd lane stride keeps a wavefront's accesses off shared LDS banks.
template <int O>
inline size_t rollout_scratch_floats(int chunk)
{
  return 2 * (size_t)chunk * RC_LANES * (O | 1) + 2 * (size_t)RC_CONSUMERS * RC_LANES + 3 * RC_LANES;
}

// Largest chunk (a power of two, 4..32 steps) whose buffers fit in `avail` floats; 0 if none.
template <int O>
inline int rollout_chunk(size_t avail)
{
  for (int ch = 32; ch >= 4; ch >>= 1)
    if (rollout_scratch_floats<O>(ch) <= avail) return ch;
  return 0;
}

// Scratch floats a model's rollout needs for `chunk` (the BF16 network also stages its parameters there), and the
// largest chunk that fits in `avail` floats (0 if none).
template <int MODEL, int O>
inline size_t rollout_scratch_for(int chunk)
{
  const size_t s = rollout_scratch_floats<O>(chunk);
  return (MODEL == MODEL_RACER_DUBINS_NN_BF16 && s < (size_t)NN_PARAMS) ? (size_t)NN_PARAMS : s;
}
template <int MODEL, int O>
inline int rollout_chunk_for(size_t avail)
{
  for (int ch = 32; ch >= 4; ch >>= 1)
    if (rollout_scratch_for<MODEL, O>(ch) <= avail) return ch;
  return 0;
}

// Workgroup size of the rollout kernels: 16 wavefronts, except for the BF16-network model, whose producers keep
// their network fragments (about 90 registers) in registers and therefore run 8 wavefronts (2 per SIMD, 256
// registers each): 4 producer wavefronts of 16 samples and 4 consumers.
template <int MODEL>
constexpr int rollout_threads()
{
  return MODEL == MODEL_RACER_DUBINS_NN_BF16 ? 512 : RC_THREADS;
}
constexpr int NN_THREADS = rollout_threads<MODEL_RACER_DUBINS_NN_BF16>();
constexpr int NN_PRODUCERS = RC_LANES / 16;
constexpr int NN_CONSUMERS = NN_THREADS / 64 - NN_PRODUCERS;

// Importance-weight correction of a proposal sample drawn around mu + dmu instead of the base mean mu: the
// mean-shift term lambda (log N(u; mu + dmu, sig_p^2) - log N(u; mu, sig_p^2)) summed over the controls of one step,
// lambda sum_c ((u - mu)^2 - (u - mu - dmu)^2) / (2 sig_p^2), linear in u (the control-cost term of MPPI's
// derivation for a shifted sampling mean). The proposal's own spread sig_q sets the exploration width and, like the
// base sampler's std_dev_decay, carries no density ratio (a variance ratio would weight the proposal's outliers by
// exp(chi^2 / 2)). `dmu` (LDS) holds the shift per (t, c), hp[c] = 1 / (2 sig_p^2).
struct PropLds {
  const float* dmu;
  float hp[MPPI_MAX_CONTROL];
};

template <int C>
__device__ __forceinline__ float is_term(const PropLds& p, const float* u, const float* mu, int j)
{
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const float d0 = u[c] - mu[c];
    const float d1 = d0 - p.dmu[j + c];
    s += (d0 * d0 - d1 * d1) * p.hp[c];
  }
  return s;
}

// The generic pipeline below is written against two traits types:
//   M  -- S, C, O; deriv(dp, x, u, x_dot); update(dp, x, x_next, x_dot, dt); output(dp, x_next, y) (the reference's
//         computeStateDeriv / updateState / stateToOutput, dynamics/dynamics.cuh:369-432, over a parameter block dp);
//   CF -- running(cp, y, u, t) and terminal(cp, y) (computeRunningCost / terminalCost, cost_functions/cost.cu:39-53).
// BuiltinModel / BuiltinCost adapt the engine's own plugins (models.hpp); user plugins are adapted in
// plugin_kernels.hpp.
template <int MODEL>
struct BuiltinModel : Model<MODEL> {
  static_assert(Model<MODEL>::O == Model<MODEL>::S, "the built-in models output their state");
  MPPI_HD static void output(const mppi_dynamics_params&, const float* x, float* y)
  {
#pragma unroll
    for (int i = 0; i < Model<MODEL>::O; ++i) y[i] = x[i];
  }
};

template <int COST, int O>
struct BuiltinCost {
  MPPI_HD static float running(const mppi_cost_params& c, const float* y, const float*, int t)
  {
    return Cost<COST, O>::running(c, y, t);
  }
  MPPI_HD static float terminal(const mppi_cost_params& c, const float* y) { return Cost<COST, O>::terminal(c, y); }
};

// The producer's steps t0 .. t1 - 1 of one chunk: un holds the (already constrained) controls of the next PG steps,
// read one group ahead so that the LDS latency passes while the current group is stepped; each step advances the
// dynamics and stores the output to the chunk buffer yb (lane's row of step t0 at yb). Full groups run without
// per-step guards; a partial group only occurs at the end of the horizon.
template <class M, int PG, class DP>
__device__ __forceinline__ void produce_chunk(const DP& dyn, float dt, const float* row, float* yb, int t0, int t1,
                                              float* x, float* y, float (&un)[PG][M::C])
{
  constexpr int S = M::S, C = M::C, O = M::O;
  auto step = [&](const float* u, int t) {
    float xd[S], xn[S];
    M::deriv(dyn, x, u, xd);
    M::update(dyn, x, xn, xd, dt);
    M::output(dyn, xn, y);
#pragma unroll
    for (int i = 0; i < O; ++i) yb[(t - t0) * RC_LANES * (O | 1) + i] = y[i];
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = xn[i];
  };
  int tg = t0;
  for (; tg + PG <= t1; tg += PG) {
    float ug[PG][C];
#pragma unroll
    for (int j = 0; j < PG; ++j)
#pragma unroll
      for (int c = 0; c < C; ++c) ug[j][c] = un[j][c];
    // the next group's controls; past the horizon these read the row's padding and the following LDS words
    // (inside the tile or the buffers behind it), never used
#pragma unroll
    for (int j = 0; j < PG; ++j)
#pragma unroll
      for (int c = 0; c < C; ++c) un[j][c] = row[(tg + PG + j) * C + c];
#pragma unroll
    for (int j = 0; j < PG; ++j) step(ug[j], tg + j);
  }
  if (tg < t1) {
#pragma unroll
    for (int j = 0; j < PG; ++j)
      if (tg + j < t1) step(un[j], tg + j);
  }
}

// The same pipeline for MODEL_RACER_DUBINS_NN_BF16 (NN_THREADS threads): four producer wavefronts, wavefront w owning
// samples 16 w .. 16 w + 15 of the tile (lane 16 g + n, every lane group g of a sample carrying its state), each
// evaluating the network for its 16 samples per step on the matrix cores; the other wavefronts evaluate the costs.
// `scratch` must hold at least NN_PARAMS floats: the parameters are staged there before the chunk buffers use it.
template <int COST>
__device__ __forceinline__ void rollout_tile_nn(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                                                const mppi_sampler_params& sp, int T, float dt, float lambda,
                                                float alpha, const float* __restrict__ x0, float* tile, int ROW,
                                                const float* mu, int pstart, int k0, float* scratch, int chunk,
                                                float*& cost_l_out, float*& raw_l_out, const PropLds* prop)
{
  using M = Model<MODEL_RACER_DUBINS_NN_BF16>;
  using CF = Cost<COST, M::O>;
  constexpr int S = M::S, C = M::C, O = M::O;
  static_assert(O == S, "the bicycle's output is its state");
  float* ybuf = scratch;                                  // [2][chunk][64][O]
  float* part = ybuf + 2 * chunk * RC_LANES * (O | 1);          // [NN_CONSUMERS][64]
  float* part_is = part + NN_CONSUMERS * RC_LANES;        // [NN_CONSUMERS][64]
  float* term_l = part_is + NN_CONSUMERS * RC_LANES;      // [64]
  float* cost_l = term_l + RC_LANES;                      // [64]
  float* raw_l = cost_l + RC_LANES;                       // [64]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int nch = (T + chunk - 1) / chunk;
  const bool producer = wave < NN_PRODUCERS;

  // network parameters: staged in LDS (coalesced), then each producer lane builds its BF16 fragments
  for (int i = tid; i < NN_PARAMS; i += NN_THREADS) scratch[i] = dyn.nn_theta[i];
  __syncthreads();
  NnFrags F;
  if (producer) nn_load_frags(scratch, dyn.nn_output_states, lane, F);
  __syncthreads();

  const int g = lane >> 4;
  const int r = wave * 16 + (lane & 15);  // the producer lane's sample row in the tile
  float x[S], un[C];
  float* row = tile + r * ROW;
  if (producer) {
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = x0[i];
#pragma unroll
    for (int c = 0; c < C; ++c) un[c] = row[c];
  }
  float acc = 0.0f, acc_is = 0.0f;
  const bool pure = (k0 + lane) >= pstart;
  const float* crow = tile + lane * ROW;

  for (int it = 0; it <= nch; ++it) {
    if (producer && it < nch) {
      const int t0 = it * chunk, t1 = min(T, t0 + chunk);
      float* yb = ybuf + (it & 1) * chunk * RC_LANES * (O | 1) + r * (O | 1);
      for (int t = t0; t < t1; ++t) {
        float u[C];
#pragma unroll
        for (int c = 0; c < C; ++c) u[c] = un[c];
#pragma unroll
        for (int c = 0; c < C; ++c) un[c] = row[(t + 1) * C + c];
        float xd[S], xn[S];
        M::deriv(dyn, x, u, xd);
        const NnOut o = nn_eval_bf16(F, x, u, g);
#pragma unroll
        for (int s = 0; s < S; ++s) xd[s] += s < 4 ? o.lo[s] : o.hi[s - 4];
        M::update(dyn, x, xn, xd, dt);
        if (g == 0) {
#pragma unroll
          for (int i = 0; i < O; ++i) yb[(t - t0) * RC_LANES * (O | 1) + i] = xn[i];
        }
#pragma unroll
        for (int i = 0; i < S; ++i) x[i] = xn[i];
      }
      if (it == nch - 1 && g == 0) term_l[r] = CF::terminal(cost, x);
    } else if (!producer && it > 0) {
      const int t0 = (it - 1) * chunk, t1 = min(T, t0 + chunk);
      const float* yb = ybuf + ((it - 1) & 1) * chunk * RC_LANES * (O | 1) + lane * (O | 1);
      for (int t = t0 + wave - NN_PRODUCERS; t < t1; t += NN_CONSUMERS) {
        acc += CF::running(cost, yb + (t - t0) * RC_LANES * (O | 1), t) +
               lr_cost<C>(sp, crow + t * C, mu + t * C, pure, lambda, alpha, t);
        if (prop) acc_is += is_term<C>(*prop, crow + t * C, mu + t * C, t * C);
      }
    }
    __syncthreads();
  }
  if (!producer) {
    part[(wave - NN_PRODUCERS) * RC_LANES + lane] = acc;
    part_is[(wave - NN_PRODUCERS) * RC_LANES + lane] = acc_is;
  }
  __syncthreads();
  if (wave == 0) {
    float run = 0.0f, isv = 0.0f;
#pragma unroll
    for (int w = 0; w < NN_CONSUMERS; ++w) {
      run += part[w * RC_LANES + lane];
      isv += part_is[w * RC_LANES + lane];
    }
    const float raw = run / (float)T + term_l[lane] / (float)T;
    raw_l[lane] = raw;
    // the warm start (sample 0) is not a draw from the proposal and carries no density ratio
    cost_l[lane] = (prop && k0 + lane != 0) ? raw + lambda * isv : raw;
  }
  __syncthreads();
  cost_l_out = cost_l;
  raw_l_out = raw_l;
}

// Small-grid variant for MODEL_RACER_DUBINS_NN_BF16: a workgroup of NS_THREADS threads rolls out NS_ROWS = 16
// samples, so that a small K spreads over four times as many CUs. Its four wavefronts share each step's network:
// wavefront q evaluates hidden tile q (units 16 q .. 16 q + 15) of layers 1 and 2 for all 16 samples and hands its
// BF16 activations to the others through LDS (one barrier per layer); every wavefront then evaluates the output
// layer and steps the dynamics of the 16 samples itself (lane 16 g + n carries sample n). The MFMA operands and
// rounding points are those of nn_eval_bf16, so the states equal the 64-sample pipeline's bit for bit. Costs: wavefront
// 0 keeps the outputs of the last NS_BLOCK steps in LDS; after each block every lane evaluates the running and
// likelihood-ratio cost of one (sample, step) of it -- lane 16 g + n of wavefront q takes step 4 q + g of sample n.
// `scratch` must hold nn_split_scratch_floats<O>() floats.
constexpr int NS_ROWS = 16;
constexpr int NS_THREADS = 256;
constexpr int NS_BLOCK = 16;
static_assert(NS_BLOCK == NS_THREADS / NS_ROWS, "one (sample, step) of a block per lane");

template <int O>
constexpr size_t nn_split_scratch_floats()
{
  // activation exchange (2 layers x 4 tiles x 64 lanes x 4 BF16), output history, partial sums, terminal costs,
  // costs and uncorrected costs (64 each: the caller reads a full wavefront)
  constexpr size_t s = 2 * 4 * 64 * 2 + (size_t)NS_BLOCK * NS_ROWS * (O | 1) + 2 * 4 * NS_ROWS + 3 * 64;
  return s < (size_t)NN_PARAMS ? (size_t)NN_PARAMS : s;
}

template <int COST>
__device__ __forceinline__ void rollout_tile_nn_split(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                                                      const mppi_sampler_params& sp, int T, float dt, float lambda,
                                                      float alpha, const float* __restrict__ x0, float* tile, int ROW,
                                                      const float* mu, int pstart, int k0, float* scratch,
                                                      float*& cost_l_out, float*& raw_l_out, const PropLds* prop)
{
  using M = Model<MODEL_RACER_DUBINS_NN_BF16>;
  using CF = Cost<COST, M::O>;
  constexpr int S = M::S, C = M::C, O = M::O;
  static_assert(O == S, "the bicycle's output is its state");
  bf16x4* xh1 = reinterpret_cast<bf16x4*>(scratch);        // [4 tiles][64 lanes]
  bf16x4* xh2 = xh1 + 4 * 64;                                // [4 tiles][64 lanes]
  float* hist = scratch + 2 * 4 * 64 * 2;                    // [NS_BLOCK][NS_ROWS][O | 1]
  float* part = hist + NS_BLOCK * NS_ROWS * (O | 1);         // [4][NS_ROWS]
  float* part_is = part + 4 * NS_ROWS;                       // [4][NS_ROWS]
  float* term_l = part_is + 4 * NS_ROWS;                     // [64]
  float* cost_l = term_l + 64;                               // [64]
  float* raw_l = cost_l + 64;                                // [64]
  const int tid = threadIdx.x;
  const int q = tid >> 6, lane = tid & 63;
  const int g = lane >> 4, n = lane & 15;

  // network parameters: staged in LDS (coalesced), then each lane builds its wavefront's hidden tile and the output
  // layer; the staging area is reused once every lane has its fragments
  for (int i = tid; i < NN_PARAMS; i += NS_THREADS) scratch[i] = dyn.nn_theta[i];
  __syncthreads();
  bf16x8 a1, a2[2], a3[2][2];
  f32x4 b1, b2, b3[2];
  nn_load_hidden_tile(scratch, lane, q, a1, a2, b1, b2);
  nn_load_out_tiles(scratch, dyn.nn_output_states, lane, a3, b3);
  __syncthreads();

  float x[S], un[C];
  const float* row = tile + n * ROW;
#pragma unroll
  for (int i = 0; i < S; ++i) x[i] = x0[i];
#pragma unroll
  for (int c = 0; c < C; ++c) un[c] = row[c];
  const bool pure = (k0 + n) >= pstart;
  float acc = 0.0f, acc_is = 0.0f;

  for (int t = 0; t < T; ++t) {
    float u[C];
#pragma unroll
    for (int c = 0; c < C; ++c) u[c] = un[c];
    // past the horizon this reads the row's padding word and the next row (or the mean behind the tile), unused
#pragma unroll
    for (int c = 0; c < C; ++c) un[c] = row[(t + 1) * C + c];
    xh1[q * 64 + lane] = act4(mfma_bf16(a1, nn_input(x, u, g), b1));
    // the bicycle's own derivatives are computed while the activations are in flight (held here, before the
    // barrier, rather than sunk to their use)
    float xd[S], xn[S];
    M::deriv(dyn, x, u, xd);
#pragma unroll
    for (int s = 0; s < S; ++s) asm volatile("" : "+v"(xd[s]));
    __syncthreads();
    const bf16x8 h1a = cat_bf16(xh1[lane], xh1[64 + lane]);
    const bf16x8 h1b = cat_bf16(xh1[128 + lane], xh1[192 + lane]);
    asm volatile("" ::"v"(h1a), "v"(h1b));  // both reads in flight before the first wait
    xh2[q * 64 + lane] = act4(mfma_bf16(a2[1], h1b, mfma_bf16(a2[0], h1a, b2)));
    __syncthreads();
    const NnOut o = nn_output(a3, b3, cat_bf16(xh2[lane], xh2[64 + lane]), cat_bf16(xh2[128 + lane], xh2[192 + lane]));
#pragma unroll
    for (int s = 0; s < S; ++s) xd[s] += s < 4 ? o.lo[s] : o.hi[s - 4];
    M::update(dyn, x, xn, xd, dt);
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = xn[i];
    if (q == 0 && g == 0) {
#pragma unroll
      for (int i = 0; i < O; ++i) hist[((t % NS_BLOCK) * NS_ROWS + n) * (O | 1) + i] = xn[i];
    }
    if (t % NS_BLOCK == NS_BLOCK - 1 || t == T - 1) {
      // the block's history is complete; it is next written after the following step's barriers
      __syncthreads();
      const int tb = t - t % NS_BLOCK;
      const int ts = tb + 4 * q + g;
      if (ts <= t) {
        acc += CF::running(cost, hist + ((ts - tb) * NS_ROWS + n) * (O | 1), ts) +
               lr_cost<C>(sp, row + ts * C, mu + ts * C, pure, lambda, alpha, ts);
        if (prop) acc_is += is_term<C>(*prop, row + ts * C, mu + ts * C, ts * C);
      }
    }
  }
  if (q == 0 && g == 0) term_l[n] = CF::terminal(cost, x);
  // per-sample sums: the four lane groups of a wavefront, then the four wavefronts, in a fixed order
  acc += partner<4>(acc);
  acc += partner<5>(acc);
  acc_is += partner<4>(acc_is);
  acc_is += partner<5>(acc_is);
  if (g == 0) {
    part[q * NS_ROWS + n] = acc;
    part_is[q * NS_ROWS + n] = acc_is;
  }
  __syncthreads();
  if (tid < NS_ROWS) {
    float run = 0.0f, isv = 0.0f;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      run += part[w * NS_ROWS + tid];
      isv += part_is[w * NS_ROWS + tid];
    }
    const float raw = run / (float)T + term_l[tid] / (float)T;
    raw_l[tid] = raw;
    // the warm start (sample 0) is not a draw from the proposal and carries no density ratio
    cost_l[tid] = (prop && k0 + tid != 0) ? raw + lambda * isv : raw;
  }
  __syncthreads();
  cost_l_out = cost_l;
  raw_l_out = raw_l;
}

// Runs the pipelined rollout of the tile for the traits types M and CF (above); on return (after a barrier)
// cost_l[r] holds the cost of sample k0 + r for every r < 64 (garbage for r >= nrows) and raw_l[r] the same cost
// without the importance-weight correction (equal to it without a proposal). `scratch` provides
// rollout_scratch_floats<O>(chunk) floats. A tile of fewer than 64 rows (`rows`) has its last row repeated for the
// lanes beyond it.
template <class M, class CF, class DP, class CP>
__device__ __forceinline__ void rollout_tile_t(const DP& dyn, const CP& cost, const mppi_sampler_params& sp, int T,
                                               float dt, float lambda, float alpha, const float* __restrict__ x0,
                                               float* tile, int ROW, const float* mu, int pstart, int k0,
                                               float* scratch, int chunk, float*& cost_l_out, float*& raw_l_out,
                                               const PropLds* prop = nullptr, int rows = RC_LANES)
{
  constexpr int S = M::S, C = M::C, O = M::O;
  float* ybuf = scratch;                                  // [2][chunk][64][O]
  float* part = ybuf + 2 * chunk * RC_LANES * (O | 1);          // [15][64]
  float* part_is = part + RC_CONSUMERS * RC_LANES;        // [15][64]
  float* term_l = part_is + RC_CONSUMERS * RC_LANES;      // [64]
  float* cost_l = term_l + RC_LANES;                      // [64]
  float* raw_l = cost_l + RC_LANES;                       // [64]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int nch = (T + chunk - 1) / chunk;

  // producer state; un holds the controls of the next group of PG steps (chunks are multiples of PG)
  constexpr int PG = 4;
  float x[S], y[O], un[PG][C];
  const int lrow = min(lane, rows - 1);
  float* row = tile + lrow * ROW;
  if (wave == 0) {
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = x0[i];
    M::output(dyn, x, y);
#pragma unroll
    for (int j = 0; j < PG; ++j)
#pragma unroll
      for (int c = 0; c < C; ++c) un[j][c] = row[j * C + c];
  }
  // consumer state
  float acc = 0.0f, acc_is = 0.0f;
  const bool pure = (k0 + lane) >= pstart;
  const float* crow = tile + lrow * ROW;

  for (int it = 0; it <= nch; ++it) {
    if (wave == 0 && it < nch) {
      const int t0 = it * chunk, t1 = min(T, t0 + chunk);
      float* yb = ybuf + (it & 1) * chunk * RC_LANES * (O | 1) + lane * (O | 1);
      produce_chunk<M, PG>(dyn, dt, row, yb, t0, t1, x, y, un);
      if (it == nch - 1) term_l[lane] = CF::terminal(cost, y);
    } else if (wave > 0 && it > 0) {
      const int t0 = (it - 1) * chunk, t1 = min(T, t0 + chunk);
      const float* yb = ybuf + ((it - 1) & 1) * chunk * RC_LANES * (O | 1) + lane * (O | 1);
      for (int t = t0 + wave - 1; t < t1; t += RC_CONSUMERS) {
        // y and u are read in place from LDS (runtime-indexed cost terms become LDS addresses)
        acc += CF::running(cost, yb + (t - t0) * RC_LANES * (O | 1), crow + t * C, t) +
               lr_cost<C>(sp, crow + t * C, mu + t * C, pure, lambda, alpha, t);
        if (prop) acc_is += is_term<C>(*prop, crow + t * C, mu + t * C, t * C);
      }
    }
    __syncthreads();
  }
  if (wave > 0) {
    part[(wave - 1) * RC_LANES + lane] = acc;
    part_is[(wave - 1) * RC_LANES + lane] = acc_is;
  }
  __syncthreads();
  if (wave == 0) {
    float run = 0.0f, isv = 0.0f;
#pragma unroll
    for (int w = 0; w < RC_CONSUMERS; ++w) {
      run += part[w * RC_LANES + lane];
      isv += part_is[w * RC_LANES + lane];
    }
    const float raw = run / (float)T + term_l[lane] / (float)T;
    raw_l[lane] = raw;
    // the warm start (sample 0) is not a draw from the proposal and carries no density ratio
    cost_l[lane] = (prop && k0 + lane != 0) ? raw + lambda * isv : raw;
  }
  __syncthreads();
  cost_l_out = cost_l;
  raw_l_out = raw_l;
}

// The pipelined rollout of a built-in (model, cost) pair; the BF16-network model has its own pipeline.
template <int MODEL, int COST>
__device__ __forceinline__ void rollout_tile(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                                             const mppi_sampler_params& sp, int T, float dt, float lambda,
                                             float alpha, const float* __restrict__ x0, float* tile, int ROW,
                                             const float* mu, int pstart, int k0, float* scratch, int chunk,
                                             float*& cost_l_out, float*& raw_l_out, const PropLds* prop = nullptr,
                                             int rows = RC_LANES)
{
  if constexpr (MODEL == MODEL_RACER_DUBINS_NN_BF16) {
    rollout_tile_nn<COST>(dyn, cost, sp, T, dt, lambda, alpha, x0, tile, ROW, mu, pstart, k0, scratch, chunk,
                          cost_l_out, raw_l_out, prop);
  } else {
    using M = BuiltinModel<MODEL>;
    rollout_tile_t<M, BuiltinCost<COST, M::O>>(dyn, cost, sp, T, dt, lambda, alpha, x0, tile, ROW, mu, pstart, k0,
                                               scratch, chunk, cost_l_out, raw_l_out, prop, rows);
  }
}

}  // namespace mppi_amd
