// Fused MPPI iteration for the controller path:
//   rollout_update_kernel  -- per workgroup of samples (16, 32 or 64: the smallest with at most one workgroup per
//                             CU, else 64; 16 for the small-grid BF16-network variant, 64 for its large grids; the
//                             colored sampler's 64 / C series for in-place colored samples): the control tile is
//                             drawn into LDS (Gaussian from Philox, colored by the sampler's workgroup body) or loaded
//                             from the sample buffer, the samples are rolled out over the horizon, then the workgroup
//                             reduces its clamped tile against a local baseline into (b_g, argmin_g, eta_g,
//                             sum w^2_g, S_g[T*C]).
//   combine_kernel         -- log-sum-exp combination of the partials: global baseline (first strict minimum),
//                             FP64 normalizer, new mean S / eta.
// The result equals fullGPUcomputeWeights + weightedReductionKernelDeviceNorm (core/mppi_common.cu:746-773,
// 1067-1089) on the same costs: sum_k exp(-(c_k - b)/lambda) u_k = sum_g exp(-(b_g - b)/lambda) S_g.
#include <algorithm>
#include <atomic>
#include <climits>

#include "colored_core.hpp"
#include "fused.hpp"
#include "iteration_core.hpp"
#include "models.hpp"
#include "rollout_core.hpp"
#include "tile.hpp"

namespace mppi_amd {

__device__ __forceinline__ IterArgs load_iter_f(const IterRef& r) { return load_iter_args(r); }

// Samples per workgroup and threads per workgroup of the fused kernel; SPLIT selects the small-grid BF16-network
// rollout (rollout_tile_nn_split).
template <bool SPLIT>
constexpr int ru_rows()
{
  return SPLIT ? NS_ROWS : RC_LANES;
}
template <int MODEL, bool SPLIT>
constexpr int ru_threads()
{
  return SPLIT ? NS_THREADS : rollout_threads<MODEL>();
}

// sized for the smaller workgroup tile, which gives the most partials
size_t fused_iteration_workspace_bytes(int K, int T, int C)
{
  const size_t G = (size_t)(K + NS_ROWS - 1) / NS_ROWS;
  return G * sizeof(Partial) + G * (size_t)T * C * sizeof(float);
}

// With PROP the samples come from the learned proposal: u = mu + dmu + sig_q eps (k >= 1; sample 0 and t < stride
// keep mu; no pure-noise samples) and every cost carries the mean-shift correction (rollout_core.hpp). With CNTT > 0
// the workgroup's spw samples are colored noise generated in place (colored_stage with NTT = CNTT: the same series,
// arithmetic and values as colored_mfma_kernel's workgroup `blockIdx.x`) and written into the control tile.
template <int MODEL, int COST, bool GEN, bool WRITE, bool PROP, bool SPLIT, int CNTT>
__global__ void __launch_bounds__((ru_threads<MODEL, SPLIT>())) rollout_update_kernel(
    mppi_dynamics_params dyn, mppi_cost_params cost, mppi_sampler_params sp, int K, int T, float dt, float lambda,
    float alpha, const float* __restrict__ x0, const float* __restrict__ mean, float* __restrict__ samples,
    IterRef itr, int pstart, float* __restrict__ costs_out, Partial* __restrict__ parts, float* __restrict__ sums,
    int chunk, const float* __restrict__ prop_dmu, const float* __restrict__ prop_sig, ColoredPlan cpl, int rows_arg)
{
  using M = Model<MODEL>;
  constexpr int C = M::C;
  constexpr int NT = ru_threads<MODEL, SPLIT>();
  constexpr int ROWS = ru_rows<SPLIT>();
  constexpr int MV = 1024 / NT;  // mean elements per thread: T C <= MPPI_MAX_TIMESTEPS MPPI_MAX_CONTROL = 1024
  static_assert(MV * NT == 1024 && MPPI_MAX_TIMESTEPS * MPPI_MAX_CONTROL <= 1024, "mean staging");
  static_assert(CNTT == 0 || (NT == CTHREADS && !GEN && !PROP && !SPLIT), "in-place colored samples");
  extern __shared__ float lds[];
  const int TC = T * C;
  const int TC4 = (TC + 3) & ~3;
  const int ROW = TC + 1;
  // samples per workgroup: the colored shape's, the split variant's 16, else the launcher's choice (16, 32 or 64)
  const int rows = CNTT > 0 ? cpl.sh.spw : (SPLIT ? ROWS : rows_arg);
  float* tile = lds;
  float* mu = tile + ((rows * ROW + 3) & ~3);
  float* pl = mu + TC4;                    // PROP: dmu, sig_q [2][TC4]
  float* wl = pl + (PROP ? 2 * TC4 : 0);   // [64] local weights
  float* scratch = wl + RC_LANES;          // pipelined-rollout buffers
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int g = blockIdx.x;
  const int k0 = g * rows;
  const int nrows = min(rows, K - k0);
  // the mean and x0 are requested first: their loads then retire with the tile's, whose waits come later; x0 is
  // staged in the local-weight slots, which are free until the rollout has read it. x0 may live in mapped host
  // memory, so only wavefront 0 reads it (one request per workgroup).
  float mv[MV];
  if constexpr (CNTT == 0) {
#pragma unroll
    for (int i = 0; i < MV; ++i) mv[i] = mean[min(tid + i * NT, TC - 1)];
  }
  const float xv = tid < 64 ? x0[min(tid, M::S - 1)] : 0.0f;
  float* x0_l = wl;
  bool deadband = false;
#pragma unroll
  for (int c = 0; c < C; ++c) deadband |= dyn.control_deadband[c] != 0.0f;
  PropLds pv{};
  if (PROP) {
    const float decay_p = powf(sp.std_dev_decay, (float)load_iter_f(itr).iteration);
    for (int j = tid; j < TC; j += NT) {
      pl[j] = prop_dmu[j];
      pl[TC4 + j] = prop_sig[j];
    }
    pv.dmu = pl;
#pragma unroll
    for (int c = 0; c < MPPI_MAX_CONTROL; ++c) {
      const float sp_c = decay_p * sp.std_dev[c];
      pv.hp[c] = c < C ? 0.5f / (sp_c * sp_c) : 0.0f;
    }
  }
  if (GEN) {
    // u = transform(eps) for the tile's elements; the tile is a contiguous span starting at element k0*T*C, which
    // is a multiple of 4, so Philox block b covers tile elements 4(b - base) .. 4(b - base) + 3.
#pragma unroll
    for (int i = 0; i < MV; ++i)
      if (tid + i * NT < TC) mu[tid + i * NT] = mv[i];
    if (tid < M::S) x0_l[tid] = xv;
    __syncthreads();
    gaussian_tile_fill<C, NT, PROP>(sp, load_iter_f(itr), k0, nrows, TC, ROW, pstart, mu, pl, TC4, tile,
                                    [&](int col, float v) { return constrain_control(dyn, col % C, v, deadband); });
  } else if constexpr (CNTT > 0) {
    // colored samples of this workgroup's span, generated in the scratch area (the rollout reuses it afterwards),
    // constrained into the tile; the mean comes with them
    const ColoredTile ct = colored_stage<CNTT>(sp, K, T, C, cpl.sh.n, cpl.sh.F, cpl.sh.Kpad, cpl.sh.KPH, cpl.sh.MT,
                                               cpl.sh.S, cpl.sh.spw, cpl.sh.TCS, mean, itr, nullptr, cpl.sig, cpl.bhi,
                                               cpl.blo, g, scratch);
    // thread (row group, column): column col of rows r0, r0 + RG, ... (RG = NT / TC row groups), so that the mean,
    // the component and the stride test are per thread
    const int RG = NT / TC;
    const int r0 = tid / TC, col = tid - r0 * TC;
    if (r0 < RG) {
      const float mu_c = ct.mean_l[col];
      const bool head = col < ct.sc;
      auto fill = [&](auto f) {
        for (int r = r0; r < nrows; r += RG)
          tile[r * ROW + col] = f(col, colored_transform(mu_c, ct.stage[r * ct.TCS + col], k0 + r, head, pstart));
      };
      if (deadband) fill(ControlConstrain<C>{&dyn});
      else fill(ControlClamp<C>(dyn));
      if (r0 == 0) mu[col] = mu_c;
    }
    if (tid < M::S) x0_l[tid] = xv;
  } else {
    if (deadband) tile_load<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid, ControlConstrain<C>{&dyn});
    else tile_load<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid, ControlClamp<C>(dyn));
#pragma unroll
    for (int i = 0; i < MV; ++i)
      if (tid + i * NT < TC) mu[tid + i * NT] = mv[i];
    if (tid < M::S) x0_l[tid] = xv;
  }
  __syncthreads();
  float* cost_l;
  float* raw_l;
  if constexpr (SPLIT)
    rollout_tile_nn_split<COST>(dyn, cost, sp, T, dt, lambda, alpha, x0_l, tile, ROW, mu, pstart, k0, scratch, cost_l,
                                raw_l, PROP ? &pv : nullptr);
  else
    rollout_tile<MODEL, COST>(dyn, cost, sp, T, dt, lambda, alpha, x0_l, tile, ROW, mu, pstart, k0, scratch, chunk,
                              cost_l, raw_l, PROP ? &pv : nullptr, rows);
  // with a proposal, sample 0 (the warm start) is only a probe for the fallback test: it is not a draw from the
  // proposal and takes no part in the weights
  tile_partials<NT, WRITE, PROP>(cost_l, raw_l, wl, tile, ROW, TC, k0, nrows, g, lambda, costs_out, parts, sums,
                                 samples);
}

// Combination of G partials: workgroup q owns COLS consecutive columns, read by THREADS / COLS row groups; every
// workgroup recomputes the global baseline and the scales e_g = exp(-(b_g - b)/lambda) (G <= MAX_PARTIALS values),
// workgroup 0 also writes the statistics. Each thread's first COMBINE_PREFETCH column values are requested before the
// partials, so that their latency overlaps the baseline and normalizer reductions. The kernel is latency-bound (one
// dependent memory round trip, two block reductions): the launcher takes 8 columns on 256 threads up to 256 partials
// and 16 columns on 512 threads above, the fastest measured shapes; wider column segments on more threads fetch fewer
// bytes (each row line by fewer workgroups) but take longer.
constexpr int COMBINE_PREFETCH = 8;
constexpr int NUM_XCDS = 8;

// Workgroup w of a grid of n is dispatched to XCD w % 8 (round-robin). The logical block it takes is chosen so that
// each XCD gets a contiguous run of blocks -- XCD x takes blocks start(x) .. start(x) + count(x) - 1 -- and the
// workgroups whose column segments share a cache line of the partial rows sit behind the same L2.
__device__ __forceinline__ int xcd_contiguous_block(int w, int n)
{
  const int x = w % NUM_XCDS, slot = w / NUM_XCDS;
  const int base = n / NUM_XCDS, extra = n % NUM_XCDS;
  return x * base + min(x, extra) + slot;
}

template <int COLS, int THREADS>
__global__ void __launch_bounds__(THREADS) combine_kernel(const Partial* __restrict__ parts,
                                                          const float* __restrict__ sums, int G, int K, int TC,
                                                          float lambda, mppi_weight_stats* __restrict__ stats,
                                                          float* __restrict__ mean_out, IterArgs* advance,
                                                          int advance_step, ResultCopies copies, int pstride)
{
  // system blockIdx.y of a multi-system combine: its partials, rows and results lie one stride further on
  const int sys = blockIdx.y;
  parts += (size_t)sys * pstride;
  sums += (size_t)sys * pstride * TC;
  stats += sys;
  mean_out += (size_t)sys * TC;
  if (copies.mean != nullptr) copies.mean += (size_t)sys * TC;
  if (copies.stats != nullptr) copies.stats += sys;
  constexpr int COMBINE_COLS = COLS;
  constexpr int COMBINE_THREADS = THREADS;
  __shared__ float escale[MAX_PARTIALS];
  __shared__ float sv[COMBINE_THREADS / 64];
  __shared__ int si[COMBINE_THREADS / 64];
  __shared__ double se[COMBINE_THREADS / 64], se2[COMBINE_THREADS / 64], sec[COMBINE_THREADS / 64];
  __shared__ float acc_l[COMBINE_THREADS / COMBINE_COLS][COMBINE_COLS];
  static_assert(MAX_PARTIALS % COMBINE_THREADS == 0, "partials per thread");
  constexpr int PER = MAX_PARTIALS / COMBINE_THREADS;
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  const double inv = 1.0 / (double)lambda;
  // columns of this workgroup: thread (grp, col) sums partials g = grp, grp + NG, ...
  constexpr int NG = COMBINE_THREADS / COMBINE_COLS;
  const int col = tid % COMBINE_COLS, grp = tid / COMBINE_COLS;
  const int j = xcd_contiguous_block(blockIdx.x, gridDim.x) * COMBINE_COLS + col;
  float pre[COMBINE_PREFETCH];
#pragma unroll
  for (int i = 0; i < COMBINE_PREFETCH; ++i) {
    const int q = grp + i * NG;
    pre[i] = (j < TC && q < G) ? sums[(size_t)q * TC + j] : 0.0f;
  }
  Partial mine[PER];
  float v = INFINITY;
  int idx = INT_MAX;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int q = tid + i * COMBINE_THREADS;
    if (q < G) {
      mine[i] = parts[q];
      argmin_combine(v, idx, mine[i].baseline, mine[i].argmin);
    }
  }
  wave_argmin(v, idx);
  if (lane == 0) {
    sv[wid] = v;
    si[wid] = idx;
  }
  __syncthreads();
  float b = sv[0];
  int bi = si[0];
  for (int q = 1; q < COMBINE_THREADS / 64; ++q) argmin_combine(b, bi, sv[q], si[q]);
  if (bi == INT_MAX) bi = 0;
  double e = 0.0, e2 = 0.0, ec = 0.0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int q = tid + i * COMBINE_THREADS;
    if (q < G) {
      const double sc = (mine[i].baseline <= FLT_MAX) ? exp(-inv * ((double)mine[i].baseline - (double)b)) : 0.0;
      escale[q] = (float)sc;
      e += mine[i].eta * sc;
      e2 += mine[i].s2 * sc * sc;
      ec += mine[i].wc * sc;
    }
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
  for (int q = 0; q < COMBINE_THREADS / 64; ++q) {
    eta += se[q];
    s2 += se2[q];
    wcost += sec[q];
  }
  float acc = 0.0f;
  if (j < TC) {
#pragma unroll
    for (int i = 0; i < COMBINE_PREFETCH; ++i) {
      const int q = grp + i * NG;
      if (q < G) acc += escale[q] * pre[i];
    }
#pragma unroll 8
    for (int q = grp + COMBINE_PREFETCH * NG; q < G; q += NG) acc += escale[q] * sums[(size_t)q * TC + j];
  }
  acc_l[grp][col] = acc;
  __syncthreads();
  if (grp == 0 && j < TC) {
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < NG; ++q) s += acc_l[q][col];
    const float m = (float)((double)s / eta);
    mean_out[j] = m;
    if (copies.mean != nullptr) copies.mean[j] = m;
  }
  if (blockIdx.x == 0 && tid == 0) {
    mppi_weight_stats st;
    st.baseline = b;
    st.argmin = bi;
    st.normalizer = eta;
    st.sum_w2 = s2;
    st.lambda = lambda;
    st.num_rollouts = K;
    st.weighted_cost = wcost / eta;
    *stats = st;
    if (copies.stats != nullptr) *copies.stats = st;
    if (sys == 0 && copies.cost0 != nullptr) *copies.cost0 = *copies.cost0_src;
    if (sys == 0 && advance != nullptr) advance->rng_offset += (uint64_t)advance_step;
  }
  if (copies.signal.flag != nullptr && last_block_arrive(copies.signal)) {
    // the last workgroup, which sees every column of the new mean: computeControl's post-processing
    // (controllers/controller.cuh:573-607 smoothing, controllers/MPPI/mppi_controller.cu:224-227 constraints) into
    // the host's copy and the next call's nominal
    if (copies.control != nullptr) {
      for (int q = tid; q < TC; q += COMBINE_THREADS) {
        const int t = q / copies.C, c = q - t * copies.C;
        const float u = constrain_with(copies.limits, c, sg_smooth_at(mean_out, t, c, copies.T, copies.C));
        copies.control[q] = u;
        copies.nominal[q] = u;
      }
    }
    signal_publish(copies.signal);
  }
}

// Compute units of the current device (cached per device; 0 if the query fails).
static int device_cu_count()
{
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  int n = cus[dev].load(std::memory_order_relaxed);
  if (n == 0 && hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
    cus[dev].store(n, std::memory_order_relaxed);
  return n;
}

// LDS floats of the fused kernel for `rows` samples per workgroup: the control tile (rounded to 16 bytes), the mean
// (and the proposal's shift and spread), the local weights, then the rollout scratch, which for in-place colored
// samples also holds the colored stage (colored_bytes). 0 when it does not fit; *chunk_out gets the rollout chunk.
template <int MODEL, bool PROP, bool SPLIT>
static size_t ru_lds_floats(int T, int rows, size_t colored_bytes, int* chunk_out)
{
  constexpr int C = Model<MODEL>::C, O = Model<MODEL>::O;
  const size_t TC4 = (size_t)((T * C + 3) & ~3);
  const size_t tile_f = ((size_t)rows * (T * C + 1) + 3) & ~(size_t)3;
  const size_t base = tile_f + TC4 * (PROP ? 3 : 1) + RC_LANES;
  const size_t limit = LDS_LIMIT / sizeof(float);
  *chunk_out = 0;
  if (base >= limit) return 0;
  if constexpr (SPLIT) {
    const size_t total = base + nn_split_scratch_floats<O>();
    return total <= limit ? total : 0;
  } else {
    const int chunk = rollout_chunk_for<MODEL, O>(limit - base);
    if (chunk == 0) return 0;
    const size_t scratch = std::max(rollout_scratch_for<MODEL, O>(chunk), (colored_bytes + 15) / 16 * 4);
    if (base + scratch > limit) return 0;
    *chunk_out = chunk;
    return base + scratch;
  }
}

// Whether the fused kernel can generate the colored samples in place: 1024-thread workgroups (every model but the
// BF16 network), the 64-series colored shape with at least 16 samples per workgroup (the partial workspace is sized
// for 16), and the LDS for tile + colored stage.
static bool colored_inline_fits(int model, int T, const ColoredPlan& pl)
{
  if (pl.sh.NS != 64 || pl.sh.spw < NS_ROWS) return false;
  int chunk = 0;
  const int rows = pl.sh.spw;
  const size_t cb = pl.sh.lds_bytes;
  switch (model) {
    case MPPI_MODEL_CARTPOLE: return ru_lds_floats<MPPI_MODEL_CARTPOLE, false, false>(T, rows, cb, &chunk) > 0;
    case MPPI_MODEL_DOUBLE_INTEGRATOR:
      return ru_lds_floats<MPPI_MODEL_DOUBLE_INTEGRATOR, false, false>(T, rows, cb, &chunk) > 0;
    case MPPI_MODEL_RACER_DUBINS: return ru_lds_floats<MPPI_MODEL_RACER_DUBINS, false, false>(T, rows, cb, &chunk) > 0;
    case MPPI_MODEL_RACER_DUBINS_NN:
      return ru_lds_floats<MPPI_MODEL_RACER_DUBINS_NN, false, false>(T, rows, cb, &chunk) > 0;
    default: return false;
  }
}

template <int MODEL, int COST, bool GEN, bool WRITE, bool PROP, bool SPLIT, int CNTT>
static int launch_ru(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp,
                     int K, int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                     float* samples, IterRef it, float* costs_out, Partial* parts, float* sums, const float* prop_dmu,
                     const float* prop_sig, const ColoredPlan* cpl, int rows_sel, hipStream_t s)
{
  const int rows = CNTT > 0 ? cpl->sh.spw : (SPLIT ? ru_rows<SPLIT>() : rows_sel);
  int chunk = 0;
  const size_t lf = ru_lds_floats<MODEL, PROP, SPLIT>(T, rows, CNTT > 0 ? cpl->sh.lds_bytes : 0, &chunk);
  if (lf == 0) return fail(MPPI_E_SIZE, "fused rollout: horizon too long for the LDS control tile");
  MPPI_CHECK(allow_max_dynamic_lds((const void*)rollout_update_kernel<MODEL, COST, GEN, WRITE, PROP, SPLIT, CNTT>));
  const int pstart = PROP ? K : pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  const int G = (K + rows - 1) / rows;
  rollout_update_kernel<MODEL, COST, GEN, WRITE, PROP, SPLIT, CNTT>
      <<<G, ru_threads<MODEL, SPLIT>(), sizeof(float) * lf, s>>>(dyn, cost, sp, K, T, dt, lambda, alpha, x0, mean,
                                                                 samples, it, pstart, costs_out, parts, sums, chunk,
                                                                 prop_dmu, prop_sig,
                                                                 CNTT > 0 ? *cpl : ColoredPlan{}, rows);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

template <int MODEL, int COST, bool SPLIT, int CNTT>
static int dispatch_flags(bool gen, bool write, const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                          const mppi_sampler_params& sp, int K, int T, float dt, float lambda, float alpha,
                          const float* x0, const float* mean, float* samples, IterRef it, float* costs_out,
                          Partial* parts, float* sums, const float* prop_dmu, const float* prop_sig,
                          const ColoredPlan* cpl, int rows, hipStream_t s)
{
#define MPPI_LAUNCH_RU(GEN, WRITE, PROP)                                                                               \
  launch_ru<MODEL, COST, GEN, WRITE, PROP, SPLIT, CNTT>(dyn, cost, sp, K, T, dt, lambda, alpha, x0, mean, samples, it, \
                                                        costs_out, parts, sums, prop_dmu, prop_sig, cpl, rows, s)
  if constexpr (CNTT > 0) {
    return write ? MPPI_LAUNCH_RU(false, true, false) : MPPI_LAUNCH_RU(false, false, false);
  } else {
    if (prop_dmu != nullptr) {
      if (!gen) return fail(MPPI_E_INVALID, "fused iteration: the learned proposal draws Gaussian samples");
      return write ? MPPI_LAUNCH_RU(true, true, true) : MPPI_LAUNCH_RU(true, false, true);
    }
    if (gen) return write ? MPPI_LAUNCH_RU(true, true, false) : MPPI_LAUNCH_RU(true, false, false);
    return write ? MPPI_LAUNCH_RU(false, true, false) : MPPI_LAUNCH_RU(false, false, false);
  }
#undef MPPI_LAUNCH_RU
}

int launch_fused_iteration(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                           const mppi_sampler_params& sp, int K, int T, float dt, float lambda, float alpha,
                           const float* x0, const float* mean, float* samples, bool generate, bool write_samples,
                           IterRef it, float* costs_out, void* ws, mppi_weight_stats* stats, float* mean_out,
                           IterArgs* advance, hipStream_t s, const float* prop_dmu, const float* prop_sig,
                           const ResultCopies* copies, const ColoredPlan* colored, int advance_step,
                           hipEvent_t between, int* partials_out)
{
  const int C = dyn.control_dim;
  const int model = (dyn.model == MPPI_MODEL_RACER_DUBINS_NN && dyn.nn_precision == MPPI_NN_BF16)
                        ? MODEL_RACER_DUBINS_NN_BF16
                        : dyn.model;
  // colored samples: generated inside the rollout kernel where its workgroup shape allows, else by the colored
  // sampler into `samples` first
  bool inline_colored = false;
  if (colored != nullptr) {
    if (generate || prop_dmu != nullptr)
      return fail(MPPI_E_INVALID, "fused iteration: colored samples exclude generated and proposal samples");
    inline_colored = colored_inline_fits(model, T, *colored);
    if (!inline_colored) MPPI_CHECK(launch_sample_colored(sp, K, T, C, mean, samples, it, nullptr, s));
  }
  // the BF16-network rollout is bound by the latency of its per-step chain: while every CU can take one 16-sample
  // workgroup (one wavefront per SIMD), splitting each step's network over four wavefronts finishes sooner; beyond
  // that, the 64-sample pipeline does the same work with a quarter of the replicated dynamics
  const bool split = model == MODEL_RACER_DUBINS_NN_BF16 && (K + NS_ROWS - 1) / NS_ROWS <= device_cu_count();
  // samples per workgroup: the colored shape's or the split variant's; otherwise (1024-thread workgroups) the smallest
  // of 64, 32, 16 with at most one workgroup per CU -- the state chain takes as long for 16 samples as for 64, while
  // drawing the tile and reducing it scale with the samples
  int rows = RC_LANES;
  if (inline_colored) {
    rows = colored->sh.spw;
  } else if (split) {
    rows = NS_ROWS;
  } else if (model != MODEL_RACER_DUBINS_NN_BF16) {
    rows = fused_rows_for(K, device_cu_count());
  }
  const int G = (K + rows - 1) / rows;
  if (G > MAX_PARTIALS) return fail(MPPI_E_SIZE, "fused iteration: too many samples");
  if (partials_out != nullptr) *partials_out = G;
  Partial* parts = static_cast<Partial*>(ws);
  float* sums = reinterpret_cast<float*>(static_cast<char*>(ws) + (size_t)G * sizeof(Partial));
  int st = MPPI_E_UNSUPPORTED;
#define MPPI_FUSED_ARGS                                                                                         \
  generate, write_samples, dyn, cost, sp, K, T, dt, lambda, alpha, x0, mean, samples, it, costs_out, parts, sums, \
      prop_dmu, prop_sig, colored, rows, s
#define MPPI_FUSED_CASE(MODEL, COST)                                                                   \
  if (model == MODEL && cost.cost == COST) {                                                           \
    constexpr bool NN16 = MODEL == MODEL_RACER_DUBINS_NN_BF16;                                         \
    if (inline_colored) st = dispatch_flags<MODEL, COST, false, NN16 ? 0 : 4>(MPPI_FUSED_ARGS);        \
    else if (split) st = dispatch_flags<MODEL, COST, NN16, 0>(MPPI_FUSED_ARGS);                        \
    else st = dispatch_flags<MODEL, COST, false, 0>(MPPI_FUSED_ARGS);                                  \
  }
  MPPI_FUSED_CASE(MPPI_MODEL_CARTPOLE, MPPI_COST_CARTPOLE_QUADRATIC)
  MPPI_FUSED_CASE(MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_QUADRATIC_MAP)
  MPPI_FUSED_CASE(MPPI_MODEL_RACER_DUBINS, MPPI_COST_QUADRATIC_MAP)
  MPPI_FUSED_CASE(MPPI_MODEL_RACER_DUBINS_NN, MPPI_COST_QUADRATIC_MAP)
  MPPI_FUSED_CASE(MODEL_RACER_DUBINS_NN_BF16, MPPI_COST_QUADRATIC_MAP)
#undef MPPI_FUSED_CASE
#undef MPPI_FUSED_ARGS
  if (st == MPPI_E_UNSUPPORTED)
    return fail(MPPI_E_UNSUPPORTED, "fused iteration: model/cost combination is not compiled into the kernel library");
  if (st != MPPI_OK) return st;
  if (between != nullptr) MPPI_HIP_CHECK(hipEventRecord(between, s));
  return launch_combine(parts, sums, G, K, T * C, lambda, stats, mean_out, advance, advance_step, copies, s);
}

int launch_combine(const Partial* parts, const float* sums, int G, int K, int TC, float lambda,
                   mppi_weight_stats* stats, float* mean_out, IterArgs* advance, int advance_step,
                   const ResultCopies* copies, hipStream_t s, int systems, int pstride)
{
  if (G < 1 || G > MAX_PARTIALS) return fail(MPPI_E_SIZE, "combine: partial count out of range");
  if (systems < 1 || systems > 2 || (systems > 1 && pstride < G))
    return fail(MPPI_E_INVALID, "combine: one or two systems, partial stride >= partial count");
  const ResultCopies rc = copies != nullptr ? *copies : ResultCopies{};
  if (systems > 1 && rc.control != nullptr)
    return fail(MPPI_E_INVALID, "combine: the smoothed-control copy takes one system");
  if (G <= 256)
    combine_kernel<8, 256><<<dim3((TC + 7) / 8, systems), 256, 0, s>>>(parts, sums, G, K, TC, lambda, stats, mean_out,
                                                                       advance, advance_step, rc, pstride);
  else
    combine_kernel<16, 512><<<dim3((TC + 15) / 16, systems), 512, 0, s>>>(parts, sums, G, K, TC, lambda, stats,
                                                                          mean_out, advance, advance_step, rc, pstride);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

}  // namespace mppi_amd
