// HIP kernels of the MPPI iteration for gfx950: normal generation, Gaussian and colored-noise control sampling,
// the LDS-staged fused rollout, the weight transform and the weighted control update.
#include <climits>

#include "kernels.hpp"
#include "models.hpp"
#include "rollout_core.hpp"
#include "tile.hpp"

namespace mppi_amd {

__device__ __forceinline__ IterArgs load_iter(const IterRef& r) { return r.ptr ? *r.ptr : r.val; }

static int grid_for(int64_t work, int block, int max_blocks)
{
  int64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (int)g;
}

// =============================================================================================== normals (Philox)
template <bool ALIGNED>
__global__ void __launch_bounds__(256) normals_kernel(uint64_t seed, uint64_t offset, float* __restrict__ out,
                                                      int64_t n)
{
  const int64_t nb = (n + 3) / 4;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
    const Normal4 z = normals4(seed, offset, (uint64_t)b);
    const int64_t e = b * 4;
    if (ALIGNED && e + 3 < n) {
      *reinterpret_cast<float4*>(out + e) = make_float4(z.z[0], z.z[1], z.z[2], z.z[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (e + j < n) out[e + j] = z.z[j];
    }
  }
}

int launch_generate_normals(uint64_t seed, uint64_t offset, float* out, int64_t n, hipStream_t s)
{
  if (n <= 0) return MPPI_OK;
  const int grid = grid_for((n + 3) / 4, 256, 8192);
  if (((uintptr_t)out & 15) == 0) normals_kernel<true><<<grid, 256, 0, s>>>(seed, offset, out, n);
  else normals_kernel<false><<<grid, 256, 0, s>>>(seed, offset, out, n);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// ========================================================================================= Gaussian control samples
// setGaussianControls (sampling_distributions/gaussian/gaussian.cu:17-277) fused with the normal generation: each
// thread owns one Philox block (4 consecutive elements of the [K][T][C] buffer) and writes them as one float4.
template <bool ALIGNED>
__global__ void __launch_bounds__(256) sample_gaussian_kernel(mppi_sampler_params sp, int K, int T, int C,
                                                              const float* __restrict__ mean,
                                                              float* __restrict__ out, IterRef itr,
                                                              const float* __restrict__ eps_in, int pstart)
{
  const IterArgs a = load_iter(itr);
  const float decay = powf(sp.std_dev_decay, (float)a.iteration);
  float sd[MPPI_MAX_CONTROL];
#pragma unroll
  for (int c = 0; c < MPPI_MAX_CONTROL; ++c) sd[c] = decay * sp.std_dev[c];
  const uint32_t TC = (uint32_t)(T * C);
  const int64_t n = (int64_t)K * TC;
  const int64_t nb = (n + 3) / 4;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e0 = b * 4;
    float eps[4];
    if (eps_in != nullptr) {
#pragma unroll
      for (int j = 0; j < 4; ++j) eps[j] = (e0 + j < n) ? eps_in[e0 + j] : 0.0f;
    } else {
      const Normal4 z = normals4(sp.seed, a.rng_offset, (uint64_t)b);
#pragma unroll
      for (int j = 0; j < 4; ++j) eps[j] = z.z[j];
    }
    uint32_t k = (uint32_t)(e0 / TC);
    uint32_t r = (uint32_t)(e0 - (int64_t)k * TC);
    uint32_t t = r / (uint32_t)C;
    uint32_t c = r - t * (uint32_t)C;
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float mu = mean[t * C + c];
      const float s = sp.std_dev_t ? decay * sp.std_dev_t[t * C + c] : sd[c];
      if (k == 0 || (int)t < a.stride) v[j] = mu;
      else if ((int)k >= pstart) v[j] = s * eps[j];
      else v[j] = mu + s * eps[j];
      if (++c == (uint32_t)C) {
        c = 0;
        if (++t == (uint32_t)T) {
          t = 0;
          ++k;
        }
      }
    }
    if (ALIGNED && e0 + 3 < n) {
      *reinterpret_cast<float4*>(out + e0) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (e0 + j < n) out[e0 + j] = v[j];
    }
  }
}

int launch_sample_gaussian(const mppi_sampler_params& sp, int K, int T, int C, const float* mean, float* samples,
                           IterRef it, const float* eps_in, hipStream_t s, int rows)
{
  const int pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  const int Kw = (rows >= 0 && rows < K) ? rows : K;  // samples written
  if (Kw == 0) return MPPI_OK;
  const int64_t n = (int64_t)Kw * T * C;
  const int grid = grid_for((n + 3) / 4, 256, 8192);
  if (((uintptr_t)samples & 15) == 0)
    sample_gaussian_kernel<true><<<grid, 256, 0, s>>>(sp, Kw, T, C, mean, samples, it, eps_in, pstart);
  else
    sample_gaussian_kernel<false><<<grid, 256, 0, s>>>(sp, Kw, T, C, mean, samples, it, eps_in, pstart);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// ===================================================================================================== rollout
// rolloutKernel (core/mppi_common.cu:35-153): each workgroup stages its control tile [64][T*C] (a contiguous span of
// the sample buffer) in LDS with padded rows, runs the pipelined rollout of rollout_core.hpp, writes the costs and
// stores the clamped tile back with 16-byte stores.
template <int MODEL, int COST>
__global__ void __launch_bounds__(rollout_threads<MODEL>()) rollout_kernel(mppi_dynamics_params dyn, mppi_cost_params cost,
                                                             mppi_sampler_params sp, int K, int T, float dt,
                                                             float lambda, float alpha, const float* __restrict__ x0,
                                                             const float* __restrict__ mean,
                                                             float* __restrict__ samples, float* __restrict__ costs,
                                                             int pstart, int chunk)
{
  constexpr int C = Model<MODEL>::C;
  constexpr int NT = rollout_threads<MODEL>();
  extern __shared__ float lds[];
  const int TC = T * C;
  const int ROW = TC + 1;
  float* tile = lds;
  float* mu = tile + RC_LANES * ROW;
  float* scratch = mu + ((TC + 3) & ~3);
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * RC_LANES;
  const int nrows = min(RC_LANES, K - k0);
  // the mean is requested first (T C <= 1024 <= 2 NT): its loads then retire with the tile's
  const float mv0 = mean[min(tid, TC - 1)], mv1 = mean[min(tid + NT, TC - 1)];
  bool deadband = false;
#pragma unroll
  for (int c = 0; c < C; ++c) deadband |= dyn.control_deadband[c] != 0.0f;
  if (deadband) tile_load<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid, ControlConstrain<C>{&dyn});
  else tile_load<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid, ControlClamp<C>(dyn));
  if (tid < TC) mu[tid] = mv0;
  if (tid + NT < TC) mu[tid + NT] = mv1;
  __syncthreads();
  float* cost_l;
  float* raw_l;
  rollout_tile<MODEL, COST>(dyn, cost, sp, T, dt, lambda, alpha, x0, tile, ROW, mu, pstart, k0, scratch, chunk,
                            cost_l, raw_l);
  if (tid < nrows) costs[k0 + tid] = cost_l[tid];
  tile_store<NT>(samples + (size_t)k0 * TC, tile, nrows, TC, ROW, tid);
}

template <int MODEL, int COST>
static int launch_rollout_t(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                            const mppi_sampler_params& sp, int K, int T, float dt, float lambda, float alpha,
                            const float* x0, const float* mean, float* samples, float* costs, hipStream_t s)
{
  constexpr int C = Model<MODEL>::C, O = Model<MODEL>::O;
  const size_t base = (size_t)RC_LANES * (T * C + 1) + (size_t)((T * C + 3) & ~3);
  const size_t limit = LDS_LIMIT / sizeof(float);
  const int chunk = base < limit ? rollout_chunk_for<MODEL, O>(limit - base) : 0;
  if (chunk == 0) return fail(MPPI_E_SIZE, "rollout: horizon too long for the LDS control tile");
  const size_t lds = sizeof(float) * (base + rollout_scratch_for<MODEL, O>(chunk));
  MPPI_CHECK(allow_max_dynamic_lds((const void*)rollout_kernel<MODEL, COST>));
  const int pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  const int grid = (K + RC_LANES - 1) / RC_LANES;
  rollout_kernel<MODEL, COST><<<grid, rollout_threads<MODEL>(), lds, s>>>(dyn, cost, sp, K, T, dt, lambda, alpha, x0, mean,
                                                            samples, costs, pstart, chunk);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

int launch_rollout(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp,
                   int K, int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                   float* samples, float* costs, hipStream_t s)
{
  const int model = (dyn.model == MPPI_MODEL_RACER_DUBINS_NN && dyn.nn_precision == MPPI_NN_BF16)
                        ? MODEL_RACER_DUBINS_NN_BF16
                        : dyn.model;
#define MPPI_ROLLOUT_CASE(MODEL, COST)                                                                          \
  if (model == MODEL && cost.cost == COST)                                                                      \
    return launch_rollout_t<MODEL, COST>(dyn, cost, sp, K, T, dt, lambda, alpha, x0, mean, samples, costs, s);
  MPPI_ROLLOUT_CASE(MPPI_MODEL_CARTPOLE, MPPI_COST_CARTPOLE_QUADRATIC)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_CARTPOLE, MPPI_COST_QUADRATIC_MAP)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_QUADRATIC_MAP)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_CARTPOLE_QUADRATIC)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_RACER_DUBINS, MPPI_COST_QUADRATIC_MAP)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_RACER_DUBINS_NN, MPPI_COST_QUADRATIC_MAP)
  MPPI_ROLLOUT_CASE(MODEL_RACER_DUBINS_NN_BF16, MPPI_COST_QUADRATIC_MAP)
#undef MPPI_ROLLOUT_CASE
  return fail(MPPI_E_UNSUPPORTED, "rollout: model/cost combination is not compiled into the kernel library");
}

// =============================================================================================== weight transform
// fullGPUcomputeWeights (core/mppi_common.cu:1067-1089): one workgroup; baseline = first strict minimum, weights
// w_out[k] (may be costs itself), normalizer and sum of squares in FP64. The weights are the normExp transform
// exp(-(c - b) / lambda) (:994-1002) or, with TSALLIS, (1 - (c - b) / gamma)^(1 / (r - 1)) for c - b < gamma and 0
// beyond (TsallisTransform, :1004-1021, with its FP64 1.0 - (c - b) / gamma rounded to FP32 for logf). Optionally the
// statistics also go to host memory and the iteration's rng offset is advanced (device-resident loops).
template <bool TSALLIS>
__global__ void __launch_bounds__(1024) weights_kernel(const float* costs, float* w_out, int K, float lambda,
                                                       float gamma, float r, mppi_weight_stats* __restrict__ stats,
                                                       mppi_weight_stats* __restrict__ stats_copy,
                                                       IterArgs* __restrict__ advance, int advance_step)
{
  __shared__ float sv[16];
  __shared__ int si[16];
  __shared__ double se[16], se2[16], sec[16];
  __shared__ float sbase;
  __shared__ int sidx;
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  float v = INFINITY;
  int idx = INT_MAX;
  for (int i = tid; i < K; i += blockDim.x) argmin_combine(v, idx, sanitize_cost(costs[i]), i);
  wave_argmin(v, idx);
  if (lane == 0) {
    sv[wid] = v;
    si[wid] = idx;
  }
  __syncthreads();
  if (tid == 0) {
    float bv = sv[0];
    int bi = si[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) argmin_combine(bv, bi, sv[w], si[w]);
    if (bi == INT_MAX) bi = 0;
    sbase = bv;
    sidx = bi;
  }
  __syncthreads();
  const float b = sbase;
  const float inv = 1.0f / lambda;
  double e = 0.0, e2 = 0.0, ec = 0.0;
  for (int i = tid; i < K; i += blockDim.x) {
    const float c = costs[i];
    const bool fin = c <= FLT_MAX;
    float w;
    if (TSALLIS) {
      const float dif = c - b;
      w = dif < gamma ? expf(logf((float)(1.0 - (double)(dif / gamma))) / (r - 1.0f)) : 0.0f;
    } else {
      w = fin ? expf(-inv * (c - b)) : 0.0f;
    }
    w_out[i] = w;
    e += (double)w;
    e2 += (double)w * (double)w;
    if (fin) ec += (double)w * (double)c;
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
  if (tid == 0) {
    double te = 0.0, te2 = 0.0, tec = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      te += se[w];
      te2 += se2[w];
      tec += sec[w];
    }
    mppi_weight_stats st;
    st.baseline = b;
    st.argmin = sidx;
    st.normalizer = te;
    st.sum_w2 = te2;
    st.lambda = lambda;
    st.num_rollouts = K;
    st.weighted_cost = tec / te;
    *stats = st;
    if (stats_copy != nullptr) {
      *stats_copy = st;
      __threadfence_system();  // in host memory before any later kernel's completion signal
    }
    if (advance != nullptr) advance->rng_offset += (uint64_t)advance_step;
  }
}

int launch_compute_weights(float* costs, int K, float lambda, mppi_weight_stats* stats, hipStream_t s)
{
  weights_kernel<false><<<1, 1024, 0, s>>>(costs, costs, K, lambda, 0.0f, 0.0f, stats, nullptr, nullptr, 0);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

int launch_weights(const float* costs, float* w_out, int K, float lambda, float gamma, float r,
                   mppi_weight_stats* stats, mppi_weight_stats* stats_copy, IterArgs* advance, int advance_step,
                   hipStream_t s)
{
  if (gamma != 0.0f && r != 0.0f)
    weights_kernel<true><<<1, 1024, 0, s>>>(costs, w_out, K, lambda, gamma, r, stats, stats_copy, advance,
                                           advance_step);
  else
    weights_kernel<false><<<1, 1024, 0, s>>>(costs, w_out, K, lambda, 0.0f, 0.0f, stats, stats_copy, advance,
                                            advance_step);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// ================================================================================================ weighted update
// weightedReductionKernelDeviceNorm (core/mppi_common.cu:746-773): new_u[j] = sum_k w_k u_k[j] / normalizer, j = t*C+c.
// Two launches over an internal workspace: every workgroup reduces a block of 64 whole rows (contiguous,
// coalesced reads) into a partial row S_g[T*C]; a column pass adds the partials in fixed order and divides by the
// normalizer read from the statistics on the device.
constexpr int WU_ROWS = 64;
constexpr int WU_THREADS = 256;
constexpr int WU_COLS = 32;

__global__ void __launch_bounds__(WU_THREADS) update_partial_kernel(const float* __restrict__ w,
                                                                    const float* __restrict__ samples, int K, int TC,
                                                                    float* __restrict__ part)
{
  __shared__ float wl[WU_ROWS];
  const int tid = threadIdx.x, g = blockIdx.x, k0 = g * WU_ROWS;
  const int nr = min(WU_ROWS, K - k0);
  if (tid < WU_ROWS) wl[tid] = (tid < nr) ? w[k0 + tid] : 0.0f;
  __syncthreads();
  const float* base = samples + (size_t)k0 * TC;
  for (int j = tid; j < TC; j += WU_THREADS) {
    float acc = 0.0f;
#pragma unroll 8
    for (int r = 0; r < nr; ++r) acc += wl[r] * base[(size_t)r * TC + j];
    part[(size_t)g * TC + j] = acc;
  }
}

__global__ void __launch_bounds__(WU_THREADS) update_columns_kernel(const float* __restrict__ part, int G, int TC,
                                                                    const mppi_weight_stats* __restrict__ stats,
                                                                    float* __restrict__ out)
{
  constexpr int NG = WU_THREADS / WU_COLS;
  __shared__ float acc_l[NG][WU_COLS];
  const int tid = threadIdx.x, col = tid % WU_COLS, grp = tid / WU_COLS;
  const int j = blockIdx.x * WU_COLS + col;
  float acc = 0.0f;
  if (j < TC) {
#pragma unroll 8
    for (int q = grp; q < G; q += NG) acc += part[(size_t)q * TC + j];
  }
  acc_l[grp][col] = acc;
  __syncthreads();
  if (grp == 0 && j < TC) {
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < NG; ++q) s += acc_l[q][col];
    out[j] = (float)((double)s / stats->normalizer);
  }
}

size_t weighted_update_workspace_bytes(int K, int T, int C)
{
  return (size_t)((K + WU_ROWS - 1) / WU_ROWS) * T * C * sizeof(float);
}

int launch_weighted_update(const float* w, const mppi_weight_stats* stats, const float* samples, int K, int T, int C,
                           float* out, void* ws, hipStream_t s)
{
  const int TC = T * C;
  const int G = (K + WU_ROWS - 1) / WU_ROWS;
  float* part = static_cast<float*>(ws);
  update_partial_kernel<<<G, WU_THREADS, 0, s>>>(w, samples, K, TC, part);
  MPPI_HIP_CHECK(hipGetLastError());
  update_columns_kernel<<<(TC + WU_COLS - 1) / WU_COLS, WU_THREADS, 0, s>>>(part, G, TC, stats, out);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// One wavefront per span, so that the spans' transfers (host-memory latency each) overlap; every lane issues its
// loads of a 256-word block before its stores.
struct WordSpans {
  WordSpan s[MAX_WORD_SPANS];
};

__global__ void __launch_bounds__(64 * MAX_WORD_SPANS) copy_words_kernel(WordSpans l, int count, HostSignal sig)
{
  const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (k < count) {
    const WordSpan sp = k == 0 ? l.s[0] : (k == 1 ? l.s[1] : (k == 2 ? l.s[2] : l.s[3]));
    const uint32_t* __restrict__ src = static_cast<const uint32_t*>(sp.src);
    uint32_t* __restrict__ dst = static_cast<uint32_t*>(sp.dst);
    for (int i0 = 0; i0 < sp.words; i0 += 4 * 64) {
      uint32_t v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = i0 + 64 * j + lane;
        v[j] = i < sp.words ? src[i] : 0u;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = i0 + 64 * j + lane;
        if (i < sp.words) dst[i] = v[j];
      }
    }
  }
  signal_host_done(sig);
}

int launch_copy_words(const WordSpan* spans, int count, hipStream_t s, const HostSignal* sig)
{
  if (count < 1 || count > MAX_WORD_SPANS) return fail(MPPI_E_INVALID, "copy_words: 1 to 4 spans");
  WordSpans l{};
  for (int i = 0; i < count; ++i) {
    if (spans[i].words < 0) return fail(MPPI_E_INVALID, "copy_words: negative span");
    l.s[i] = spans[i];
  }
  copy_words_kernel<<<1, 64 * count, 0, s>>>(l, count, sig != nullptr ? *sig : HostSignal{});
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

}  // namespace mppi_amd
