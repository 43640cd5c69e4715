// HIP kernels of the MPPI iteration for gfx950: normal generation, Gaussian and colored-noise control sampling,
// the LDS-staged fused rollout, the weight transform and the weighted control update.
#include <climits>

#include "kernels.hpp"
#include "models.hpp"

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
      if (k == 0 || (int)t < a.stride) v[j] = mu;
      else if ((int)k >= pstart) v[j] = sd[c] * eps[j];
      else v[j] = mu + sd[c] * eps[j];
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
                           IterRef it, const float* eps_in, hipStream_t s)
{
  const int pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  const int64_t n = (int64_t)K * T * C;
  const int grid = grid_for((n + 3) / 4, 256, 8192);
  if (((uintptr_t)samples & 15) == 0)
    sample_gaussian_kernel<true><<<grid, 256, 0, s>>>(sp, K, T, C, mean, samples, it, eps_in, pstart);
  else
    sample_gaussian_kernel<false><<<grid, 256, 0, s>>>(sp, K, T, C, mean, samples, it, eps_in, pstart);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// ============================================================================================ colored-noise samples
// ColoredNoiseDistributionImpl::generateSamples (sampling_distributions/colored_noise/colored_noise.cu:286-392) in
// one kernel: each workgroup owns COLORED_SPB samples (their C series), builds the power-law frequency scales and
// sigma_c (:292-338), draws or loads the complex spectra [k][c][f] (:343), applies configureFrequencyNoise (:12-37),
// evaluates the inverse real DFT of length n at t = 0..T-1 and at the offset index, and writes
// (x[t] - x[stride] decay^t) / (sigma_c n) (:39-56) through the Gaussian transform.
constexpr int COLORED_SPB = 8;

__global__ void __launch_bounds__(256) sample_colored_kernel(mppi_sampler_params sp, int K, int T, int C, int n,
                                                             int F, int spb, const float* __restrict__ mean,
                                                             float* __restrict__ out, IterRef itr,
                                                             const float* __restrict__ spectra_in, int pstart)
{
  extern __shared__ float lds[];
  float* scale = lds;                        // [F][C]
  float* ct = scale + F * C;                 // [n]
  float* st = ct + n;                        // [n]
  float* sigma = st + n;                     // [C] (padded to 4)
  float* spec = sigma + 4;                   // [spb*C][F][2]
  float* xs = spec + (size_t)spb * C * F * 2;  // [spb*C][T+1]

  const IterArgs a = load_iter(itr);
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * spb;
  const int ns = min(spb, K - k0);
  const int nser = ns * C;
  const int stride = min(max(a.stride, 0), n - 1);

  // frequency scales: f_i = i / n, frequencies below the cutoff take the first admissible one
  const float cutoff = fmaxf(sp.fmin, 1.0f / n);
  int j0 = 0;
  while (j0 < F && (j0 / (1.0f * n)) < cutoff) ++j0;
  for (int i = tid; i < F * C; i += blockDim.x) {
    const int fi = i / C, c = i - fi * C;
    const int src = (fi < j0 && j0 < F) ? j0 : fi;
    scale[i] = powf(src / (1.0f * n), -sp.exponents[c] / 2.0f);
  }
  for (int m = tid; m < n; m += blockDim.x) {
    float s, c;
    sincospif((float)(2 * m) / (float)n, &s, &c);
    ct[m] = c;
    st[m] = s;
  }
  __syncthreads();
  if (tid < C) {
    float sg = 0.0f;
    for (int j = 1; j < F - 1; ++j) sg += scale[j * C + tid] * scale[j * C + tid];
    const float last = scale[(F - 1) * C + tid] * ((1.0f + (n % 2)) / 2.0f);
    sg += last * last;
    sigma[tid] = 2.0f * sqrtf(sg) / n;
  }
  // spectra of this workgroup's series: global elements [base, base + cnt) of the [K][C][F][2] buffer
  const int64_t base = (int64_t)k0 * C * F * 2;
  const int64_t cnt = (int64_t)nser * F * 2;
  for (int64_t gb = base / 4 + tid; gb * 4 < base + cnt; gb += blockDim.x) {
    float z[4];
    if (spectra_in != nullptr) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t g = gb * 4 + j;
        z[j] = (g >= base && g < base + cnt) ? spectra_in[g] : 0.0f;
      }
    } else {
      const Normal4 nz = normals4(sp.seed, a.rng_offset, (uint64_t)gb);
#pragma unroll
      for (int j = 0; j < 4; ++j) z[j] = nz.z[j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t g = gb * 4 + j;
      if (g < base || g >= base + cnt) continue;
      const int e = (int)(g - base);
      const int f = (e >> 1) % F;
      const bool zero_im = (e & 1) && (f == 0 || ((F & 1) && f == F - 1));
      spec[e] = zero_im ? 0.0f : z[j];
    }
  }
  __syncthreads();
  for (int e = tid; e < nser * F * 2; e += blockDim.x) {
    const int f = (e >> 1) % F;
    const int c = ((e >> 1) / F) % C;
    spec[e] *= scale[f * C + c];
  }
  __syncthreads();
  // inverse real DFT at t = 0..T-1 and at the offset index
  for (int idx = tid; idx < nser * (T + 1); idx += blockDim.x) {
    const int sidx = idx / (T + 1);
    const int o = idx - sidx * (T + 1);
    const int tt = (o < T) ? o : stride;
    const float* X = spec + (size_t)sidx * F * 2;
    float acc = X[0];
    int m = 0;
    for (int f = 1; f < F; ++f) {
      m += tt;
      if (m >= n) m -= n;
      const float re = X[2 * f], im = X[2 * f + 1];
      if (2 * f == n) acc += re * ct[m];
      else acc += 2.0f * (re * ct[m] - im * st[m]);
    }
    xs[idx] = acc;
  }
  __syncthreads();
  const float decay = powf(sp.std_dev_decay, (float)a.iteration);
  const int TC = T * C;
  for (int e = tid; e < ns * TC; e += blockDim.x) {
    const int r = e / TC;
    const int rem = e - r * TC;
    const int t = rem / C;
    const int c = rem - t * C;
    const int k = k0 + r;
    const float* xser = xs + (size_t)(r * C + c) * (T + 1);
    const float dec = (sp.offset_decay_rate == 0.0f) ? 0.0f : powf(sp.offset_decay_rate, (float)t);
    const float eps = (xser[t] - xser[T] * dec) / (sigma[c] * n);
    const float sd = decay * sp.std_dev[c];
    const float mu = mean[t * C + c];
    float v;
    if (k == 0 || t < a.stride) v = mu;
    else if (k >= pstart) v = sd * eps;
    else v = mu + sd * eps;
    out[(size_t)k * TC + rem] = v;
  }
}

static size_t colored_lds_bytes(int T, int C, int n, int F, int spb)
{
  return sizeof(float) * ((size_t)F * C + 2 * n + 4 + (size_t)spb * C * F * 2 + (size_t)spb * C * (T + 1));
}

int launch_sample_colored(const mppi_sampler_params& sp, int K, int T, int C, const float* mean, float* samples,
                          IterRef it, const float* spectra_in, hipStream_t s)
{
  const int pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  const int n = (sp.fft_len_mode == MPPI_FFT_LEN_T) ? T : 2 * T;
  const int F = n / 2 + 1;
  const int spb = (C > 2) ? COLORED_SPB / 2 : COLORED_SPB;
  const size_t lds = colored_lds_bytes(T, C, n, F, spb);
  if (lds > 160 * 1024) return fail(MPPI_E_SIZE, "colored sampler: horizon too long for the LDS tile");
  static bool attr_set = false;
  if (!attr_set) {
    MPPI_HIP_CHECK(hipFuncSetAttribute((const void*)sample_colored_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  const int grid = (K + spb - 1) / spb;
  sample_colored_kernel<<<grid, 256, lds, s>>>(sp, K, T, C, n, F, spb, mean, samples, it, spectra_in, pstart);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// ===================================================================================================== rollout
// rolloutKernel (core/mppi_common.cu:35-153) with one sample per lane. The workgroup's control tile [64][T*C]
// (a contiguous span of the sample buffer) is staged in LDS with a padded row, each lane steps its sample over the
// whole horizon without barriers, and the clamped tile is written back with coalesced stores.
constexpr int ROLLOUT_SPB = 64;

template <int MODEL, int COST>
__global__ void __launch_bounds__(ROLLOUT_SPB) rollout_kernel(mppi_dynamics_params dyn, mppi_cost_params cost,
                                                              mppi_sampler_params sp, int K, int T, float dt,
                                                              float lambda, float alpha,
                                                              const float* __restrict__ x0,
                                                              const float* __restrict__ mean,
                                                              float* __restrict__ samples, float* __restrict__ costs,
                                                              int pstart)
{
  using M = Model<MODEL>;
  constexpr int S = M::S, C = M::C, O = M::O;
  extern __shared__ float lds[];
  const int TC = T * C;
  const int ROW = TC + 1;
  float* tile = lds;
  float* mu = lds + ROLLOUT_SPB * ROW;
  const int lane = threadIdx.x;
  const int k0 = blockIdx.x * ROLLOUT_SPB;
  const int nrows = min(ROLLOUT_SPB, K - k0);
  const float* src = samples + (size_t)k0 * TC;
  for (int r = 0; r < nrows; ++r)
    for (int j = lane; j < TC; j += ROLLOUT_SPB) tile[r * ROW + j] = src[(size_t)r * TC + j];
  for (int j = lane; j < TC; j += ROLLOUT_SPB) mu[j] = mean[j];
  __syncthreads();

  const int k = k0 + lane;
  const bool pure = k >= pstart;
  float x[S], y[O];
#pragma unroll
  for (int i = 0; i < S; ++i) x[i] = x0[i];
#pragma unroll
  for (int i = 0; i < O; ++i) y[i] = x[i];
  float* row = tile + lane * ROW;
  float run = 0.0f;
  for (int t = 0; t < T; ++t) {
    float u[C];
#pragma unroll
    for (int c = 0; c < C; ++c) u[c] = row[t * C + c];
    enforce_constraints<C>(dyn, u);
#pragma unroll
    for (int c = 0; c < C; ++c) row[t * C + c] = u[c];
    float xd[S], xn[S];
    M::deriv(dyn, x, u, xd);
    M::update(dyn, x, xn, xd, dt);
#pragma unroll
    for (int i = 0; i < O; ++i) y[i] = xn[i];
    run += Cost<COST, O>::running(cost, y) + lr_cost<C>(sp, u, mu + t * C, pure, lambda, alpha);
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = xn[i];
  }
  if (lane < nrows) costs[k] = run / (float)T + Cost<COST, O>::terminal(cost, y) / (float)T;
  __syncthreads();
  float* dst = samples + (size_t)k0 * TC;
  for (int r = 0; r < nrows; ++r)
    for (int j = lane; j < TC; j += ROLLOUT_SPB) dst[(size_t)r * TC + j] = tile[r * ROW + j];
}

template <int MODEL, int COST>
static int launch_rollout_t(const mppi_dynamics_params& dyn, const mppi_cost_params& cost,
                            const mppi_sampler_params& sp, int K, int T, float dt, float lambda, float alpha,
                            const float* x0, const float* mean, float* samples, float* costs, hipStream_t s)
{
  constexpr int C = Model<MODEL>::C;
  const size_t lds = sizeof(float) * ((size_t)ROLLOUT_SPB * (T * C + 1) + (size_t)T * C);
  if (lds > 160 * 1024) return fail(MPPI_E_SIZE, "rollout: horizon too long for the LDS control tile");
  static bool attr_set = false;
  if (!attr_set) {
    MPPI_HIP_CHECK(hipFuncSetAttribute((const void*)rollout_kernel<MODEL, COST>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  const int pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  const int grid = (K + ROLLOUT_SPB - 1) / ROLLOUT_SPB;
  rollout_kernel<MODEL, COST><<<grid, ROLLOUT_SPB, lds, s>>>(dyn, cost, sp, K, T, dt, lambda, alpha, x0, mean,
                                                            samples, costs, pstart);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

int launch_rollout(const mppi_dynamics_params& dyn, const mppi_cost_params& cost, const mppi_sampler_params& sp,
                   int K, int T, float dt, float lambda, float alpha, const float* x0, const float* mean,
                   float* samples, float* costs, hipStream_t s)
{
#define MPPI_ROLLOUT_CASE(MODEL, COST)                                                                          \
  if (dyn.model == MODEL && cost.cost == COST)                                                                  \
    return launch_rollout_t<MODEL, COST>(dyn, cost, sp, K, T, dt, lambda, alpha, x0, mean, samples, costs, s);
  MPPI_ROLLOUT_CASE(MPPI_MODEL_CARTPOLE, MPPI_COST_CARTPOLE_QUADRATIC)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_CARTPOLE, MPPI_COST_QUADRATIC_MAP)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_QUADRATIC_MAP)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_DOUBLE_INTEGRATOR, MPPI_COST_CARTPOLE_QUADRATIC)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_RACER_DUBINS, MPPI_COST_QUADRATIC_MAP)
  MPPI_ROLLOUT_CASE(MPPI_MODEL_RACER_DUBINS_NN, MPPI_COST_QUADRATIC_MAP)
#undef MPPI_ROLLOUT_CASE
  return fail(MPPI_E_UNSUPPORTED, "rollout: model/cost combination is not compiled into the kernel library");
}

// =============================================================================================== weight transform
// fullGPUcomputeWeights (core/mppi_common.cu:1067-1089): one workgroup; baseline = first strict minimum, weights in
// place, normalizer and sum of squares in FP64.
__global__ void __launch_bounds__(1024) weights_kernel(float* __restrict__ costs, int K, float lambda,
                                                       mppi_weight_stats* __restrict__ stats)
{
  __shared__ float sv[16];
  __shared__ int si[16];
  __shared__ double se[16], se2[16];
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
  double e = 0.0, e2 = 0.0;
  for (int i = tid; i < K; i += blockDim.x) {
    const float c = costs[i];
    const float w = (c <= FLT_MAX) ? expf(-inv * (c - b)) : 0.0f;
    costs[i] = w;
    e += (double)w;
    e2 += (double)w * (double)w;
  }
  e = wave_sum(e);
  e2 = wave_sum(e2);
  if (lane == 0) {
    se[wid] = e;
    se2[wid] = e2;
  }
  __syncthreads();
  if (tid == 0) {
    double te = 0.0, te2 = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      te += se[w];
      te2 += se2[w];
    }
    mppi_weight_stats st;
    st.baseline = b;
    st.argmin = sidx;
    st.normalizer = te;
    st.sum_w2 = te2;
    st.lambda = lambda;
    st.num_rollouts = K;
    *stats = st;
  }
}

int launch_compute_weights(float* costs, int K, float lambda, mppi_weight_stats* stats, hipStream_t s)
{
  weights_kernel<<<1, 1024, 0, s>>>(costs, K, lambda, stats);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// ================================================================================================ weighted update
// weightedReductionKernelDeviceNorm (core/mppi_common.cu:746-773): new_u[j] = sum_k w_k u_k[j] / normalizer, j = t*C+c.
// A workgroup owns 64 consecutive columns; its 16 wavefronts take every 16th sample and are summed in fixed order.
__global__ void __launch_bounds__(1024) weighted_update_kernel(const float* __restrict__ w,
                                                               const mppi_weight_stats* __restrict__ stats,
                                                               const float* __restrict__ samples, int K, int TC,
                                                               float* __restrict__ out)
{
  __shared__ float part[16][64];
  const int tid = threadIdx.x, col = tid & 63, rg = tid >> 6;
  const int j = blockIdx.x * 64 + col;
  float acc = 0.0f;
  if (j < TC)
    for (int k = rg; k < K; k += 16) acc += w[k] * samples[(size_t)k * TC + j];
  part[rg][col] = acc;
  __syncthreads();
  if (rg == 0 && j < TC) {
    float s = 0.0f;
#pragma unroll
    for (int g = 0; g < 16; ++g) s += part[g][col];
    out[j] = (float)((double)s / stats->normalizer);
  }
}

int launch_weighted_update(const float* w, const mppi_weight_stats* stats, const float* samples, int K, int T, int C,
                           float* out, hipStream_t s)
{
  const int TC = T * C;
  weighted_update_kernel<<<(TC + 63) / 64, 1024, 0, s>>>(w, stats, samples, K, TC, out);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

}  // namespace mppi_amd
