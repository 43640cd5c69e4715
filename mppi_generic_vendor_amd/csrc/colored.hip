// Colored-noise (power-law) control sampler for gfx950
// (ColoredNoiseDistributionImpl::generateSamples, sampling_distributions/colored_noise/colored_noise.cu:286-392).
//
// For series s = (k, c) with raw complex spectrum X_s[f] (f < F = n/2 + 1, drawn as normals or fed) the reference
// computes, with scales s_c(f) = f^(-beta_c/2) (low frequencies clamped to the cutoff, :300-326), the unnormalised
// inverse real FFT x = C2R(s_c X_s) of length n = 2T (or T), and
//     eps[k][t][c] = (x[t] - x[stride] * decay^t) / (sigma_c * n),      t < T                      (:39-56)
// which it then feeds through the Gaussian transform. Here the whole map from the scaled spectrum to eps is one
// linear operator per workgroup:
//     OUT[t][s] = sum_j B[t][j] * Y[j][s],   Y[2f][s] = s_c(f) Re X_s[f],  Y[2f+1][s] = s_c(f) Im X_s[f]
//     B[t][2f]   =  w_f (cos(2 pi f t/n) - decay^t cos(2 pi f stride/n))
//     B[t][2f+1] = -w_f (sin(2 pi f t/n) - decay^t sin(2 pi f stride/n)),   w_f = 1 at f = 0 and f = n/2, else 2
// (the imaginary parts the C2R transform ignores -- DC and Nyquist -- and those configureFrequencyNoise zeroes
// (:26-33) have B = 0), evaluated as an FP32 MFMA product (v_mfma_f32_16x16x4f32): every workgroup of 16 wavefronts
// owns up to 64 series, stages their scaled spectra in LDS, builds the basis in K-chunks of 32 columns; wavefront w
// accumulates the 16 x 16 output tiles (M-tile mg + i*MG, N-tile nt) in registers. The epilogue divides by
// sigma_c * n, applies the Gaussian transform and writes the workgroup's contiguous span of the [K][T][C] sample
// buffer with 16-byte stores.
#include <vector>

#include "kernels.hpp"
#include "models.hpp"

namespace mppi_amd {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ IterArgs load_iter_c(const IterRef& r) { return r.ptr ? *r.ptr : r.val; }

constexpr int CKC = 32;        // basis K-chunk (columns)
constexpr int CTHREADS = 1024; // 16 wavefronts
constexpr int CKCP = CKC + 1;  // padded LDS row of the basis chunk

struct Sigma4 {
  float v[MPPI_MAX_CONTROL];
};

// sigma_c exactly as the reference's host code computes it (colored_noise.cu:292-338, FP32, sequential sums).
static Sigma4 colored_sigma(const mppi_sampler_params& sp, int C, int n)
{
  const int F = n / 2 + 1;
  std::vector<float> f(F), sc((size_t)F * C);
  for (int i = 0; i < F; ++i) f[i] = i / (1.0f * n);
  const float cutoff = fmaxf(sp.fmin, 1.0f / n);
  int smaller = 0;
  for (int i = 0; i < F; ++i) {
    if (f[i] < cutoff) {
      smaller++;
    } else if (smaller < F) {
      for (int j = 0; j < smaller; ++j) {
        f[j] = f[smaller];
        for (int c = 0; c < C; ++c) sc[j * C + c] = powf(f[smaller], -sp.exponents[c] / 2.0f);
      }
    }
    for (int c = 0; c < C; ++c) sc[i * C + c] = powf(f[i], -sp.exponents[c] / 2.0f);
  }
  Sigma4 s{};
  for (int c = 0; c < C; ++c) {
    float sg = 0.0f;
    for (int j = 1; j < F - 1; ++j) sg += sc[j * C + c] * sc[j * C + c];
    const float last = sc[(F - 1) * C + c] * ((1.0f + (n % 2)) / 2.0f);
    sg += last * last;
    s.v[c] = 2.0f * sqrtf(sg) / n;
  }
  return s;
}

struct ColoredShape {
  int n, F, Kpad, KP, MT, MTP, NS, spw, threads;
  size_t lds_floats;
};

static ColoredShape colored_shape(int T, int C, int mode)
{
  ColoredShape sh;
  sh.n = (mode == MPPI_FFT_LEN_T) ? T : 2 * T;
  sh.F = sh.n / 2 + 1;
  sh.Kpad = (2 * sh.F + CKC - 1) / CKC * CKC;
  sh.KP = sh.Kpad + 1;
  sh.MT = (T + 15) / 16;
  for (int ns = 64; ns >= 16; ns /= 2) {
    sh.NS = ns;
    sh.spw = ns / C;
    sh.threads = CTHREADS;
    // every wavefront owns MPW M-tiles of one N-tile; the basis rows are padded to MPW * MG * 16 (zero rows)
    const int mg = (CTHREADS / 64) / (ns / 16);
    sh.MTP = (sh.MT + mg - 1) / mg * mg;
    const size_t tables = 2 * (size_t)sh.n + T + (((size_t)sh.F * C + 3) & ~(size_t)3);
    sh.lds_floats = tables + (size_t)ns * sh.KP + (size_t)sh.MTP * 16 * CKCP;
    if (sh.lds_floats * sizeof(float) <= LDS_LIMIT && sh.spw >= 1) break;
  }
  return sh;
}

template <int MPW>
__global__ void __launch_bounds__(CTHREADS) colored_mfma_kernel(mppi_sampler_params sp, int K, int T, int C, int n,
                                                                int F, int MTP, int Kpad, int KP, int NS, int spw,
                                                           const float* __restrict__ mean, float* __restrict__ out,
                                                           IterRef itr, const float* __restrict__ spectra_in,
                                                           int pstart, Sigma4 sig)
{
  extern __shared__ float lds[];
  float* ct = lds;               // [n]
  float* st = ct + n;            // [n]
  float* dec = st + n;           // [T]
  float* scale = dec + T;        // [F][C]
  float* X = scale + ((F * C + 3) & ~3);  // [NS][KP]   scaled spectra; reused as the output staging [spw][T][C]
  float* Bc = X + NS * KP;       // [MTP*16][CKCP]

  const IterArgs a = load_iter_c(itr);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int k0 = blockIdx.x * spw;
  const int nsamp = min(spw, K - k0);
  const int nser = nsamp * C;
  const int TF = 2 * F;
  const int stride = min(max(a.stride, 0), n - 1);

  // tables: cos/sin of 2 pi m / n, decay^t, power-law scales (low frequencies take the first admissible one)
  for (int m = tid; m < n; m += nthr) {
    float s, c;
    sincospif((float)(2 * m) / (float)n, &s, &c);
    ct[m] = c;
    st[m] = s;
  }
  for (int t = tid; t < T; t += nthr) dec[t] = (sp.offset_decay_rate == 0.0f) ? 0.0f : powf(sp.offset_decay_rate, (float)t);
  const float cutoff = fmaxf(sp.fmin, 1.0f / n);
  int j0 = 0;
  while (j0 < F && (j0 / (1.0f * n)) < cutoff) ++j0;
  for (int i = tid; i < F * C; i += nthr) {
    const int fi = i / C, c = i - fi * C;
    const int src = (fi < j0 && j0 < F) ? j0 : fi;
    scale[i] = powf(src / (1.0f * n), -sp.exponents[c] / 2.0f);
  }
  __syncthreads();
  // scaled spectra of series s0 .. s0 + nser - 1 (a contiguous span of the [K][C][F][2] normal stream)
  const int64_t base = (int64_t)k0 * C * TF;
  const int64_t cnt = (int64_t)nser * TF;
  for (int64_t gb = (base >> 2) + tid; gb * 4 < base + cnt; gb += nthr) {
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
    const int e0 = (int)(gb * 4 - base);  // may be negative for the first block
    int ns_ = (e0 >= 0) ? e0 / TF : -1;
    int jj = e0 - ns_ * TF;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = e0 + j;
      if (e >= 0 && e < cnt) X[ns_ * KP + jj] = z[j] * scale[(jj >> 1) * C + (ns_ % C)];
      if (++jj == TF) {
        jj = 0;
        ++ns_;
      }
    }
  }
  // zero padding: columns [2F, KP) of the valid series, whole rows of the unused series
  const int padc = KP - TF;
  for (int i = tid; i < nser * padc; i += nthr) {
    const int r = i / padc;
    X[r * KP + TF + (i - r * padc)] = 0.0f;
  }
  for (int i = tid; i < (NS - nser) * KP; i += nthr) X[nser * KP + i] = 0.0f;

  const int wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int NTT = NS / 16;               // N-tiles (16 series each)
  const int MG = (CTHREADS / 64) / NTT;  // wavefronts sharing one N-tile
  const int nt = wave % NTT, mg = wave / NTT;
  floatx4 acc[MPW];
#pragma unroll
  for (int i = 0; i < MPW; ++i) acc[i] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int kc0 = 0; kc0 < Kpad; kc0 += CKC) {
    __syncthreads();  // spectra staged / previous chunk consumed
    // basis chunk: thread -> (row t, 8 consecutive columns); the phase index advances by t per frequency
    for (int item = tid; item < MTP * 16 * (CKC / 8); item += nthr) {
      const int t = item / (CKC / 8);
      const int cb = (item - t * (CKC / 8)) * 8;
      float* brow = Bc + t * CKCP + cb;
      if (t >= T) {
#pragma unroll
        for (int q = 0; q < 8; ++q) brow[q] = 0.0f;
        continue;
      }
      const float dt_ = dec[t];
      const int jbeg = kc0 + cb;  // even
      int f = jbeg >> 1;
      int mt_ = (int)(((long long)f * t) % n);
      int ms_ = (int)(((long long)f * stride) % n);
#pragma unroll
      for (int q = 0; q < 8; q += 2) {
        float re = 0.0f, im = 0.0f;
        if (f < F) {
          const bool edge = (f == 0) || (2 * f == n);
          const float w = edge ? 1.0f : 2.0f;
          re = w * (ct[mt_] - dt_ * ct[ms_]);
          const bool zero_im = edge || ((F & 1) && f == F - 1);
          im = zero_im ? 0.0f : -w * (st[mt_] - dt_ * st[ms_]);
        }
        brow[q] = re;
        brow[q + 1] = im;
        ++f;
        mt_ += t;
        if (mt_ >= n) mt_ -= n;
        ms_ += stride;
        if (ms_ >= n) ms_ -= n;
      }
    }
    __syncthreads();
    const float* xrow = X + (nt * 16 + lr) * KP + kc0 + lk;
#pragma unroll
    for (int ks = 0; ks < CKC / 4; ++ks) {
      const float bv = xrow[ks * 4];
#pragma unroll
      for (int i = 0; i < MPW; ++i) {
        const float av = Bc[((mg + i * MG) * 16 + lr) * CKCP + ks * 4 + lk];
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[i], 0, 0, 0);
      }
    }
  }
  __syncthreads();  // X is reused as the output staging below

  // epilogue: lane holds OUT[t = 16 mt + 4 lk + v][series 16 nt + lr] for mt = mg + i * MG
  const float decay_it = powf(sp.std_dev_decay, (float)a.iteration);
  const int TC = T * C;
  const int sidx = nt * 16 + lr;
  if (sidx < nser) {
    const int c = sidx % C, r = sidx / C, k = k0 + r;
    const float inv = 1.0f / (sig.v[c] * n);
    const float sd = decay_it * sp.std_dev[c];
#pragma unroll
    for (int i = 0; i < MPW; ++i) {
      const int mt = mg + i * MG;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int t = mt * 16 + lk * 4 + v;
        if (t < T) {
          const float eps = acc[i][v] * inv;
          const float mu = mean[t * C + c];
          float u;
          if (k == 0 || t < a.stride) u = mu;
          else if (k >= pstart) u = sd * eps;
          else u = mu + sd * eps;
          X[r * TC + t * C + c] = u;
        }
      }
    }
  }
  __syncthreads();
  float* dst = out + (size_t)k0 * TC;
  const int total = nsamp * TC;
  if ((((uintptr_t)dst) & 15) == 0) {
    const int n4 = total >> 2;
    for (int i = tid; i < n4; i += nthr)
      reinterpret_cast<float4*>(dst)[i] = make_float4(X[4 * i], X[4 * i + 1], X[4 * i + 2], X[4 * i + 3]);
    for (int e = (n4 << 2) + tid; e < total; e += nthr) dst[e] = X[e];
  } else {
    for (int e = tid; e < total; e += nthr) dst[e] = X[e];
  }
}

template <int MPW>
static int launch_colored_t(const mppi_sampler_params& sp, int K, int T, int C, const float* mean, float* samples,
                            IterRef it, const float* spectra_in, int pstart, const ColoredShape& sh, Sigma4 sig,
                            hipStream_t s)
{
  static bool attr_set = false;
  if (!attr_set) {
    MPPI_HIP_CHECK(hipFuncSetAttribute((const void*)colored_mfma_kernel<MPW>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, LDS_LIMIT));
    attr_set = true;
  }
  const int grid = (K + sh.spw - 1) / sh.spw;
  colored_mfma_kernel<MPW><<<grid, sh.threads, sh.lds_floats * sizeof(float), s>>>(
      sp, K, T, C, sh.n, sh.F, sh.MTP, sh.Kpad, sh.KP, sh.NS, sh.spw, mean, samples, it, spectra_in, pstart, sig);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

int launch_sample_colored(const mppi_sampler_params& sp, int K, int T, int C, const float* mean, float* samples,
                          IterRef it, const float* spectra_in, hipStream_t s)
{
  const int pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  const ColoredShape sh = colored_shape(T, C, sp.fft_len_mode);
  const Sigma4 sig = colored_sigma(sp, C, sh.n);
  if (sh.lds_floats * sizeof(float) > LDS_LIMIT || sh.spw < 1)
    return fail(MPPI_E_SIZE, "colored sampler: horizon too long for the LDS staging");
  const int MG = (CTHREADS / 64) / (sh.NS / 16);
  switch (sh.MTP / MG) {
#define MPPI_CMT(M) \
  case M: return launch_colored_t<M>(sp, K, T, C, mean, samples, it, spectra_in, pstart, sh, sig, s);
    MPPI_CMT(1) MPPI_CMT(2) MPPI_CMT(3) MPPI_CMT(4)
#undef MPPI_CMT
    default: return fail(MPPI_E_SIZE, "colored sampler: horizon exceeds 256");
  }
}

}  // namespace mppi_amd
