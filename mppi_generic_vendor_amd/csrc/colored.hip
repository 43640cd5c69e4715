// Colored-noise (power-law) control sampler for gfx950
// (ColoredNoiseDistributionImpl::generateSamples, sampling_distributions/colored_noise/colored_noise.cu:286-392).
//
// For series s = (k, c) with raw complex spectrum X_s[f] (f < F = n/2 + 1, drawn as normals or fed) the reference
// computes, with scales s_c(f) = f^(-beta_c/2) (low frequencies clamped to the cutoff, :300-326), the unnormalised
// inverse real FFT x = C2R(s_c X_s) of length n = 2T (or T), and
//     eps[k][t][c] = (x[t] - x[stride] * decay^t) / (sigma_c * n),      t < T                      (:39-56)
// which it then feeds through the Gaussian transform. Here the inverse transform of a workgroup's series is one
// matrix product
//     X[t][s] = sum_j B[t][j] * Y[j][s],   Y[2f][s] = s_c(f) Re X_s[f],  Y[2f+1][s] = s_c(f) Im X_s[f]
//     B[t][2f] = w_f cos(2 pi f t/n),  B[t][2f+1] = -w_f sin(2 pi f t/n),   w_f = 1 at f = 0 and f = n/2, else 2
// (the imaginary parts the C2R transform ignores -- DC and Nyquist -- and those configureFrequencyNoise zeroes
// (:26-33) have B = 0). B depends on n and T only: it is built once per shape (colored_basis_kernel) as FP16 high and
// low parts. Every workgroup of 16 wavefronts owns up to 64 series, stages their scaled spectra in LDS as FP16 high and
// low parts and forms the product on v_mfma_f32_16x16x32_f16 as Bh Yh + Bh Yl + Bl Yh; the epilogue subtracts
// decay^t x[stride], divides by sigma_c * n, applies the Gaussian transform and writes the workgroup's contiguous
// span of the [K][T][C] sample buffer with 16-byte stores. The device code up to the Gaussian transform
// (colored_stage, colored_core.hpp) is shared with the controller's fused iteration, which writes the samples into
// its control tile in LDS instead.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "colored_core.hpp"

namespace mppi_amd {

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

// Largest number of a wavefront's 64 lanes on one LDS bank (32 or 64 banks) when the lanes write the staging buffer
// with row stride TCS: lane (lr, lk) writes series 16 q + lr (sample r, component c) at t = 4 lk (+ a constant).
static int staging_conflict(int TCS, int C, int NTT)
{
  int worst = 0;
  for (int banks = 32; banks <= 64; banks *= 2)
    for (int q = 0; q < NTT; ++q) {
      int cnt[64] = {0};
      for (int lane = 0; lane < 64; ++lane) {
        const int sidx = 16 * q + (lane & 15), r = sidx / C, c = sidx % C;
        const int addr = r * TCS + 4 * (lane >> 4) * C + c;
        worst = std::max(worst, ++cnt[addr % banks]);
      }
    }
  return worst;
}

static ColoredShape colored_shape(int T, int C, int mode)
{
  ColoredShape sh;
  sh.n = (mode == MPPI_FFT_LEN_T) ? T : 2 * T;
  sh.F = sh.n / 2 + 1;
  sh.Kpad = (2 * sh.F + CKC - 1) / CKC * CKC;
  sh.KPH = sh.Kpad + 8;  // LDS row of a series' spectrum in halves: 16-byte aligned, consecutive rows shift banks
  sh.MT = (T + 15) / 16;
  sh.S = sh.MT > 0 ? std::max(1, std::min(16 / sh.MT, sh.Kpad / CKC)) : 1;
  for (int ns = 64; ns >= 16; ns /= 2) {
    sh.NS = ns;
    sh.spw = ns / C;
    // scale table, factors, mean, decay powers, x[stride] per series, the stride's basis row
    const size_t tables = ((((size_t)sh.F * C + 3) & ~(size_t)3) + 8 + (size_t)T * C + T + 64 + sh.Kpad + 3) &
                          ~(size_t)3;
    // high and low spectra; the output staging [spw][T][C] reuses the same space (KPH > T)
    const size_t xh = 2 * (size_t)ns * sh.KPH * sizeof(_Float16);
    const size_t part = (size_t)(sh.S - 1) * sh.MT * (ns / 16) * 64 * 16;  // k-group partial sums
    sh.lds_bytes = tables * sizeof(float) + xh + part;
    if (sh.lds_bytes <= LDS_LIMIT && sh.spw >= 1) break;
  }
  // staging row stride: T C padded (within the spectra space it reuses) to spread the stores over the banks
  const int TC = T * C;
  sh.TCS = TC;
  int best = staging_conflict(TC, C, sh.NS / 16);
  for (int pad = 1; pad < 32 && (size_t)sh.spw * (TC + pad) <= (size_t)sh.NS * sh.KPH; ++pad) {
    const int cf = staging_conflict(TC + pad, C, sh.NS / 16);
    if (cf < best) {
      best = cf;
      sh.TCS = TC + pad;
    }
  }
  return sh;
}

// The transform's basis for rows t < T (rows T .. rows - 1 zero) times B_SCALE, FP16 high and low parts, columns
// j < Kpad (2f: real part, 2f + 1: imaginary part). It depends on n and T only and is built once per shape.
__global__ void __launch_bounds__(256) colored_basis_kernel(int n, int F, int T, int rows, int Kpad,
                                                            _Float16* __restrict__ bhi, _Float16* __restrict__ blo)
{
  const int Fp = Kpad >> 1;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= rows * Fp) return;
  const int t = i / Fp, f = i - t * Fp;
  float re = 0.0f, im = 0.0f;
  if (t < T) basis_pair(n, F, t, f, re, im);
  _Float16 h0, l0, h1, l1;
  split_f16(re * B_SCALE, h0, l0);
  split_f16(im * B_SCALE, h1, l1);
  const half2v h = {h0, h1}, l = {l0, l1};
  *reinterpret_cast<half2v*>(bhi + (size_t)t * Kpad + 2 * f) = h;
  *reinterpret_cast<half2v*>(blo + (size_t)t * Kpad + 2 * f) = l;
}

template <int NTT>
__global__ void __launch_bounds__(CTHREADS) colored_mfma_kernel(mppi_sampler_params sp, int K, int T, int C, int n,
                                                                int F, int Kpad, int KPH, int MT, int S, int spw,
                                                                int TCS, const float* __restrict__ mean,
                                                                float* __restrict__ out, IterRef itr,
                                                                const float* __restrict__ spectra_in, int pstart,
                                                                Sigma4 sig, const _Float16* __restrict__ bhi,
                                                                const _Float16* __restrict__ blo)
{
  extern __shared__ float lds[];
  const ColoredTile ct = colored_stage<NTT>(sp, K, T, C, n, F, Kpad, KPH, MT, S, spw, TCS, mean, itr, spectra_in, sig,
                                            bhi, blo, blockIdx.x, lds);
  // the Gaussian transform on the way out, by every wavefront: element e of the span belongs to sample r = e / TC at
  // rem = e % TC = t C + c
  const int tid = threadIdx.x, nthr = blockDim.x, TC = ct.TC;
  auto transform = [&](int e) -> float {
    const int r = colored_row(ct, e);
    return colored_value(ct, r, e - r * TC, pstart);
  };
  float* dst = out + (size_t)ct.k0 * TC;
  const int total = ct.nsamp * TC;
  if ((((uintptr_t)dst) & 15) == 0) {
    const int n4 = total >> 2;
    for (int i = tid; i < n4; i += nthr)
      reinterpret_cast<float4*>(dst)[i] =
          make_float4(transform(4 * i), transform(4 * i + 1), transform(4 * i + 2), transform(4 * i + 3));
    for (int e = (n4 << 2) + tid; e < total; e += nthr) dst[e] = transform(e);
  } else {
    for (int e = tid; e < total; e += nthr) dst[e] = transform(e);
  }
}

// The basis of one shape (n, T): built once per device on first use and kept until the process ends, read-only
// afterwards (so any stream may use it, and a graph captured with it keeps a valid pointer). The first use of a
// shape on a device must not be inside a graph capture (the controllers run their work once eagerly first).
struct ColoredBasis {
  _Float16* hi;
  _Float16* lo;
};

static int colored_basis(int n, int T, int F, int rows, int Kpad, hipStream_t s, ColoredBasis* out)
{
  int dev = 0;
  MPPI_HIP_CHECK(hipGetDevice(&dev));
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, ColoredBasis> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({dev, n, T});
  if (it == cache.end()) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MPPI_HIP_CHECK(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return fail(MPPI_E_INVALID, "colored sampler: the first launch of a horizon cannot be graph-captured");
    const size_t nb = (size_t)rows * Kpad;
    void* p = nullptr;
    MPPI_HIP_CHECK(hipMalloc(&p, 2 * nb * sizeof(_Float16)));
    ColoredBasis b{static_cast<_Float16*>(p), static_cast<_Float16*>(p) + nb};
    const int items = rows * (Kpad >> 1);
    colored_basis_kernel<<<(items + 255) / 256, 256, 0, s>>>(n, F, T, rows, Kpad, b.hi, b.lo);
    MPPI_HIP_CHECK(hipGetLastError());
    MPPI_HIP_CHECK(hipStreamSynchronize(s));  // complete before any stream reads it
    it = cache.emplace(std::make_tuple(dev, n, T), b).first;
  }
  *out = it->second;
  return MPPI_OK;
}

template <int NTT>
static int launch_colored_t(const mppi_sampler_params& sp, int K, int T, int C, const float* mean, float* samples,
                            IterRef it, const float* spectra_in, const ColoredPlan& pl, hipStream_t s)
{
  MPPI_CHECK(allow_max_dynamic_lds((const void*)colored_mfma_kernel<NTT>));
  const ColoredShape& sh = pl.sh;
  const int grid = (K + sh.spw - 1) / sh.spw;
  colored_mfma_kernel<NTT><<<grid, CTHREADS, sh.lds_bytes, s>>>(sp, K, T, C, sh.n, sh.F, sh.Kpad, sh.KPH, sh.MT,
                                                               sh.S, sh.spw, sh.TCS, mean, samples, it, spectra_in,
                                                               pl.pstart, pl.sig, pl.bhi, pl.blo);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// colored_shape per (T, C, transform length mode), computed once: the staging-stride search is not free
static ColoredShape colored_shape_cached(int T, int C, int mode)
{
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, ColoredShape> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({T, C, mode});
  if (it == cache.end()) it = cache.emplace(std::make_tuple(T, C, mode), colored_shape(T, C, mode)).first;
  return it->second;
}

int colored_plan(const mppi_sampler_params& sp, int K, int T, int C, hipStream_t s, ColoredPlan* out)
{
  ColoredPlan pl;
  pl.pstart = pure_noise_start(K, sp.pure_noise_trajectories_percentage);
  pl.sh = colored_shape_cached(T, C, sp.fft_len_mode);
  if (pl.sh.MT > 16) return fail(MPPI_E_SIZE, "colored sampler: horizon exceeds 256");
  if (pl.sh.lds_bytes > LDS_LIMIT || pl.sh.spw < 1)
    return fail(MPPI_E_SIZE, "colored sampler: horizon too long for the LDS staging");
  pl.sig = colored_sigma(sp, C, pl.sh.n);
  ColoredBasis b;
  MPPI_CHECK(colored_basis(pl.sh.n, T, pl.sh.F, pl.sh.MT * 16, pl.sh.Kpad, s, &b));
  pl.bhi = b.hi;
  pl.blo = b.lo;
  *out = pl;
  return MPPI_OK;
}

int launch_sample_colored(const mppi_sampler_params& sp, int K, int T, int C, const float* mean, float* samples,
                          IterRef it, const float* spectra_in, hipStream_t s)
{
  ColoredPlan pl;
  MPPI_CHECK(colored_plan(sp, K, T, C, s, &pl));
  switch (pl.sh.NS / 16) {
    case 4: return launch_colored_t<4>(sp, K, T, C, mean, samples, it, spectra_in, pl, s);
    case 2: return launch_colored_t<2>(sp, K, T, C, mean, samples, it, spectra_in, pl, s);
    default: return launch_colored_t<1>(sp, K, T, C, mean, samples, it, spectra_in, pl, s);
  }
}

}  // namespace mppi_amd
