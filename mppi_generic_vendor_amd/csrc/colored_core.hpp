// Colored-noise sampler core, shared by the standalone sampler (colored.hip) and the controller's fused iteration
// (fused.hip, which generates a workgroup's samples straight into its control tile). The algorithm and its reference
// (sampling_distributions/colored_noise/colored_noise.cu:286-392) are described in colored.hip.
#pragma once

#include "kernels.hpp"
#include "models.hpp"

namespace mppi_amd {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ IterArgs load_iter_c(const IterRef& r) { return r.ptr ? *r.ptr : r.val; }

constexpr int CKC = 32;        // k-step of the FP16 product (basis columns)
constexpr int CTHREADS = 1024; // 16 wavefronts

struct Sigma4 {
  float v[MPPI_MAX_CONTROL];
};

// Launch shape: NS series (NS / C samples) per workgroup of 16 wavefronts, MT = ceil(T / 16) row tiles, Kpad = 2F
// rounded up to the 32-wide k-step; wavefront w owns row tile w % MT against all NS / 16 series tiles over the k-group
// w / MT of S = min(16 / MT, Kpad / 32) k-groups (groups 1 .. S - 1 pass their sums to group 0 through LDS).
struct ColoredShape {
  int n, F, Kpad, KPH, MT, NS, spw, S, TCS;
  size_t lds_bytes;
};

// Everything a launch needs besides the buffers: the shape, sigma_c, the cached basis and the first pure-noise sample.
struct ColoredPlan {
  ColoredShape sh;
  Sigma4 sig;
  const _Float16* bhi;
  const _Float16* blo;
  int pstart;
};

// Plan for K samples of horizon T and C controls (builds the basis on the shape's first use, synchronously: not
// inside a graph capture).
int colored_plan(const mppi_sampler_params& sp, int K, int T, int C, hipStream_t s, ColoredPlan* out);

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// v = hi + lo to about 2^-22 relative: hi is v rounded to FP16, lo the rounded remainder. v is pinned to one FP32
// register first (an empty asm with a register operand, no memory access): otherwise the product that computes v is
// fused into one of the conversions (v_fma_mix rounds the exact product to FP16 once) and not into the other, and at
// FP16 ties the two roundings differ by one unit, which lo no longer compensates.
__device__ __forceinline__ void split_f16(float v, _Float16& hi, _Float16& lo)
{
  asm volatile("" : "+v"(v));
  hi = (_Float16)v;
  lo = (_Float16)(v - (float)hi);
}

// Power-of-two scales of the FP16 operands (exact): the spectra |Y| <= 6-ish and the basis |B| <= 2 are multiplied by
// 2^8 and 2^12, so that the low part (about 2^-11 of the value) stays a normal FP16 number down to |Y| ~ 5e-4 and
// |B| ~ 3e-5, and the largest products stay far inside the FP32 range. Fed spectra must stay below 255 in magnitude.
constexpr float Y_SCALE = 256.0f, B_SCALE = 4096.0f;

// Column pair f of the inverse real transform's row t: w_f cos(2 pi f t / n) and -w_f sin(2 pi f t / n), the phase
// f t reduced modulo n in integers; zero where the transform ignores the term (f >= F, the imaginary parts at DC and
// Nyquist and the last bin that configureFrequencyNoise zeroes when F is odd).
__device__ __forceinline__ void basis_pair(int n, int F, int t, int f, float& re, float& im)
{
  re = 0.0f;
  im = 0.0f;
  if (f >= F) return;
  const int m = (int)(((long long)f * t) % n);
  float s, c;
  sincospif((float)(2 * m) / (float)n, &s, &c);
  const bool edge = (f == 0) || (2 * f == n);
  const float w = edge ? 1.0f : 2.0f;
  re = w * c;
  const bool zero_im = edge || ((F & 1) && f == F - 1);
  im = zero_im ? 0.0f : -w * s;
}

// Stages the 4-block of normals z starting at element e0 (even; -2 for the first block when the span does not start
// on a 4-boundary) of the workgroup's span of cnt spectrum values: TF is even, so (z0, z1) and (z2, z3) are each one
// (Re, Im) pair of one series, scaled and split into FP16 high and low parts.
__device__ __forceinline__ void stage_spectra(const float* z, int e0, int cnt, int TF, int C, int KPH, float invTF,
                                              float invC, const float* scale, _Float16* Xh, _Float16* Xl)
{
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int e = e0 + 2 * p;
    if (e >= 0 && e < cnt) {
      // e / TF and ns % C by a float reciprocal: exact, the quotients' fractional parts are >= 1 / (2 TF) from 1
      const int ns = (int)(((float)e + 0.5f) * invTF);
      const int jj = e - ns * TF;
      const int c = ns - (int)(((float)ns + 0.5f) * invC) * C;
      const float s = scale[(jj >> 1) * C + c];
      _Float16 h0, l0, h1, l1;
      split_f16(z[2 * p] * s, h0, l0);
      split_f16(z[2 * p + 1] * s, h1, l1);
      const half2v h = {h0, h1}, l = {l0, l1};
      *reinterpret_cast<half2v*>(Xh + ns * KPH + jj) = h;
      *reinterpret_cast<half2v*>(Xl + ns * KPH + jj) = l;
    }
  }
}

constexpr int KPRE = 4;  // k-steps whose basis fragments are held in registers from the kernel's start

// One 32-wide k-step at column kc0: acc[q] += Bh Yh + Bh Yl + Bl Yh for the series tiles q < NTT (Y from LDS).
template <int NTT>
__device__ __forceinline__ void product_step(const half8& ah, const half8& al, const _Float16* xh, const _Float16* xl,
                                             int KPH, int kc0, floatx4* acc)
{
#pragma unroll
  for (int q = 0; q < NTT; ++q) {
    const half8 bh = *reinterpret_cast<const half8*>(xh + q * 16 * KPH + kc0);
    const half8 bl = *reinterpret_cast<const half8*>(xl + q * 16 * KPH + kc0);
    acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[q], 0, 0, 0);
    acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[q], 0, 0, 0);
    acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[q], 0, 0, 0);
  }
}

// What colored_stage leaves in LDS for the Gaussian transform: sd_c eps of sample row r at stage[r TCS + t C + c]
// and the sampling mean at mean_l[t C + c], for the nsamp samples k0 .. k0 + nsamp - 1.
struct ColoredTile {
  const float* stage;
  const float* mean_l;
  int TCS, TC, k0, nsamp, sc;
  float invTC;
};

// Per workgroup (CTHREADS threads, `block` = the index of its span of spw samples, `lds` its LDS region of
// lds_bytes): the scaled spectra of its series (every scale divided by the component's largest, so the values stay
// far inside the FP16 range; the factor returns in the epilogue), staged in LDS as FP16 high and low parts; then
// X = B Y on v_mfma_f32_16x16x32_f16 as Bh Yh + Bh Yl + Bl Yh (the dropped Bl Yl term is about 2^-22 relative),
// accumulated in FP32, with B the stride-free basis. Each basis fragment is read once per workgroup (the first KPRE
// k-steps' fragments before the spectra are generated) and applied to all NTT series tiles. The epilogue forms
// x[t] - decay^t x[stride]: for stride < T, x[stride] is the product's own row (at t = stride the difference then
// carries no rounding of its own beyond one FMA); for stride >= T it is a dot product of the staged spectra with
// that row of the basis, done in FP32. Ends with a barrier.
template <int NTT>
__device__ __forceinline__ ColoredTile colored_stage(const mppi_sampler_params& sp, int K, int T, int C, int n, int F,
                                                     int Kpad, int KPH, int MT, int S, int spw, int TCS,
                                                     const float* __restrict__ mean, IterRef itr,
                                                     const float* __restrict__ spectra_in, const Sigma4& sig,
                                                     const _Float16* __restrict__ bhi,
                                                     const _Float16* __restrict__ blo, int block, float* lds)
{
  constexpr int NS = NTT * 16;
  const int TC = T * C;
  const int FCp = (F * C + 3) & ~3;
  float* scale = lds;                  // [F][C]  Y_SCALE s_c(f) / smax_c
  float* fac = scale + FCp;            // [8]     inv[4] = smax_c / (sigma_c n Y_SCALE B_SCALE), sd[4]
  float* mean_l = fac + 8;             // [T][C]
  float* dpow = mean_l + TC;           // [T]     decay^t (0 for a zero decay rate)
  float* xs = dpow + T;                // [64]    x[stride] per series, in the product's units
  float* bs = xs + 64;                 // [Kpad]  the basis row of the stride, used when stride >= T
  const int NTAB = (FCp + 8 + TC + T + 64 + Kpad + 3) & ~3;
  _Float16* Xh = reinterpret_cast<_Float16*>(lds + NTAB);  // [NS][KPH] high parts
  _Float16* Xl = Xh + NS * KPH;                            // [NS][KPH] low parts
  float* stage = lds + NTAB;                               // [spw][TCS], after the product

  const IterArgs a = load_iter_c(itr);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int k0 = block * spw;
  const int nsamp = min(spw, K - k0);
  const int nser = nsamp * C;
  const int TF = 2 * F;
  const int stride = min(max(a.stride, 0), n - 1);
  const bool decay = sp.offset_decay_rate != 0.0f;

  // wavefront w: row tile m = w % MT against all series tiles over the k-steps [kb, ke) of k-group w / MT; lane
  // (lr, lk) supplies row lr and k-columns 8 lk .. 8 lk + 7 of each 32-wide k-step
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int m = wave % MT, sg = wave / MT;
  const bool active = sg < S;
  const int ks = Kpad / CKC;
  const int kb = active ? (sg * ks) / S : 0, ke = active ? ((sg + 1) * ks) / S : 0;
  const _Float16* gh = bhi + (size_t)(m * 16 + lr) * Kpad + 8 * lk;
  const _Float16* gl = blo + (size_t)(m * 16 + lr) * Kpad + 8 * lk;
  // the mean first (T C <= 1024 threads; the memory counter retires in order, so its wait below does not cover the
  // basis loads), then the basis fragments of the first KPRE k-steps, whose latency passes while the tables and the
  // spectra are computed
  const float mv = mean[min(tid, TC - 1)];
  __builtin_amdgcn_sched_barrier(0);
  half8 ph[KPRE], pl[KPRE];
#pragma unroll
  for (int i = 0; i < KPRE; ++i) {
    const int k = min(kb + i, ks - 1);
    ph[i] = *reinterpret_cast<const half8*>(gh + k * CKC);
    pl[i] = *reinterpret_cast<const half8*>(gl + k * CKC);
  }
  floatx4 acc[NTT];
#pragma unroll
  for (int q = 0; q < NTT; ++q) acc[q] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  // tables: power-law scales (low frequencies take the first admissible one; f^(-beta/2) is monotone in f, so the
  // largest scale is at the lowest or the highest admissible frequency), epilogue factors, decay powers
  {
    const float cutoff = fmaxf(sp.fmin, 1.0f / n);
    int j0 = 0;
    while (j0 < F && (j0 / (1.0f * n)) < cutoff) ++j0;
    const int lo = (j0 < F) ? j0 : 0;
    for (int i = tid; i < F * C; i += nthr) {
      const int fi = i / C, c = i - fi * C;
      const int src = (fi < j0 && j0 < F) ? j0 : fi;
      const float e = -sp.exponents[c] / 2.0f;
      const float smax = fmaxf(powf(lo / (1.0f * n), e), powf((F - 1) / (1.0f * n), e));
      scale[i] = powf(src / (1.0f * n), e) / smax * Y_SCALE;
      if (fi == 0) {
        fac[c] = smax / (sig.v[c] * n) / (Y_SCALE * B_SCALE);
        fac[4 + c] = powf(sp.std_dev_decay, (float)a.iteration) * sp.std_dev[c];
      }
    }
    for (int t = tid; t < T; t += nthr) dpow[t] = decay ? powf(sp.offset_decay_rate, (float)t) : 0.0f;
    if (decay && stride >= T) {
      for (int f = tid; 2 * f < Kpad; f += nthr) {
        float re, im;
        basis_pair(n, F, stride, f, re, im);
        bs[2 * f] = re * B_SCALE;
        bs[2 * f + 1] = im * B_SCALE;
      }
    }
    if (!decay && tid < 64) xs[tid] = 0.0f;
    // unconditional store (threads past the end write the last entry's own value again), so that the load is not
    // sunk behind the basis loads
    mean_l[min(tid, TC - 1)] = mv;
  }

  // the first 4-block of normals is generated while the basis loads are in flight
  const int64_t base = (int64_t)k0 * C * TF;
  const int64_t cnt = (int64_t)nser * TF;
  const int64_t gb0 = (base >> 2) + tid;
  Normal4 nz0;
  if (spectra_in == nullptr) nz0 = normals4(sp.seed, a.rng_offset, (uint64_t)gb0);
  __syncthreads();
  // scaled spectra of series s0 .. s0 + nser - 1 (a contiguous span of the [K][C][F][2] normal stream). TF is even
  // and so is base, so every (2j, 2j + 1) pair of a 4-block is one (Re, Im) pair of one series.
  const float invTF = 1.0f / TF, invC = 1.0f / C;
  // one loop per source, so that the generated one holds no memory load (a load in the other path would make every
  // iteration wait for the basis fragments in flight)
  if (spectra_in != nullptr) {
    for (int64_t gb = (base >> 2) + tid; gb * 4 < base + cnt; gb += nthr) {
      float z[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t g = gb * 4 + j;
        z[j] = (g >= base && g < base + cnt) ? spectra_in[g] : 0.0f;
      }
      stage_spectra(z, (int)(gb * 4 - base), (int)cnt, TF, C, KPH, invTF, invC, scale, Xh, Xl);
    }
  } else {
    if (gb0 * 4 < base + cnt) stage_spectra(nz0.z, (int)(gb0 * 4 - base), (int)cnt, TF, C, KPH, invTF, invC, scale, Xh, Xl);
    for (int64_t gb = gb0 + nthr; gb * 4 < base + cnt; gb += nthr) {
      const Normal4 nz = normals4(sp.seed, a.rng_offset, (uint64_t)gb);
      stage_spectra(nz.z, (int)(gb * 4 - base), (int)cnt, TF, C, KPH, invTF, invC, scale, Xh, Xl);
    }
  }
  // zero padding: columns [2F, Kpad) of the valid series, whole rows of the unused series (pairs of halves)
  {
    const int padp = (Kpad - TF) >> 1;
    const half2v zz = {(_Float16)0.0f, (_Float16)0.0f};
    for (int i = tid; i < nser * padp; i += nthr) {
      const int r = i / padp, col = TF + 2 * (i - r * padp);
      *reinterpret_cast<half2v*>(Xh + r * KPH + col) = zz;
      *reinterpret_cast<half2v*>(Xl + r * KPH + col) = zz;
    }
    for (int i = tid; i < ((NS - nser) * KPH) >> 1; i += nthr) {
      *reinterpret_cast<half2v*>(Xh + nser * KPH + 2 * i) = zz;
      *reinterpret_cast<half2v*>(Xl + nser * KPH + 2 * i) = zz;
    }
  }
  __syncthreads();

  // x[stride] for stride >= T: 16 lanes per series, each over the columns j = part (mod 16), summed in a fixed order
  if (decay && stride >= T) {
    for (int sidx = tid >> 4; sidx < NS; sidx += nthr >> 4) {
      const int part = tid & 15;
      float v = 0.0f;
      for (int j = part; j < Kpad; j += 16)
        v += bs[j] * ((float)Xh[sidx * KPH + j] + (float)Xl[sidx * KPH + j]);
#pragma unroll
      for (int o = 8; o >= 1; o >>= 1) v += __shfl_xor(v, o, 16);
      if (part == 0) xs[sidx] = v;
    }
  }

  // product: the k-steps beyond the prefetched ones stream their basis fragments from memory
  if (active) {
    const _Float16* xh = Xh + lr * KPH + 8 * lk;
    const _Float16* xl = Xl + lr * KPH + 8 * lk;
#pragma unroll
    for (int i = 0; i < KPRE; ++i)
      if (kb + i < ke) product_step<NTT>(ph[i], pl[i], xh, xl, KPH, (kb + i) * CKC, acc);
    for (int k = kb + KPRE; k < ke; ++k)
      product_step<NTT>(*reinterpret_cast<const half8*>(gh + k * CKC), *reinterpret_cast<const half8*>(gl + k * CKC),
                        xh, xl, KPH, k * CKC, acc);
  }
  __syncthreads();  // the spectra space is reused as the output staging below

  // k-groups 1 .. S - 1 hand their partial sums to group 0 through LDS (lane-major 16-byte slots), which adds them in
  // order. Lane holds X[t = 16 m + 4 lk + v][series 16 q + lr] in acc[q][v].
  floatx4* part = reinterpret_cast<floatx4*>(Xl + NS * KPH);  // [S - 1][MT][NTT][64]
  if (S > 1) {
    if (active && sg > 0) {
#pragma unroll
      for (int q = 0; q < NTT; ++q) part[(((sg - 1) * MT + m) * NTT + q) * 64 + lane] = acc[q];
    }
    __syncthreads();
  }
  if (active && sg == 0) {
    for (int g = 1; g < S; ++g) {
#pragma unroll
      for (int q = 0; q < NTT; ++q) acc[q] += part[(((g - 1) * MT + m) * NTT + q) * 64 + lane];
    }
  }
  // x[stride] for stride < T: the row's values, from the wavefront and lanes that hold it
  if (decay && stride < T) {
    if (active && sg == 0 && m == (stride >> 4) && lk == ((stride & 15) >> 2)) {
#pragma unroll
      for (int q = 0; q < NTT; ++q) xs[q * 16 + lr] = acc[q][stride & 3];
    }
    __syncthreads();
  }
  if (active && sg == 0) {
    const int cdiv = 65536 / C + 1;  // (x * cdiv) >> 16 == x / C for 0 <= x < 64
#pragma unroll
    for (int q = 0; q < NTT; ++q) {
      const int sidx = q * 16 + lr;
      const int r = (sidx * cdiv) >> 16, c = sidx - r * C;
      const float inv = fac[c], sd = fac[4 + c], xv = xs[sidx];
      const bool valid = sidx < nser;
      float* srow = stage + r * TCS + c;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int t = m * 16 + lk * 4 + v;
        // per-step standard deviations scale each step's noise (setGaussianControls after the colored draw,
        // colored_noise.cu:367-386 with time_specific_std_dev)
        const float sdt = sp.std_dev_t ? powf(sp.std_dev_decay, (float)a.iteration) * sp.std_dev_t[min(t, T - 1) * C + c]
                                       : sd;
        if (valid && t < T) srow[t * C] = sdt * ((acc[q][v] - dpow[t] * xv) * inv);
      }
    }
  }
  __syncthreads();
  ColoredTile ct;
  ct.stage = stage;
  ct.mean_l = mean_l;
  ct.TCS = TCS;
  ct.TC = TC;
  ct.k0 = k0;
  ct.nsamp = nsamp;
  ct.sc = a.stride * C;
  ct.invTC = 1.0f / TC;  // e / TC by the reciprocal is exact: e < 64 TC, the fraction is >= 1 / (2 TC) from 1
  return ct;
}

// Sample row of element e of the workgroup's span.
__device__ __forceinline__ int colored_row(const ColoredTile& ct, int e) { return (int)(((float)e + 0.5f) * ct.invTC); }

// The Gaussian transform of sample k's element with mean mu and staged noise v: sample 0 and t < stride (`head`) keep
// the mean, pure-noise samples (k >= pstart) are the noise alone.
__device__ __forceinline__ float colored_transform(float mu, float v, int k, bool head, int pstart)
{
  const float u = (k >= pstart) ? v : mu + v;
  return (k == 0 || head) ? mu : u;
}

// The same for element rem = t C + c of sample row r of the tile.
__device__ __forceinline__ float colored_value(const ColoredTile& ct, int r, int rem, int pstart)
{
  return colored_transform(ct.mean_l[rem], ct.stage[r * ct.TCS + rem], ct.k0 + r, rem < ct.sc, pstart);
}

}  // namespace mppi_amd
