// Learned proposal network on the GPU: forward, loss, backward and Adam with hand-written kernels.
//
// Every matrix product goes through gemm_bf16_kernel: 64 x 64 output tiles per 256-thread workgroup, the A and B
// tiles staged in LDS as BF16 (converted from the FP32 tensors on load, either operand optionally transposed through
// its strides), v_mfma_f32_16x16x32_bf16 with FP32 accumulation, and a fused epilogue (bias; bias + GELU keeping the
// pre-activation for the backward pass; or multiplication by GELU'(z) of a saved pre-activation). The loss is the
// Gaussian negative log-likelihood of the optimal-sequence residual (see the oracle's header for its exact form).
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "host.hpp"
#include "proposal.hpp"

namespace mppi_amd {

PropDims prop_dims(int S, int C, int T)
{
  PropDims d;
  d.S = S;
  d.C = C;
  d.T = T;
  d.din = S + PROP_KNOTS * C;
  d.dout = 2 * T * C;
  d.w1 = 0;
  d.b1 = d.w1 + (size_t)PROP_H * d.din;
  d.w2 = d.b1 + PROP_H;
  d.b2 = d.w2 + (size_t)PROP_H * PROP_H;
  d.w3 = d.b2 + PROP_H;
  d.b3 = d.w3 + (size_t)d.dout * PROP_H;
  d.total = d.b3 + d.dout;
  return d;
}

typedef __bf16 pbf16x8 __attribute__((ext_vector_type(8)));
typedef float pf32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float dgelu_f(float z)
{
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

enum { EPI_BIAS = 0, EPI_BIAS_GELU = 1, EPI_DGELU = 2 };
constexpr int GT = 64, GK = 32, GKP = GK + 8;  // tile, k step, padded LDS row (80 bytes: 16-byte aligned rows)

// C[m][n] = sum_k A(m, k) B(k, n) with A(m, k) = A[m sam + k sak], B(k, n) = B[k sbk + n sbn], then the epilogue:
//   EPI_BIAS       C = acc (+ bias[n])
//   EPI_BIAS_GELU  C = z = acc + bias[n], aux = gelu(z)
//   EPI_DGELU      C = acc * gelu'(aux)
template <int EPI>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const float* __restrict__ A, int64_t sam,
                                                         int64_t sak, const float* __restrict__ B, int64_t sbk,
                                                         int64_t sbn, const float* __restrict__ bias,
                                                         float* __restrict__ C, int ldc, float* __restrict__ aux,
                                                         int ldaux)
{
  __shared__ __attribute__((aligned(16))) __bf16 As[GT][GKP];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[GT][GKP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int g = lane >> 4, l16 = lane & 15;
  pf32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = pf32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const bool a_kfast = sak == 1, b_kfast = sbk == 1;
  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      const int mm = a_kfast ? e / GK : e % GT, kk = a_kfast ? e % GK : e / GT;
      const int m = m0 + mm, k = k0 + kk;
      As[mm][kk] = (__bf16)((m < M && k < K) ? A[m * sam + k * sak] : 0.0f);
      const int nn = b_kfast ? e / GK : e % GT, kb = b_kfast ? e % GK : e / GT;
      const int n = n0 + nn, k2 = k0 + kb;
      Bs[nn][kb] = (__bf16)((n < N && k2 < K) ? B[k2 * sbk + n * sbn] : 0.0f);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const pbf16x8 a = *reinterpret_cast<const pbf16x8*>(&As[wm + 16 * i + l16][8 * g]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const pbf16x8 b = *reinterpret_cast<const pbf16x8*>(&Bs[wn + 16 * j + l16][8 * g]);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * g + r, n = n0 + wn + 16 * j + l16;
        if (m >= M || n >= N) continue;
        float v = acc[i][j][r];
        if (EPI == EPI_BIAS) {
          C[(int64_t)m * ldc + n] = v + (bias ? bias[n] : 0.0f);
        } else if (EPI == EPI_BIAS_GELU) {
          v += bias[n];
          C[(int64_t)m * ldc + n] = v;
          aux[(int64_t)m * ldaux + n] = gelu_f(v);
        } else {
          C[(int64_t)m * ldc + n] = v * dgelu_f(aux[(int64_t)m * ldaux + n]);
        }
      }
}

template <int EPI>
static int gemm(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk,
                int64_t sbn, const float* bias, float* C, int ldc, float* aux, int ldaux, hipStream_t s)
{
  const dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
  gemm_bf16_kernel<EPI><<<grid, 256, 0, s>>>(M, N, K, A, sam, sak, B, sbk, sbn, bias, C, ldc, aux, ldaux);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// Loss terms and dL/dY; one thread per (sample, output pair j < TC); per-block FP64 partial sums of the loss.
__global__ void __launch_bounds__(256) proposal_loss_kernel(const float* __restrict__ Y, int dout,
                                                            const float* __restrict__ R, int B, int TC,
                                                            float* __restrict__ dY, double* __restrict__ partial)
{
  __shared__ double red[4];
  const int idx = blockIdx.x * 256 + threadIdx.x;
  const double inv = 1.0 / ((double)TC * B);
  double l = 0.0;
  if (idx < B * TC) {
    const int b = idx / TC, j = idx - b * TC;
    const float y = Y[(int64_t)b * dout + j];
    const float s = Y[(int64_t)b * dout + TC + j];
    const float sc = fminf(fmaxf(s, PROP_S_LO), PROP_S_HI);
    const float e = R[(int64_t)b * TC + j] - y;
    const float iv = expf(-2.0f * sc);
    l = (0.5 * (double)e * e * iv + sc) * inv;
    dY[(int64_t)b * dout + j] = (float)(-(double)e * iv * inv);
    const bool inside = s > PROP_S_LO && s < PROP_S_HI;
    dY[(int64_t)b * dout + TC + j] = inside ? (float)((1.0 - (double)e * e * iv) * inv) : 0.0f;
  }
  l = wave_sum(l);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) sum_partials_kernel(const double* __restrict__ partial, int n,
                                                           double* __restrict__ out)
{
  __shared__ double red[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += partial[i];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

// out[n] = sum_m X[m][n] (column sums in row order: the bias gradients)
__global__ void __launch_bounds__(256) colsum_kernel(const float* __restrict__ X, int M, int N, int ld,
                                                     float* __restrict__ out)
{
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += X[(int64_t)m * ld + n];
  out[n] = s;
}

__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ th, float* __restrict__ m,
                                                   float* __restrict__ v, const float* __restrict__ g, int P, float lr,
                                                   float b1, float b2, float eps, float c1, float c2)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] = th[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of the three weight matrices for the controller's inference, rows padded with zeros to din8 (a multiple
// of 8: 16-byte rows) for W1: W1 [PROP_H][din8], W2 [PROP_H][PROP_H], W3 [dout][PROP_H], each element the round to
// nearest even of the FP32 parameter (the operand rounding of every BF16 product here).
__global__ void __launch_bounds__(256) mirror_bf16_kernel(const float* __restrict__ th, PropDims d, int din8,
                                                          __bf16* __restrict__ wb)
{
  const int n1 = PROP_H * din8, n2 = PROP_H * PROP_H, n3 = d.dout * PROP_H;
  const int i = blockIdx.x * 256 + threadIdx.x;
  float v = 0.0f;
  if (i < n1) {
    const int r = i / din8, c = i - r * din8;
    v = c < d.din ? th[d.w1 + (size_t)r * d.din + c] : 0.0f;
  } else if (i < n1 + n2) {
    v = th[d.w2 + (i - n1)];
  } else if (i < n1 + n2 + n3) {
    v = th[d.w3 + (i - n1 - n2)];
  } else {
    return;
  }
  wb[i] = (__bf16)v;
}

// Network input from the state and the nominal sequence (16 knots t_j = j (T - 1) / 15), normalized.
__global__ void proposal_input_kernel(int S, int C, int T, const float* __restrict__ x0,
                                      const float* __restrict__ nominal, const float* __restrict__ shift,
                                      const float* __restrict__ scale, float* __restrict__ out)
{
  const int o = threadIdx.x;
  const int din = S + PROP_KNOTS * C;
  if (o >= din) return;
  float v;
  if (o < S) {
    v = x0[o];
  } else {
    const int j = (o - S) / C, c = (o - S) - j * C;
    v = nominal[(j * (T - 1) / 15) * C + c];
  }
  out[o] = (v - shift[o]) * scale[o];
}

// Absolute mean shift and proposal std from the network output of one state.
__global__ void proposal_head_kernel(const float* __restrict__ y, int TC, int C, const float* __restrict__ sigma_base,
                                     float* __restrict__ dmu, float* __restrict__ sig)
{
  for (int j = threadIdx.x; j < TC; j += blockDim.x) {
    const float sb = sigma_base[j % C];
    dmu[j] = sb * y[j];
    sig[j] = sb * expf(fminf(fmaxf(y[TC + j], PROP_S_LO), PROP_S_HI));
  }
}

// Inference for one state: workgroup b of a grid of ceil(2 T C / 16) evaluates the two hidden layers (wavefront w
// holds hidden tile w) and output tile b (wavefront 0), then the head of those 16 outputs. The hidden layers are
// repeated per workgroup so that no workgroup waits for another; what bounds the kernel is the weight traffic one
// CU can keep in flight, and this split brings it from all three layers down to the hidden layers plus one output
// tile per CU. The operands are rounded to BF16 and accumulated in FP32 over 32-wide k-steps in order, the arithmetic
// of gemm_bf16_kernel, so the result equals the batched forward pass.
constexpr int PI_WAVES = PROP_H / 16;
constexpr int PI_THREADS = 64 * PI_WAVES;
constexpr int PI_MAX_TILES = 64;  // output tiles (grid size): 2 T C <= 1024

// Weights of the 16-row tile `mt` of an N x Kreal layer for this lane (row 16 mt + lane % 16, K slots of the lane
// group), loaded as FP32 (16-byte loads when the row stride allows) and rounded to BF16 A-operand fragments.
template <int KS>
__device__ __forceinline__ void infer_load_tile(const float* __restrict__ W, int N, int Kreal, int mt,
                                                pbf16x8 (&a)[KS])
{
  const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
  const int row = 16 * mt + l16;
  if ((Kreal & 3) == 0) {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = 32 * ks + 8 * g;
      float4 v0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v1 = v0;
      if (row < N && k < Kreal) {
        const float4* p4 = reinterpret_cast<const float4*>(W + (int64_t)row * Kreal + k);
        v0 = p4[0];
        v1 = p4[1];
      }
      a[ks][0] = (__bf16)v0.x;
      a[ks][1] = (__bf16)v0.y;
      a[ks][2] = (__bf16)v0.z;
      a[ks][3] = (__bf16)v0.w;
      a[ks][4] = (__bf16)v1.x;
      a[ks][5] = (__bf16)v1.y;
      a[ks][6] = (__bf16)v1.z;
      a[ks][7] = (__bf16)v1.w;
    }
  } else {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = 32 * ks + 8 * g + j;
        a[ks][j] = (__bf16)((row < N && k < Kreal) ? W[(int64_t)row * Kreal + k] : 0.0f);
      }
  }
}

// The same fragments from a BF16 copy with row stride ld (a multiple of 8, zero beyond the layer's width): one 16-byte
// load per k-step.
template <int KS>
__device__ __forceinline__ void infer_load_tile_b(const __bf16* __restrict__ W, int N, int ld, int mt,
                                                  pbf16x8 (&a)[KS])
{
  const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
  const int row = 16 * mt + l16;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k = 32 * ks + 8 * g;
    pbf16x8 v = {};
    if (row < N && k < ld) v = *reinterpret_cast<const pbf16x8*>(W + (int64_t)row * ld + k);
    a[ks] = v;
  }
}

// One layer for the input vector vin (LDS, broadcast over the 16 B columns): vout[r - row0] = act(W[r] . vin + bias[r])
// for the tiles whose A fragments `a` this wave holds (tile mt[i] for i < NT).
template <int KS, int NT>
__device__ __forceinline__ void infer_apply(const pbf16x8 (&a)[NT][KS], const int (&mt)[NT], const float* bias,
                                            int N, int Kreal, const float* vin, float* vout, bool gelu, int row0 = 0)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
  pbf16x8 b[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 32 * ks + 8 * g + j;
      b[ks][j] = (__bf16)(k < Kreal ? vin[k] : 0.0f);
    }
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    if (mt[i] * 16 >= N) continue;
    pf32x4 acc = pf32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][ks], b[ks], acc, 0, 0, 0);
    if (l16 == 0) {  // column 0 of the result
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const int r = 16 * mt[i] + 4 * g + r4;
        if (r < N) {
          const float z = acc[r4] + bias[r];
          vout[r - row0] = gelu ? gelu_f(z) : z;
        }
      }
    }
  }
}

template <int DIN>
__global__ void __launch_bounds__(PI_THREADS) proposal_infer_kernel(
    int S, int C, int T, const float* __restrict__ th, int64_t w1, int64_t b1, int64_t w2, int64_t b2, int64_t w3,
    int64_t b3, const float* __restrict__ x0, const float* __restrict__ nominal, const float* __restrict__ shift,
    const float* __restrict__ scale, const float* __restrict__ sigma_base, float* __restrict__ dmu,
    float* __restrict__ sig, const __bf16* __restrict__ wb, int din8, InferStaging stg)
{
  __shared__ float xin[DIN];
  __shared__ float h1[PROP_H], h2[PROP_H];
  __shared__ float yout[16];
  constexpr int KS1 = DIN / 32, KSH = PROP_H / 32;
  const int wave = threadIdx.x >> 6;
  const int din = S + PROP_KNOTS * C, TC = T * C, dout = 2 * TC;
  const int tile = blockIdx.x;
  // every wave requests all the weights it will multiply before anything else (they do not depend on the input)
  const int mth[1] = {wave}, mt3[1] = {tile};
  pbf16x8 a1[1][KS1], a2[1][KSH], a3[1][KSH];
  // weights from the BF16 copy (the same operand values as rounding the FP32 parameters, half the bytes)
  const __bf16* w1b = wb;
  const __bf16* w2b = w1b + PROP_H * din8;
  const __bf16* w3b = w2b + PROP_H * PROP_H;
  infer_load_tile_b<KS1>(w1b, PROP_H, din8, wave, a1[0]);
  infer_load_tile_b<KSH>(w2b, PROP_H, PROP_H, wave, a2[0]);
  if (wave == 0) infer_load_tile_b<KSH>(w3b, dout, PROP_H, tile, a3[0]);
  for (int o = threadIdx.x; o < DIN; o += PI_THREADS) {
    float v = 0.0f;
    if (o < S) {
      v = (x0[o] - shift[o]) * scale[o];
    } else if (o < din) {
      const int j = (o - S) / C, c = (o - S) - j * C;
      v = (nominal[(j * (T - 1) / 15) * C + c] - shift[o]) * scale[o];
    }
    xin[o] = v;
  }
  if (stg.nominal_dst != nullptr && tile == 0) {
    for (int j = threadIdx.x; j < TC; j += PI_THREADS) stg.nominal_dst[j] = nominal[j];
    if (threadIdx.x < S) stg.x0_dst[threadIdx.x] = x0[threadIdx.x];
    if (threadIdx.x == 0) *stg.iter_dst = *stg.iter_src;
  }
  __syncthreads();
  infer_apply<KS1, 1>(a1, mth, th + b1, PROP_H, din, xin, h1, true);
  __syncthreads();
  infer_apply<KSH, 1>(a2, mth, th + b2, PROP_H, PROP_H, h1, h2, true);
  __syncthreads();
  if (wave == 0) {
    // output rows 16 tile .. 16 tile + 15, stored at yout[row - 16 tile]
    infer_apply<KSH, 1>(a3, mt3, th + b3, dout, PROP_H, h2, yout, false, 16 * tile);
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int r = 16 * tile + threadIdx.x;
    if (r < TC) {
      dmu[r] = sigma_base[r % C] * yout[threadIdx.x];
    } else if (r < dout) {
      const int j = r - TC;
      sig[j] = sigma_base[j % C] * expf(fminf(fmaxf(yout[threadIdx.x], PROP_S_LO), PROP_S_HI));
    }
  }
}

struct ProposalNet {
  PropDims d{};
  float* theta = nullptr;
  __bf16* wb = nullptr;  // BF16 weight copy for the inference (mirror_bf16_kernel), rewritten with theta
  int din8 = 0;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grad = nullptr;
  float* shift = nullptr;
  float* scale = nullptr;
  int step = 0;
  int cap = 0;  // batch capacity of the activation buffers
  float *z1 = nullptr, *h1 = nullptr, *z2 = nullptr, *h2 = nullptr, *y = nullptr, *dy = nullptr;
  float *dz2 = nullptr, *dz1 = nullptr, *xin = nullptr;
  double* partial = nullptr;
  double* loss = nullptr;
  double* loss_h = nullptr;
};

static void prop_free_acts(ProposalNet* p)
{
  for (float* q : {p->z1, p->h1, p->z2, p->h2, p->y, p->dy, p->dz2, p->dz1, p->xin})
    if (q) (void)hipFree(q);
  if (p->partial) (void)hipFree(p->partial);
  p->z1 = p->h1 = p->z2 = p->h2 = p->y = p->dy = p->dz2 = p->dz1 = p->xin = nullptr;
  p->partial = nullptr;
  p->cap = 0;
}

static int prop_reserve(ProposalNet* p, int B)
{
  if (B <= p->cap) return MPPI_OK;
  MPPI_HIP_CHECK(hipDeviceSynchronize());
  prop_free_acts(p);
  const size_t hb = sizeof(float) * (size_t)B * PROP_H;
  int st = MPPI_OK;
  auto A = [&](int s2) {
    if (st == MPPI_OK) st = s2;
  };
  A(dev_alloc(&p->z1, hb));
  A(dev_alloc(&p->h1, hb));
  A(dev_alloc(&p->z2, hb));
  A(dev_alloc(&p->h2, hb));
  A(dev_alloc(&p->dz1, hb));
  A(dev_alloc(&p->dz2, hb));
  A(dev_alloc(&p->y, sizeof(float) * (size_t)B * p->d.dout));
  A(dev_alloc(&p->dy, sizeof(float) * (size_t)B * p->d.dout));
  A(dev_alloc(&p->xin, sizeof(float) * (size_t)p->d.din));
  A(dev_alloc(&p->partial, sizeof(double) * (((size_t)B * p->d.T * p->d.C + 255) / 256)));
  if (st != MPPI_OK) return st;
  p->cap = B;
  return MPPI_OK;
}

static int prop_forward(ProposalNet* p, const float* X, int B, hipStream_t s)
{
  const PropDims& d = p->d;
  const float* th = p->theta;
  MPPI_CHECK(gemm<EPI_BIAS_GELU>(B, PROP_H, d.din, X, d.din, 1, th + d.w1, 1, d.din, th + d.b1, p->z1, PROP_H, p->h1,
                                 PROP_H, s));
  MPPI_CHECK(gemm<EPI_BIAS_GELU>(B, PROP_H, PROP_H, p->h1, PROP_H, 1, th + d.w2, 1, PROP_H, th + d.b2, p->z2, PROP_H,
                                 p->h2, PROP_H, s));
  return gemm<EPI_BIAS>(B, d.dout, PROP_H, p->h2, PROP_H, 1, th + d.w3, 1, PROP_H, th + d.b3, p->y, d.dout, nullptr,
                        0, s);
}

static int prop_loss_grad(ProposalNet* p, const float* X, const float* R, int B, float* grad, hipStream_t s)
{
  const PropDims& d = p->d;
  const float* th = p->theta;
  const int TC = d.T * d.C;
  MPPI_CHECK(prop_forward(p, X, B, s));
  const int nb = (B * TC + 255) / 256;
  proposal_loss_kernel<<<nb, 256, 0, s>>>(p->y, d.dout, R, B, TC, p->dy, p->partial);
  sum_partials_kernel<<<1, 256, 0, s>>>(p->partial, nb, p->loss);
  MPPI_HIP_CHECK(hipGetLastError());
  // layer 3
  MPPI_CHECK(gemm<EPI_BIAS>(d.dout, PROP_H, B, p->dy, 1, d.dout, p->h2, PROP_H, 1, nullptr, grad + d.w3, PROP_H,
                            nullptr, 0, s));
  colsum_kernel<<<(d.dout + 255) / 256, 256, 0, s>>>(p->dy, B, d.dout, d.dout, grad + d.b3);
  MPPI_CHECK(gemm<EPI_DGELU>(B, PROP_H, d.dout, p->dy, d.dout, 1, th + d.w3, PROP_H, 1, nullptr, p->dz2, PROP_H, p->z2,
                             PROP_H, s));
  // layer 2
  MPPI_CHECK(gemm<EPI_BIAS>(PROP_H, PROP_H, B, p->dz2, 1, PROP_H, p->h1, PROP_H, 1, nullptr, grad + d.w2, PROP_H,
                            nullptr, 0, s));
  colsum_kernel<<<1, 256, 0, s>>>(p->dz2, B, PROP_H, PROP_H, grad + d.b2);
  MPPI_CHECK(gemm<EPI_DGELU>(B, PROP_H, PROP_H, p->dz2, PROP_H, 1, th + d.w2, PROP_H, 1, nullptr, p->dz1, PROP_H,
                             p->z1, PROP_H, s));
  // layer 1
  MPPI_CHECK(gemm<EPI_BIAS>(PROP_H, d.din, B, p->dz1, 1, PROP_H, X, d.din, 1, nullptr, grad + d.w1, d.din, nullptr, 0,
                            s));
  colsum_kernel<<<1, 256, 0, s>>>(p->dz1, B, PROP_H, PROP_H, grad + d.b1);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

int proposal_infer(ProposalNet* p, const float* x0, const float* nominal, const float* sigma_base, float* dmu,
                   float* sig, hipStream_t s, const InferStaging* staging)
{
  const InferStaging stg = staging != nullptr ? *staging : InferStaging{nullptr, nullptr, nullptr, nullptr};
  const PropDims& d = p->d;
  const int tiles = (d.dout + 15) / 16;
  if (tiles > PI_MAX_TILES) return fail(MPPI_E_SIZE, "proposal: output too wide for the inference kernel");
#define MPPI_INFER(DIN)                                                                                              \
  proposal_infer_kernel<DIN><<<tiles, PI_THREADS, 0, s>>>(d.S, d.C, d.T, p->theta, d.w1, d.b1, d.w2, d.b2, d.w3, d.b3, \
                                                          x0, nominal, p->shift, p->scale, sigma_base, dmu, sig,  \
                                                          p->wb, p->din8, stg)
  if (d.din <= 32) MPPI_INFER(32);
  else if (d.din <= 64) MPPI_INFER(64);
  else MPPI_INFER(96);
#undef MPPI_INFER
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

// The same inference through the batched kernels (input kernel, three GEMM launches, head), kept for checking the
// fused kernel against them.
int proposal_infer_unfused(ProposalNet* p, const float* x0, const float* nominal, const float* sigma_base,
                           float* dmu, float* sig, hipStream_t s)
{
  const PropDims& d = p->d;
  proposal_input_kernel<<<1, 128, 0, s>>>(d.S, d.C, d.T, x0, nominal, p->shift, p->scale, p->xin);
  MPPI_HIP_CHECK(hipGetLastError());
  MPPI_CHECK(prop_forward(p, p->xin, 1, s));
  proposal_head_kernel<<<1, 256, 0, s>>>(p->y, d.T * d.C, d.C, sigma_base, dmu, sig);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

}  // namespace mppi_amd

using namespace mppi_amd;

struct mppi_proposal {
  ProposalNet net;
};

ProposalNet* mppi_amd::proposal_net_of(mppi_proposal* q) { return &q->net; }
PropDims mppi_amd::proposal_dims(const ProposalNet* p) { return p->d; }

// Rewrites the BF16 weight copy from theta on stream s.
static int prop_mirror(ProposalNet* p, hipStream_t s)
{
  const int n = PROP_H * p->din8 + PROP_H * PROP_H + p->d.dout * PROP_H;
  mirror_bf16_kernel<<<(n + 255) / 256, 256, 0, s>>>(p->theta, p->d, p->din8, p->wb);
  MPPI_HIP_CHECK(hipGetLastError());
  return MPPI_OK;
}

static void proposal_free(mppi_proposal* q)
{
  if (!q) return;
  ProposalNet* p = &q->net;
  prop_free_acts(p);
  if (p->wb) (void)hipFree(p->wb);
  for (float* x : {p->theta, p->adam_m, p->adam_v, p->grad, p->shift, p->scale})
    if (x) (void)hipFree(x);
  if (p->loss) (void)hipFree(p->loss);
  if (p->loss_h) (void)hipHostFree(p->loss_h);
  delete q;
}

extern "C" {

int mppi_proposal_num_params(int32_t S, int32_t C, int32_t T, int32_t* out)
{
  if (!out) return fail(MPPI_E_INVALID, "null argument");
  if (S < 1 || S > MPPI_MAX_STATE || C < 1 || C > MPPI_MAX_CONTROL || T < 2 || T > MPPI_MAX_TIMESTEPS)
    return fail(MPPI_E_INVALID, "proposal: dimensions out of range");
  *out = (int32_t)prop_dims(S, C, T).total;
  return MPPI_OK;
}

int mppi_proposal_create(int32_t S, int32_t C, int32_t T, const float* params_host, mppi_proposal** out)
{
  if (!out) return fail(MPPI_E_INVALID, "null argument");
  *out = nullptr;
  int32_t np = 0;
  MPPI_CHECK(mppi_proposal_num_params(S, C, T, &np));
  mppi_proposal* q = new (std::nothrow) mppi_proposal();
  if (!q) return fail(MPPI_E_NOMEM, "proposal: out of host memory");
  ProposalNet* p = &q->net;
  p->d = prop_dims(S, C, T);
  const size_t P = p->d.total;
  int st = MPPI_OK;
  auto A = [&](int s2) {
    if (st == MPPI_OK) st = s2;
  };
  A(dev_alloc(&p->theta, sizeof(float) * P));
  p->din8 = (p->d.din + 7) & ~7;
  A(dev_alloc(&p->wb, sizeof(__bf16) * ((size_t)PROP_H * p->din8 + (size_t)PROP_H * PROP_H +
                                        (size_t)p->d.dout * PROP_H)));
  A(dev_alloc(&p->adam_m, sizeof(float) * P));
  A(dev_alloc(&p->adam_v, sizeof(float) * P));
  A(dev_alloc(&p->grad, sizeof(float) * P));
  A(dev_alloc(&p->shift, sizeof(float) * p->d.din));
  A(dev_alloc(&p->scale, sizeof(float) * p->d.din));
  A(dev_alloc(&p->loss, sizeof(double)));
  A(host_alloc(&p->loss_h, sizeof(double)));
  if (st == MPPI_OK) st = prop_reserve(p, 64);
  if (st != MPPI_OK) {
    proposal_free(q);
    return st;
  }
  hipError_t e = hipMemset(p->adam_m, 0, sizeof(float) * P);
  if (e == hipSuccess) e = hipMemset(p->adam_v, 0, sizeof(float) * P);
  if (e == hipSuccess) e = hipMemset(p->shift, 0, sizeof(float) * p->d.din);
  if (e == hipSuccess && params_host)
    e = hipMemcpy(p->theta, params_host, sizeof(float) * P, hipMemcpyHostToDevice);
  else if (e == hipSuccess)
    e = hipMemset(p->theta, 0, sizeof(float) * P);
  if (e == hipSuccess) {
    std::vector<float> ones(p->d.din, 1.0f);
    e = hipMemcpy(p->scale, ones.data(), sizeof(float) * p->d.din, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && prop_mirror(p, nullptr) != MPPI_OK) e = hipErrorLaunchFailure;
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    proposal_free(q);
    return hip_fail(e, "proposal: init");
  }
  *out = q;
  return MPPI_OK;
}

int mppi_proposal_destroy(mppi_proposal* q)
{
  (void)hipDeviceSynchronize();
  proposal_free(q);
  return MPPI_OK;
}

int mppi_proposal_get_params(const mppi_proposal* q, float* host)
{
  if (!q || !host) return fail(MPPI_E_INVALID, "null argument");
  MPPI_HIP_CHECK(hipDeviceSynchronize());
  MPPI_HIP_CHECK(hipMemcpy(host, q->net.theta, sizeof(float) * q->net.d.total, hipMemcpyDeviceToHost));
  return MPPI_OK;
}

int mppi_proposal_set_params(mppi_proposal* q, const float* host)
{
  if (!q || !host) return fail(MPPI_E_INVALID, "null argument");
  MPPI_HIP_CHECK(hipDeviceSynchronize());
  MPPI_HIP_CHECK(hipMemcpy(q->net.theta, host, sizeof(float) * q->net.d.total, hipMemcpyHostToDevice));
  MPPI_CHECK(prop_mirror(&q->net, nullptr));
  MPPI_HIP_CHECK(hipDeviceSynchronize());
  MPPI_HIP_CHECK(hipMemset(q->net.adam_m, 0, sizeof(float) * q->net.d.total));
  MPPI_HIP_CHECK(hipMemset(q->net.adam_v, 0, sizeof(float) * q->net.d.total));
  q->net.step = 0;
  return MPPI_OK;
}

int mppi_proposal_set_input_normalization(mppi_proposal* q, const float* shift, const float* scale)
{
  if (!q || !shift || !scale) return fail(MPPI_E_INVALID, "null argument");
  MPPI_HIP_CHECK(hipMemcpy(q->net.shift, shift, sizeof(float) * q->net.d.din, hipMemcpyHostToDevice));
  MPPI_HIP_CHECK(hipMemcpy(q->net.scale, scale, sizeof(float) * q->net.d.din, hipMemcpyHostToDevice));
  return MPPI_OK;
}

int mppi_proposal_forward(mppi_proposal* q, const float* inputs_d, int32_t B, float* outputs_d, void* stream)
{
  if (!q || !inputs_d || !outputs_d) return fail(MPPI_E_INVALID, "null argument");
  if (B < 1) return fail(MPPI_E_INVALID, "proposal: batch must be >= 1");
  ProposalNet* p = &q->net;
  MPPI_CHECK(prop_reserve(p, B));
  const hipStream_t s = (hipStream_t)stream;
  MPPI_CHECK(prop_forward(p, inputs_d, B, s));
  MPPI_HIP_CHECK(hipMemcpyAsync(outputs_d, p->y, sizeof(float) * (size_t)B * p->d.dout, hipMemcpyDeviceToDevice, s));
  return MPPI_OK;
}

int mppi_proposal_infer(mppi_proposal* q, const float* x0_d, const float* nominal_d, const float* sigma_base_d,
                        float* dmu_d, float* sig_d, int32_t fused, void* stream)
{
  if (!q || !x0_d || !nominal_d || !sigma_base_d || !dmu_d || !sig_d) return fail(MPPI_E_INVALID, "null argument");
  const hipStream_t s = (hipStream_t)stream;
  return fused ? proposal_infer(&q->net, x0_d, nominal_d, sigma_base_d, dmu_d, sig_d, s)
               : proposal_infer_unfused(&q->net, x0_d, nominal_d, sigma_base_d, dmu_d, sig_d, s);
}

int mppi_proposal_loss_grad(mppi_proposal* q, const float* inputs_d, const float* targets_d, int32_t B, float* grad_d,
                            double* loss_host, void* stream)
{
  if (!q || !inputs_d || !targets_d || !grad_d) return fail(MPPI_E_INVALID, "null argument");
  if (B < 1) return fail(MPPI_E_INVALID, "proposal: batch must be >= 1");
  ProposalNet* p = &q->net;
  MPPI_CHECK(prop_reserve(p, B));
  const hipStream_t s = (hipStream_t)stream;
  MPPI_CHECK(prop_loss_grad(p, inputs_d, targets_d, B, grad_d, s));
  if (loss_host) {
    MPPI_HIP_CHECK(hipMemcpyAsync(p->loss_h, p->loss, sizeof(double), hipMemcpyDeviceToHost, s));
    MPPI_HIP_CHECK(hipStreamSynchronize(s));
    *loss_host = *p->loss_h;
  }
  return MPPI_OK;
}

int mppi_proposal_train_step(mppi_proposal* q, const float* inputs_d, const float* targets_d, int32_t B, float lr,
                             double* loss_host, void* stream)
{
  if (!q || !inputs_d || !targets_d) return fail(MPPI_E_INVALID, "null argument");
  if (B < 1) return fail(MPPI_E_INVALID, "proposal: batch must be >= 1");
  if (!(lr > 0.0f)) return fail(MPPI_E_INVALID, "proposal: learning rate must be > 0");
  ProposalNet* p = &q->net;
  MPPI_CHECK(prop_reserve(p, B));
  const hipStream_t s = (hipStream_t)stream;
  MPPI_CHECK(prop_loss_grad(p, inputs_d, targets_d, B, p->grad, s));
  p->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)p->step), c2 = 1.0f - std::pow(b2, (float)p->step);
  const int P = (int)p->d.total;
  adam_kernel<<<(P + 255) / 256, 256, 0, s>>>(p->theta, p->adam_m, p->adam_v, p->grad, P, lr, b1, b2, 1e-8f, c1, c2);
  MPPI_HIP_CHECK(hipGetLastError());
  MPPI_CHECK(prop_mirror(p, s));
  if (loss_host) {
    MPPI_HIP_CHECK(hipMemcpyAsync(p->loss_h, p->loss, sizeof(double), hipMemcpyDeviceToHost, s));
    MPPI_HIP_CHECK(hipStreamSynchronize(s));
    *loss_host = *p->loss_h;
  }
  return MPPI_OK;
}

}  // extern "C"
