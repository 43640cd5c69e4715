// Residual MLP [x(7), u(2)] -> 64 -> 64 -> 4 (tanh, tanh, linear) on BF16 matrix cores with FP32 accumulation, for
// the bicycle + NN model (the FNN forward of utils/nn_helpers/fnn_helper.cu:419-494, as an MFMA chain).
//
// One wavefront evaluates the network for 16 samples per call. The weights are the A operand of
// v_mfma_f32_16x16x32_bf16 (rows = output units) and stay in registers for the whole rollout; the activations are
// the B operand (columns = samples). Lane l = 16 g + n holds, for sample n, the accumulator rows 4 g + i (i < 4) of
// every 16-row tile, and the B-operand K slots (g, j < 8) -- the same (g, j) slots in the A operand of the lanes
// 16 g + m. A layer's reduction index may therefore be permuted freely as long as A and B agree: the next layer's
// K slot (ks, g, j) is taken to be hidden unit 16 (2 ks + j / 4) + 4 g + j % 4, which is exactly accumulator element
// j % 4 of tile 2 ks + j / 4 in the same lane. Layer outputs feed the next layer without leaving the registers.
// Biases enter as the accumulator input (FP32). The output layer is laid out by state instead of by output: two
// 16-row tiles whose row 4 g + i of tile q carries the network output added to the derivative of state 4 q + i (zero
// weights for states that receive none), so every lane group receives, in place, the increments of all states of
// its sample.
#pragma once

#include "models.hpp"

namespace mppi_amd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FNN parameter offsets (utils/nn_helpers/fnn_helper.cu:130-205: W_l row-major [out][in], then b_l).
constexpr int NN_W1 = 0;
constexpr int NN_B1 = NN_W1 + NN_IN * NN_H1;
constexpr int NN_W2 = NN_B1 + NN_H1;
constexpr int NN_B2 = NN_W2 + NN_H1 * NN_H2;
constexpr int NN_W3 = NN_B2 + NN_H2;
constexpr int NN_B3 = NN_W3 + NN_H2 * NN_OUT;
static_assert(NN_B3 + NN_OUT == NN_PARAMS, "FNN layout");
static_assert(NN_H1 == 64 && NN_H2 == 64 && NN_IN <= 16 && NN_OUT == 4, "MFMA tiling assumes 9-64-64-4");

// tanh(v) = 1 - 2 / (1 + 2^(2 v log2 e)) with the hardware exp2 and reciprocal; saturates to +-1 (2^+inf -> rcp 0,
// 2^-inf -> 0), absolute error a few 1e-7, far below the BF16 rounding that follows.
__device__ __forceinline__ float tanh_fast(float v)
{
  const float e = __builtin_amdgcn_exp2f(v * 2.8853900817779268f);
  return fmaf(-2.0f, __builtin_amdgcn_rcpf(1.0f + e), 1.0f);
}

__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x4& c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x2 cvt_bf16x2(float a, float b) { return __builtin_convertvector(f32x2{a, b}, bf16x2); }

__device__ __forceinline__ bf16x8 cat_bf16(bf16x2 a, bf16x2 b, bf16x2 c, bf16x2 d)
{
  const auto ab = __builtin_shufflevector(a, b, 0, 1, 2, 3);
  const auto cd = __builtin_shufflevector(c, d, 0, 1, 2, 3);
  return __builtin_shufflevector(ab, cd, 0, 1, 2, 3, 4, 5, 6, 7);
}

// One accumulator tile's activations: tanh, then round to BF16.
__device__ __forceinline__ bf16x4 act4(const f32x4& d)
{
  bf16x4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = (__bf16)tanh_fast(d[i]);
  return r;
}

__device__ __forceinline__ bf16x8 cat_bf16(bf16x4 a, bf16x4 b)
{
  return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

// B-operand slots of the next layer from two accumulator tiles.
__device__ __forceinline__ bf16x8 act_pack(const f32x4& lo, const f32x4& hi) { return cat_bf16(act4(lo), act4(hi)); }

// Hidden unit carried by K slot (ks, g, j) of layers 2 and 3.
__device__ __forceinline__ int nn_hidden_slot(int ks, int g, int j) { return 16 * (2 * ks + j / 4) + 4 * g + (j & 3); }

struct NnFrags {
  bf16x8 a1[4];     // layer 1, tile mt: row 16 mt + m, K slot (g, j) = input 8 g + j (zero beyond NN_IN)
  bf16x8 a2[4][2];  // layer 2, tile mt, k-step ks: row 16 mt + m, K slot = hidden nn_hidden_slot(ks, g, j)
  bf16x8 a3[2][2];  // layer 3, tile q, k-step ks: row m -> the output feeding state 4 q + m % 4
  f32x4 b1[4], b2[4], b3[2];
};

// Lane `lane`'s fragments of hidden tile mt (units 16 mt .. 16 mt + 15) of layers 1 and 2, from the FP32 parameter
// vector `th` (LDS or global).
__device__ __forceinline__ void nn_load_hidden_tile(const float* th, int lane, int mt, bf16x8& a1, bf16x8 (&a2)[2],
                                                    f32x4& b1, f32x4& b2)
{
  const int m = lane & 15, g = lane >> 4;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int f = 8 * g + j;
    a1[j] = (__bf16)(f < NN_IN ? th[NN_W1 + (16 * mt + m) * NN_IN + f] : 0.0f);
  }
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int j = 0; j < 8; ++j) a2[ks][j] = (__bf16)th[NN_W2 + (16 * mt + m) * NN_H1 + nn_hidden_slot(ks, g, j)];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    b1[i] = th[NN_B1 + 16 * mt + 4 * g + i];
    b2[i] = th[NN_B2 + 16 * mt + 4 * g + i];
  }
}

// Lane `lane`'s fragments of the output layer. `out_states` are the distinct states (< 8) that the NN_OUT outputs
// are added to.
__device__ __forceinline__ void nn_load_out_tiles(const float* th, const int* out_states, int lane,
                                                  bf16x8 (&a3)[2][2], f32x4 (&b3)[2])
{
  const int m = lane & 15, g = lane >> 4;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    // output feeding the state of this lane's row (m % 4) and of its accumulator rows (4 g + i)
    int o_row = -1;
#pragma unroll
    for (int o = 0; o < NN_OUT; ++o)
      if (out_states[o] == 4 * q + (m & 3)) o_row = o;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        a3[q][ks][j] = (__bf16)(o_row >= 0 ? th[NN_W3 + o_row * NN_H2 + nn_hidden_slot(ks, g, j)] : 0.0f);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float b = 0.0f;
#pragma unroll
      for (int o = 0; o < NN_OUT; ++o)
        if (out_states[o] == 4 * q + i) b = th[NN_B3 + o];
      b3[q][i] = b;
    }
  }
}

// Builds lane `lane`'s fragments of the whole network.
__device__ __forceinline__ void nn_load_frags(const float* th, const int* out_states, int lane, NnFrags& F)
{
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) nn_load_hidden_tile(th, lane, mt, F.a1[mt], F.a2[mt], F.b1[mt], F.b2[mt]);
  nn_load_out_tiles(th, out_states, lane, F.a3, F.b3);
}

// Network increments of the derivatives of states 0..3 (.x) and 4..7 (.y) for the lane's sample (every lane group of
// the sample gets all of them).
struct NnOut {
  f32x4 lo, hi;
};

// Layer-1 B operand: inputs 0..7 (x0..x6, u0) in the K slots of lane group 0, input 8 (u1) in slot 0 of group 1.
// The other groups repeat group 0's values: their layer-1 weights are zero, so (for finite inputs) they add exact
// zeros. x: the 7 states, u: the 2 controls, g = lane / 16.
__device__ __forceinline__ bf16x8 nn_input(const float* x, const float* u, int g)
{
  return cat_bf16(cvt_bf16x2(g == 1 ? u[1] : x[0], x[1]), cvt_bf16x2(x[2], x[3]), cvt_bf16x2(x[4], x[5]),
                  cvt_bf16x2(x[6], u[0]));
}

// Output layer from the two layer-2 B operands.
__device__ __forceinline__ NnOut nn_output(const bf16x8 (&a3)[2][2], const f32x4 (&b3)[2], const bf16x8& h2a,
                                           const bf16x8& h2b)
{
  NnOut r;
  r.lo = mfma_bf16(a3[0][1], h2b, mfma_bf16(a3[0][0], h2a, b3[0]));
  r.hi = mfma_bf16(a3[1][1], h2b, mfma_bf16(a3[1][0], h2a, b3[1]));
  return r;
}

__device__ __forceinline__ NnOut nn_eval_bf16(const NnFrags& F, const float* x, const float* u, int g)
{
  const bf16x8 bin = nn_input(x, u, g);
  f32x4 d1[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) d1[mt] = mfma_bf16(F.a1[mt], bin, F.b1[mt]);
  const bf16x8 h1a = act_pack(d1[0], d1[1]);
  const bf16x8 h1b = act_pack(d1[2], d1[3]);
  f32x4 d2[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) d2[mt] = mfma_bf16(F.a2[mt][1], h1b, mfma_bf16(F.a2[mt][0], h1a, F.b2[mt]));
  const bf16x8 h2a = act_pack(d2[0], d2[1]);
  const bf16x8 h2b = act_pack(d2[2], d2[3]);
  return nn_output(F.a3, F.b3, h2a, h2b);
}

}  // namespace mppi_amd

namespace mppi_amd {

// Kernel-template id of the bicycle + NN model with the network on BF16 matrix cores (MPPI_MODEL_RACER_DUBINS_NN
// with nn_precision == MPPI_NN_BF16); the launchers map the public id to it.
constexpr int MODEL_RACER_DUBINS_NN_BF16 = 0x100 | MPPI_MODEL_RACER_DUBINS_NN;

template <>
struct Model<MODEL_RACER_DUBINS_NN_BF16> : Bicycle {};

}  // namespace mppi_amd
