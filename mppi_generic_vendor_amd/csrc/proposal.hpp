// Learned proposal network: S + 16 C inputs (state, 16-knot downsampling of the previous control sequence), two
// 128-unit GELU layers, 2 T C outputs (mean shift in units of the base sampling std, then log(sigma_q / sigma_base)).
// Forward, backward and Adam run on the GPU through a BF16 matrix-core GEMM with FP32 accumulation.
#pragma once

#include "common.hpp"
#include "kernels.hpp"

namespace mppi_amd {

constexpr int PROP_H = 128;
constexpr int PROP_KNOTS = 16;
constexpr float PROP_S_LO = -0.693147f, PROP_S_HI = 1.0f;  // sigma_q in [0.5, 2.7] sigma_base

struct PropDims {
  int S, C, T, din, dout;
  size_t w1, b1, w2, b2, w3, b3, total;
};

PropDims prop_dims(int S, int C, int T);

// Inputs that the inference kernel also copies on for the kernels after it (workgroup 0): x0 [S] to x0_dst, the
// nominal sequence [T][C] to nominal_dst and *iter_src to *iter_dst, so that a caller whose x0 and nominal live in
// mapped host memory needs no separate upload.
struct InferStaging {
  float* x0_dst;
  float* nominal_dst;
  const IterArgs* iter_src;
  IterArgs* iter_dst;
};

// Inference for one state: writes the absolute mean shift dmu[T][C] = sigma_base[c] * y and the proposal std
// sig[T][C] = sigma_base[c] * exp(clamp(s, PROP_S_LO, PROP_S_HI)). x0 [S], nominal [T][C]: device-readable arrays.
struct ProposalNet;
ProposalNet* proposal_net_of(mppi_proposal* q);
PropDims proposal_dims(const ProposalNet* p);
int proposal_infer(ProposalNet* p, const float* x0, const float* nominal, const float* sigma_base, float* dmu,
                   float* sig, hipStream_t s, const InferStaging* staging = nullptr);

}  // namespace mppi_amd
