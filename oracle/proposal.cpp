/* CPU oracle of the learned proposal network -- TEST INFRASTRUCTURE ONLY (the checker of the HIP kernels).
 *
 * The proposal MLP has no counterpart in the reference; it is specified by the task: inputs = the state (S values)
 * and a 16-knot downsampling of the previous control sequence (16 C values), two hidden layers of H = 128 GELU units,
 * output 2 T C values: the mean shift in units of the base sampling std (first T C) and log(sigma_q / sigma_base)
 * (last T C). Parameters follow the FNN layout used for the dynamics network (utils/nn_helpers/fnn_helper.cu:130-205:
 * per layer W[out][in] row-major, then b[out]).
 *
 * Loss per sample (Gaussian negative log-likelihood of the target residual r = (u* - u_prev) / sigma_base, with the
 * log-std clamped to [S_LO, S_HI]): (1 / TC) sum_j [ (r_j - y_j)^2 exp(-2 s_j) / 2 + s_j ], averaged over the batch.
 *
 * Everything here is FP64. With `bf16`, both operands of every matrix product are rounded to BF16 first (the
 * arithmetic of the matrix-core kernels); biases, activations and their derivatives stay in FP64.
 * Pins (tests/test_oracle_pins.py): the backward pass against central finite differences of the FP64 loss, GELU
 * and its derivative against closed forms, Adam's first step, the input downsampling against hand-built vectors.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int H = 128;
constexpr double S_LO = -0.693147, S_HI = 1.0;

float bf16_round_p(float f)
{
  uint32_t b;
  std::memcpy(&b, &f, 4);
  if ((b & 0x7f800000u) == 0x7f800000u) return f;
  b += 0x7fffu + ((b >> 16) & 1u);
  b &= 0xffff0000u;
  std::memcpy(&f, &b, 4);
  return f;
}

inline double rd(double v, bool bf16) { return bf16 ? (double)bf16_round_p((float)v) : v; }

inline double gelu(double z) { return 0.5 * z * (1.0 + std::erf(z / std::sqrt(2.0))); }
inline double dgelu(double z)
{
  return 0.5 * (1.0 + std::erf(z / std::sqrt(2.0))) + z * std::exp(-0.5 * z * z) / std::sqrt(2.0 * M_PI);
}

struct Dims {
  int din, dout;
  size_t w1, b1, w2, b2, w3, b3, total;
};

Dims dims(int S, int C, int T)
{
  Dims d;
  d.din = S + 16 * C;
  d.dout = 2 * T * C;
  d.w1 = 0;
  d.b1 = d.w1 + (size_t)H * d.din;
  d.w2 = d.b1 + H;
  d.b2 = d.w2 + (size_t)H * H;
  d.w3 = d.b2 + H;
  d.b3 = d.w3 + (size_t)d.dout * H;
  d.total = d.b3 + d.dout;
  return d;
}

// out[m][n] = sum_k A[m][k] * W[n][k] + b[n]   (A: M x K row-major, W: N x K row-major)
void dense(int M, int N, int K, const double* A, const float* W, const float* b, bool bf16, double* out)
{
  std::vector<double> Wr((size_t)N * K);
  for (size_t i = 0; i < Wr.size(); ++i) Wr[i] = rd(W[i], bf16);
  for (int m = 0; m < M; ++m) {
    std::vector<double> a(K);
    for (int k = 0; k < K; ++k) a[k] = rd(A[(size_t)m * K + k], bf16);
    for (int n = 0; n < N; ++n) {
      double s = 0.0;
      const double* w = &Wr[(size_t)n * K];
      for (int k = 0; k < K; ++k) s += a[k] * w[k];
      out[(size_t)m * N + n] = s + (b ? (double)b[n] : 0.0);
    }
  }
}

struct Fwd {
  std::vector<double> x, z1, h1, z2, h2, y;
};

Fwd forward(const Dims& d, const float* th, int B, const float* X, bool bf16)
{
  Fwd f;
  f.x.assign(X, X + (size_t)B * d.din);
  f.z1.resize((size_t)B * H);
  f.h1.resize((size_t)B * H);
  f.z2.resize((size_t)B * H);
  f.h2.resize((size_t)B * H);
  f.y.resize((size_t)B * d.dout);
  dense(B, H, d.din, f.x.data(), th + d.w1, th + d.b1, bf16, f.z1.data());
  for (size_t i = 0; i < f.z1.size(); ++i) f.h1[i] = gelu(f.z1[i]);
  dense(B, H, H, f.h1.data(), th + d.w2, th + d.b2, bf16, f.z2.data());
  for (size_t i = 0; i < f.z2.size(); ++i) f.h2[i] = gelu(f.z2[i]);
  dense(B, d.dout, H, f.h2.data(), th + d.w3, th + d.b3, bf16, f.y.data());
  return f;
}

}  // namespace

extern "C" {

int oracle_proposal_num_params(int S, int C, int T) { return (int)dims(S, C, T).total; }

double oracle_gelu(double z) { return gelu(z); }
double oracle_dgelu(double z) { return dgelu(z); }

/* Network input: (state, prev[t_j][c] for the 16 knots t_j = j (T - 1) / 15 in integer arithmetic), then
 * (v - shift) * scale per feature. */
void oracle_proposal_input(int S, int C, int T, const float* x, const float* prev, const float* shift,
                           const float* scale, float* out)
{
  int o = 0;
  for (int i = 0; i < S; ++i, ++o) out[o] = (x[i] - shift[o]) * scale[o];
  for (int j = 0; j < 16; ++j) {
    const int t = j * (T - 1) / 15;
    for (int c = 0; c < C; ++c, ++o) out[o] = (prev[t * C + c] - shift[o]) * scale[o];
  }
}

void oracle_proposal_forward(int S, int C, int T, const float* theta, int B, const float* X, int bf16, double* Y)
{
  const Dims d = dims(S, C, T);
  const Fwd f = forward(d, theta, B, X, bf16 != 0);
  std::copy(f.y.begin(), f.y.end(), Y);
}

/* Mean loss over the batch; grad (FP64, num_params) gets its gradient. R: targets [B][T C]. */
double oracle_proposal_loss_grad(int S, int C, int T, const float* theta, int B, const float* X, const float* R,
                                 int bf16_i, double* grad)
{
  const bool bf16 = bf16_i != 0;
  const Dims d = dims(S, C, T);
  const int TC = T * C;
  const Fwd f = forward(d, theta, B, X, bf16);
  std::vector<double> dy((size_t)B * d.dout);
  double loss = 0.0;
  const double inv = 1.0 / ((double)TC * B);
  for (int b = 0; b < B; ++b)
    for (int j = 0; j < TC; ++j) {
      const double y = f.y[(size_t)b * d.dout + j];
      const double s = f.y[(size_t)b * d.dout + TC + j];
      const double sc = std::min(std::max(s, S_LO), S_HI);
      const double e = (double)R[(size_t)b * TC + j] - y;
      const double iv = std::exp(-2.0 * sc);
      loss += (0.5 * e * e * iv + sc) * inv;
      dy[(size_t)b * d.dout + j] = -e * iv * inv;
      dy[(size_t)b * d.dout + TC + j] = (s > S_LO && s < S_HI) ? (1.0 - e * e * iv) * inv : 0.0;
    }
  std::fill(grad, grad + d.total, 0.0);
  // layer 3: Y = H2 W3^T + b3
  for (int n = 0; n < d.dout; ++n)
    for (int k = 0; k < H; ++k) {
      double s = 0.0;
      for (int b = 0; b < B; ++b) s += rd(dy[(size_t)b * d.dout + n], bf16) * rd(f.h2[(size_t)b * H + k], bf16);
      grad[d.w3 + (size_t)n * H + k] = s;
    }
  for (int n = 0; n < d.dout; ++n) {
    double s = 0.0;
    for (int b = 0; b < B; ++b) s += dy[(size_t)b * d.dout + n];
    grad[d.b3 + n] = s;
  }
  std::vector<double> dz2((size_t)B * H), dz1((size_t)B * H);
  for (int b = 0; b < B; ++b)
    for (int k = 0; k < H; ++k) {
      double s = 0.0;
      for (int n = 0; n < d.dout; ++n) s += rd(dy[(size_t)b * d.dout + n], bf16) * rd(theta[d.w3 + (size_t)n * H + k], bf16);
      dz2[(size_t)b * H + k] = s * dgelu(f.z2[(size_t)b * H + k]);
    }
  // layer 2
  for (int n = 0; n < H; ++n)
    for (int k = 0; k < H; ++k) {
      double s = 0.0;
      for (int b = 0; b < B; ++b) s += rd(dz2[(size_t)b * H + n], bf16) * rd(f.h1[(size_t)b * H + k], bf16);
      grad[d.w2 + (size_t)n * H + k] = s;
    }
  for (int n = 0; n < H; ++n) {
    double s = 0.0;
    for (int b = 0; b < B; ++b) s += dz2[(size_t)b * H + n];
    grad[d.b2 + n] = s;
  }
  for (int b = 0; b < B; ++b)
    for (int k = 0; k < H; ++k) {
      double s = 0.0;
      for (int n = 0; n < H; ++n) s += rd(dz2[(size_t)b * H + n], bf16) * rd(theta[d.w2 + (size_t)n * H + k], bf16);
      dz1[(size_t)b * H + k] = s * dgelu(f.z1[(size_t)b * H + k]);
    }
  // layer 1
  for (int n = 0; n < H; ++n)
    for (int k = 0; k < d.din; ++k) {
      double s = 0.0;
      for (int b = 0; b < B; ++b) s += rd(dz1[(size_t)b * H + n], bf16) * rd(f.x[(size_t)b * d.din + k], bf16);
      grad[d.w1 + (size_t)n * d.din + k] = s;
    }
  for (int n = 0; n < H; ++n) {
    double s = 0.0;
    for (int b = 0; b < B; ++b) s += dz1[(size_t)b * H + n];
    grad[d.b1 + n] = s;
  }
  return loss;
}

/* Importance-weight correction per sample (the caller leaves it off the warm-start sample 0): the mean-shift term
 * lambda sum_{t,c} ((u - mu)^2 - (u - mu - dmu)^2) / (2 sp_c^2) = lambda (log N(u; mu + dmu, sp^2) - log N(u; mu, sp^2)).
 * `sig` (the proposal's spread) does not enter. */
void oracle_is_cost(const float* u, const float* mean, const float* dmu, const float* sig, const float* sig_p, int K,
                    int T, int C, double lambda, double* out)
{
  (void)sig;
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int t = 0; t < T; ++t)
      for (int c = 0; c < C; ++c) {
        const int j = t * C + c;
        const double d0 = (double)u[((size_t)k * T + t) * C + c] - mean[j];
        const double d1 = d0 - dmu[j];
        const double sp = sig_p[c];
        s += (d0 * d0 - d1 * d1) / (2.0 * sp * sp);
      }
    out[k] = lambda * s;
  }
}

/* Adam (Kingma & Ba 2015) with bias correction; `step` counts from 1. */
void oracle_adam(float* theta, float* m, float* v, const float* grad, int P, float lr, float b1, float b2, float eps,
                 int step)
{
  const double c1 = 1.0 - std::pow((double)b1, step), c2 = 1.0 - std::pow((double)b2, step);
  for (int i = 0; i < P; ++i) {
    m[i] = b1 * m[i] + (1.0f - b1) * grad[i];
    v[i] = b2 * v[i] + (1.0f - b2) * grad[i] * grad[i];
    const double mh = m[i] / c1, vh = v[i] / c2;
    theta[i] = (float)(theta[i] - lr * mh / (std::sqrt(vh) + eps));
  }
}

} /* extern "C" */
