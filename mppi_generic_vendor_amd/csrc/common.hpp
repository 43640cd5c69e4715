// Shared definitions of the MI355X MPPI engine: error plumbing, the counter-based normal generator and wavefront
// reductions. Everything here is compiled by hipcc for gfx950 (device) and x86-64 (host).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>
#include <cmath>
#include <string>

#include "../../include/mppi_amd.h"

#define MPPI_HD __host__ __device__ __forceinline__

// LDS available to one workgroup on gfx950.
constexpr size_t LDS_LIMIT = 160 * 1024;

namespace mppi_amd {

// ------------------------------------------------------------------------------------------------ error handling
void set_last_error(const std::string& msg);
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// Raises kernel `fn`'s dynamic-LDS limit to LDS_LIMIT on the current device, once per (device, kernel); thread-safe.
int allow_max_dynamic_lds(const void* fn);

#define MPPI_HIP_CHECK(call)                              \
  do {                                                    \
    hipError_t _e = (call);                               \
    if (_e != hipSuccess) return ::mppi_amd::hip_fail(_e, #call); \
  } while (0)

#define MPPI_CHECK(status_expr)       \
  do {                                \
    int _s = (status_expr);           \
    if (_s != MPPI_OK) return _s;     \
  } while (0)

// ------------------------------------------------------------------------------------------- Philox4x32-10 + BM
// Counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11). Normal i of the stream (seed, offset) comes from
// Philox block i/4 with counter (i/4 lo, i/4 hi, offset lo, offset hi), key (seed lo, seed hi), and Box-Muller on
// the pairs (x0, x1) and (x2, x3). Any launch shape yields the same numbers.
struct Philox4 {
  uint32_t v[4];
};

MPPI_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  Philox4 o;
  o.v[0] = c0;
  o.v[1] = c1;
  o.v[2] = c2;
  o.v[3] = c3;
  return o;
}

// Uniform in (0, 1]: fma((float)x, 2^-32, 2^-33) -- one rounding, identical on every IEEE platform.
MPPI_HD float u32_to_unit(uint32_t x) { return fmaf((float)x, 2.3283064365386963e-10f, 1.1641532182693481e-10f); }

struct Normal4 {
  float z[4];
};

// Box-Muller with the hardware transcendentals: r = sqrt(-2 ln 2 * log2(u1)) (v_log_f32, v_sqrt_f32) and
// (cos, sin)(2 pi u2) from v_cos_f32 / v_sin_f32, which take their argument in revolutions (u2 in (0, 1]).
__device__ __forceinline__ Normal4 normals4(uint64_t seed, uint64_t offset, uint64_t block)
{
  const Philox4 o = philox4x32_10((uint32_t)block, (uint32_t)(block >> 32), (uint32_t)offset,
                                  (uint32_t)(offset >> 32), (uint32_t)seed, (uint32_t)(seed >> 32));
  Normal4 n;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const float u1 = u32_to_unit(o.v[2 * p]);
    const float u2 = u32_to_unit(o.v[2 * p + 1]);
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    n.z[2 * p] = r * __builtin_amdgcn_cosf(u2);
    n.z[2 * p + 1] = r * __builtin_amdgcn_sinf(u2);
  }
  return n;
}

// --------------------------------------------------------------------------------------------- wave reductions
// Butterfly exchanges over a full wavefront: stage 0..3 pair lanes inside each row of 16 by DPP (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror), stage 4 pairs rows r and r ^ 1 and stage 5 the two halves by
// the gfx950 row-swap permutes (v_permlane16_swap / v_permlane32_swap of a register with itself: result 0 holds row
// 2 j in rows 2 j, 2 j + 1 and result 1 row 2 j + 1; likewise for halves). After stages 0 .. s every group of 2^(s+1)
// lanes has exchanged with every other, and both lanes of a pair see the same two values, so a commutative combine
// leaves the same result in every lane. Every lane of the wavefront must be active.
template <int STAGE>
__device__ __forceinline__ uint32_t lane_partner(uint32_t v)
{
  if constexpr (STAGE == 0) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
  else if constexpr (STAGE == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
  else if constexpr (STAGE == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
  else if constexpr (STAGE == 3) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
  else if constexpr (STAGE == 4) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return ((threadIdx.x >> 4) & 1) ? r[0] : r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32) ? r[0] : r[1];
  }
}

template <int STAGE>
__device__ __forceinline__ float partner(float v)
{
  return __uint_as_float(lane_partner<STAGE>(__float_as_uint(v)));
}
template <int STAGE>
__device__ __forceinline__ int partner(int v)
{
  return (int)lane_partner<STAGE>((uint32_t)v);
}
template <int STAGE>
__device__ __forceinline__ double partner(double v)
{
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint64_t lo = lane_partner<STAGE>((uint32_t)b), hi = lane_partner<STAGE>((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(lo | (hi << 32)));
}

// Sum over the 64 lanes, the same value in every lane.
template <class Tp>
__device__ __forceinline__ Tp wave_sum(Tp v)
{
  v += partner<0>(v);
  v += partner<1>(v);
  v += partner<2>(v);
  v += partner<3>(v);
  v += partner<4>(v);
  v += partner<5>(v);
  return v;
}

// (value, index) minimum with the lower index winning ties, i.e. the first strict minimum of a scan in index order.
__device__ __forceinline__ void argmin_combine(float& v, int& i, float v2, int i2)
{
  if (v2 < v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

template <int STAGE>
__device__ __forceinline__ void argmin_stage(float& v, int& i)
{
  const float v2 = partner<STAGE>(v);
  const int i2 = partner<STAGE>(i);
  argmin_combine(v, i, v2, i2);
}

// Minimum (first strict one) over the 64 lanes, the same in every lane.
__device__ __forceinline__ void wave_argmin(float& v, int& i)
{
  argmin_stage<0>(v, i);
  argmin_stage<1>(v, i);
  argmin_stage<2>(v, i);
  argmin_stage<3>(v, i);
  argmin_stage<4>(v, i);
  argmin_stage<5>(v, i);
}

// A cost that is NaN or +inf carries no weight and never becomes the baseline.
__device__ __forceinline__ float sanitize_cost(float c) { return (c <= FLT_MAX) ? c : INFINITY; }

// Pure-noise start index: first k with (float)k >= (1 - p) * K in FP32 (sampling_distributions/gaussian/gaussian.cu:96).
int pure_noise_start(int K, float pct);

// ----------------------------------------------------------------------------------------- host completion signal
// The end of a control call as seen by a host that polls a word of mapped host memory instead of blocking on an event:
// the call's last kernel, once the results of all its workgroups are visible to the host, advances a device epoch and
// stores it to `flag` (the device-side alias of that word). `count` (zeroed device word) counts the workgroups that
// are done, `epoch` (zeroed device word) is the number of signals given. Null `flag`: no signal.
struct HostSignal {
  uint32_t* count = nullptr;
  uint32_t* epoch = nullptr;
  uint32_t* flag = nullptr;
};

// Called by every thread of every workgroup at the end of a kernel whose results the host reads. Each thread's own
// stores (results in mapped host memory among them) are made visible at system scope, then the workgroup counts
// itself done; true, in every thread of the workgroup, for the last workgroup to arrive, which then also sees every
// other workgroup's stores.
__device__ __forceinline__ bool last_block_arrive(const HostSignal& s)
{
  __shared__ uint32_t last;
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x + threadIdx.y + threadIdx.z == 0) {
    const uint32_t blocks = gridDim.x * gridDim.y * gridDim.z;
    const uint32_t before = __hip_atomic_fetch_add(s.count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (before + 1u == blocks) ? 1u : 0u;
  }
  __syncthreads();
  const bool is_last = last != 0u;
  if (is_last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return is_last;
}

// Called by every thread of the last workgroup (last_block_arrive) once its own stores are done: resets the count,
// advances the epoch and publishes it with a system-scope release store.
__device__ __forceinline__ void signal_publish(const HostSignal& s)
{
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x + threadIdx.y + threadIdx.z == 0) {
    *s.count = 0u;
    const uint32_t e = *s.epoch + 1u;
    *s.epoch = e;
    __threadfence_system();
    __hip_atomic_store(s.flag, e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The whole signal at the very end of a kernel (every thread of every workgroup calls it). Null flag: nothing.
__device__ __forceinline__ void signal_host_done(const HostSignal& s)
{
  if (s.flag == nullptr) return;
  if (last_block_arrive(s)) signal_publish(s);
}

}  // namespace mppi_amd
