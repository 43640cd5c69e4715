// Staging of a workgroup's control tile between HBM and LDS. The tile of rows k0 .. k0+nrows-1 of the [K][T*C]
// sample buffer is one contiguous span; it is moved with 16-byte accesses issued in batches of TILE_BATCH per lane
// and scattered into LDS rows padded to ROW = T*C + 1 floats (conflict-free column access by one lane per row).
#pragma once

#include "common.hpp"

namespace mppi_amd {

constexpr int TILE_BATCH = 16;

struct TileIdentity {
  __device__ float operator()(int, float v) const { return v; }
};

// f(col, v) is applied to every element on its way into LDS (col = t C + c within the row).
template <int NT, typename F = TileIdentity>
__device__ __forceinline__ void tile_load(const float* __restrict__ src, float* __restrict__ tile, int nrows, int TC,
                                          int ROW, int tid, F f = F())
{
  const int n = nrows * TC;
  if ((((uintptr_t)src) & 15) == 0) {
    const float4* s4 = reinterpret_cast<const float4*>(src);
    const int n4 = n >> 2;
    for (int base = 0; base < n4; base += NT * TILE_BATCH) {
      // unconditional loads (index clamped in bounds) so that the batch issues back to back
      float4 v[TILE_BATCH];
#pragma unroll
      for (int i = 0; i < TILE_BATCH; ++i) v[i] = s4[min(base + i * NT + tid, n4 - 1)];
#pragma unroll
      for (int i = 0; i < TILE_BATCH; ++i) {
        const int idx = base + i * NT + tid;
        if (idx < n4) {
          int e = idx * 4;
          int r = e / TC;
          int col = e - r * TC;
          const float vv[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            tile[r * ROW + col] = f(col, vv[q]);
            if (++col == TC) {
              col = 0;
              ++r;
            }
          }
        }
      }
    }
    for (int e = (n4 << 2) + tid; e < n; e += NT) {
      const int r = e / TC;
      tile[r * ROW + (e - r * TC)] = f(e - r * TC, src[e]);
    }
  } else {
    for (int e = tid; e < n; e += NT) {
      const int r = e / TC;
      tile[r * ROW + (e - r * TC)] = f(e - r * TC, src[e]);
    }
  }
}

template <int NT>
__device__ __forceinline__ void tile_store(float* __restrict__ dst, const float* __restrict__ tile, int nrows, int TC,
                                           int ROW, int tid)
{
  const int n = nrows * TC;
  if ((((uintptr_t)dst) & 15) == 0) {
    float4* d4 = reinterpret_cast<float4*>(dst);
    const int n4 = n >> 2;
    for (int idx = tid; idx < n4; idx += NT) {
      int e = idx * 4;
      int r = e / TC;
      int col = e - r * TC;
      float vv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        vv[q] = tile[r * ROW + col];
        if (++col == TC) {
          col = 0;
          ++r;
        }
      }
      d4[idx] = make_float4(vv[0], vv[1], vv[2], vv[3]);
    }
    for (int e = (n4 << 2) + tid; e < n; e += NT) {
      const int r = e / TC;
      dst[e] = tile[r * ROW + (e - r * TC)];
    }
  } else {
    for (int e = tid; e < n; e += NT) {
      const int r = e / TC;
      dst[e] = tile[r * ROW + (e - r * TC)];
    }
  }
}

}  // namespace mppi_amd
