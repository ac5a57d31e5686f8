#pragma once

// fp32 distance of one row against the query by the four lanes of a quad, in
// the engine's fixed summation order (that of the reference's AVX-512 kernels,
// src/kernels_avx512.cpp:10-58): bit-identical to host/distance.hpp.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace hnswamd {
namespace gpu {

__device__ __forceinline__ void acc4(float (&a)[4], const float4 y, const float4 x, bool l2) {
  if (l2) {
    const float d0 = y.x - x.x, d1 = y.y - x.y, d2 = y.z - x.z, d3 = y.w - x.w;
    a[0] = __builtin_fmaf(d0, d0, a[0]);
    a[1] = __builtin_fmaf(d1, d1, a[1]);
    a[2] = __builtin_fmaf(d2, d2, a[2]);
    a[3] = __builtin_fmaf(d3, d3, a[3]);
  } else {
    a[0] = __builtin_fmaf(y.x, x.x, a[0]);
    a[1] = __builtin_fmaf(y.y, x.y, a[1]);
    a[2] = __builtin_fmaf(y.z, x.z, a[2]);
    a[3] = __builtin_fmaf(y.w, x.w, a[3]);
  }
}

/// Raw l2_sq or dot of one padded row against the query, computed by the four
/// lanes of a quad. Lane g of the quad holds positions 4g..4g+3 of every
/// 16-float chunk, accumulator q = chunk index mod 4 inside each 64-float block
/// (chunks after the last full block go to accumulator 0). The result is
/// identical in all four lanes. PC = number of whole 16-float chunks when known
/// at compile time (else pc_rt). `partial`: one more zero-padded chunk holds the
/// last dim % 16 elements; it goes to accumulator 1, as a masked tail does in
/// the reference's kernels (src/kernels_avx512.cpp:32-36, :53-56).
template <int PC>
__device__ __forceinline__ float quad_raw(bool l2, const float* __restrict__ row, const float* q, uint32_t pc_rt,
                                          bool partial, bool active) {
  const uint32_t pc = PC > 0 ? static_cast<uint32_t>(PC) : pc_rt;
  const uint32_t g4 = static_cast<uint32_t>(lane_id() & 3) * 4;
  float a[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int e = 0; e < 4; ++e) a[i][e] = 0.0f;
  }
  if (active) {
    const uint32_t blocks = pc >> 2, rest = pc & 3;
#pragma unroll 2
    for (uint32_t t = 0; t < blocks; ++t) {
      const uint32_t o = 64 * t + g4;
      const float4 x0 = *reinterpret_cast<const float4*>(row + o);
      const float4 x1 = *reinterpret_cast<const float4*>(row + o + 16);
      const float4 x2 = *reinterpret_cast<const float4*>(row + o + 32);
      const float4 x3 = *reinterpret_cast<const float4*>(row + o + 48);
      acc4(a[0], *reinterpret_cast<const float4*>(q + o), x0, l2);
      acc4(a[1], *reinterpret_cast<const float4*>(q + o + 16), x1, l2);
      acc4(a[2], *reinterpret_cast<const float4*>(q + o + 32), x2, l2);
      acc4(a[3], *reinterpret_cast<const float4*>(q + o + 48), x3, l2);
    }
    for (uint32_t c = 0; c < rest; ++c) {
      const uint32_t o = 64 * blocks + 16 * c + g4;
      acc4(a[0], *reinterpret_cast<const float4*>(q + o), *reinterpret_cast<const float4*>(row + o), l2);
    }
    if (partial) {
      const uint32_t o = 64 * blocks + 16 * rest + g4;
      acc4(a[1], *reinterpret_cast<const float4*>(q + o), *reinterpret_cast<const float4*>(row + o), l2);
    }
  }
  float v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = (a[0][e] + a[1][e]) + (a[2][e] + a[3][e]);
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = v[e] + quad_xor2(v[e]);
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = v[e] + quad_xor1(v[e]);
  return (v[0] + v[2]) + (v[1] + v[3]);
}

/// quad_raw's sum for a padded row of whole chunks only (no partial chunk),
/// with the metric fixed at compile time. All of a row's loads are issued
/// before any arithmetic (PC chunks at once, or groups of 8 when the chunk
/// count pc is only known at run time), so a row costs one memory round trip.
/// The summation order is quad_raw's: chunk c of a full 64-float block goes to
/// accumulator c mod 4, later chunks to accumulator 0, each accumulator in
/// increasing chunk order. Inactive quads return 0; their `row` must still be
/// a readable row (it is loaded). G: the group size of the run-time loop
/// (fewer registers, more round trips when smaller).
template <int PC, bool L2, int G = 8>
__device__ __forceinline__ float quad_raw_rows(const float* __restrict__ row, const float* q, uint32_t pc_rt,
                                               bool active) {
  static_assert(G % 4 == 0, "groups of whole 64-float blocks");
  constexpr int kGroup = PC > 0 ? PC : G;
  const uint32_t pc = PC > 0 ? static_cast<uint32_t>(PC) : pc_rt;
  const uint32_t full = (pc >> 2) << 2;
  const uint32_t g4 = static_cast<uint32_t>(lane_id() & 3) * 4;
  float a[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int e = 0; e < 4; ++e) a[i][e] = 0.0f;
  }
  for (uint32_t c0 = 0; c0 < pc; c0 += kGroup) {
    float4 x[kGroup];
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      if (PC > 0 || c0 + j < pc) x[j] = *reinterpret_cast<const float4*>(row + 16 * (c0 + j) + g4);
    }
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      const uint32_t c = c0 + static_cast<uint32_t>(j);
      if (PC == 0 && c >= pc) break;
      // The query is read from LDS four chunks at a time: the first four chunks' sums are
      // complete before the next loads, so fewer registers are live at once.
      if (j > 0 && (j & 3) == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
          for (int e = 0; e < 4; ++e) asm volatile("" : "+v"(a[i][e])::"memory");
        }
      }
      const float4 y = *reinterpret_cast<const float4*>(q + 16 * c + g4);
      if (c < full) {
        // c0 is a multiple of kGroup, which is a multiple of 4 when pc > kGroup: c mod 4 = j mod 4.
        switch (j & 3) {
          case 0: acc4(a[0], y, x[j], L2); break;
          case 1: acc4(a[1], y, x[j], L2); break;
          case 2: acc4(a[2], y, x[j], L2); break;
          default: acc4(a[3], y, x[j], L2); break;
        }
      } else {
        acc4(a[0], y, x[j], L2);
      }
    }
  }
  float v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = (a[0][e] + a[1][e]) + (a[2][e] + a[3][e]);
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = v[e] + quad_xor2(v[e]);
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = v[e] + quad_xor1(v[e]);
  const float r = (v[0] + v[2]) + (v[1] + v[3]);
  return active ? r : 0.0f;
}

/// The loads of one fp32 row of PC whole chunks by one quad lane (lane g:
/// positions 4g..4g+3 of every chunk), issued without waiting, so the loads
/// of several rows can be in flight together; sum() is quad_raw_rows' sum of
/// them (the same summation order, bit for bit).
template <int PC>
struct F32RowRegs {
  float4 x[PC];

  __device__ __forceinline__ void load(const float* __restrict__ row) {
    const uint32_t g4 = static_cast<uint32_t>(lane_id() & 3) * 4;
#pragma unroll
    for (int j = 0; j < PC; ++j) x[j] = *reinterpret_cast<const float4*>(row + 16 * j + g4);
  }
  template <bool L2>
  __device__ __forceinline__ float sum(const float* q) const {
    constexpr int kFull = (PC >> 2) << 2;
    const uint32_t g4 = static_cast<uint32_t>(lane_id() & 3) * 4;
    float a[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int e = 0; e < 4; ++e) a[i][e] = 0.0f;
    }
#pragma unroll
    for (int j = 0; j < PC; ++j) {
      if (j > 0 && (j & 3) == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
          for (int e = 0; e < 4; ++e) asm volatile("" : "+v"(a[i][e])::"memory");
        }
      }
      const float4 y = *reinterpret_cast<const float4*>(q + 16 * j + g4);
      if (j < kFull) {
        switch (j & 3) {
          case 0: acc4(a[0], y, x[j], L2); break;
          case 1: acc4(a[1], y, x[j], L2); break;
          case 2: acc4(a[2], y, x[j], L2); break;
          default: acc4(a[3], y, x[j], L2); break;
        }
      } else {
        acc4(a[0], y, x[j], L2);
      }
    }
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = (a[0][e] + a[1][e]) + (a[2][e] + a[3][e]);
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = v[e] + quad_xor2(v[e]);
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = v[e] + quad_xor1(v[e]);
    return (v[0] + v[2]) + (v[1] + v[3]);
  }
};

}  // namespace gpu
}  // namespace hnswamd
