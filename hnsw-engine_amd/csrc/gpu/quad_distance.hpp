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

}  // namespace gpu
}  // namespace hnswamd
