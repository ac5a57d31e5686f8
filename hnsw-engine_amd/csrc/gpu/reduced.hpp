#pragma once

// Reduced-precision copies of the resident vectors and their distances.
//
//   BF16  2 bytes per element, round-to-nearest-even from fp32.
//   FP8   OCP e4m3 (gfx950's native fp8 format, largest finite 448), one byte
//         per element, with a per-vector fp32 scale s = max|x| / 448:
//         x ~ s * e4m3(x / s).
// Both keep |x^|^2 per row (of the stored, rounded vector) so that L2 can be
// evaluated as |q^|^2 + |x^|^2 - 2 q^.x^. Dot products accumulate in fp32.
// These distances only steer the traversal; the final candidates are re-ranked
// with the exact fp32 distance, so reported distances are exact.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace hnswamd {
namespace gpu {

enum Storage : int { kStoreF32 = 0, kStoreBF16 = 1, kStoreFP8 = 2 };

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t bf16_bits_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u);  // inf / nan
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ float bf16_value(uint32_t bits16) { return __uint_as_float(bits16 << 16); }

__device__ __forceinline__ float dot2_bf16(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, a), __builtin_bit_cast(bf16x2_t, b), c,
                                         false);
}

/// Dot product of a bf16 row with the bf16 query (both `padded` elements),
/// by the four lanes of a quad: lane g takes 16-byte pieces g, g+4, ...
/// (accumulator 0 the even element pairs of a piece, 1 the odd ones, pieces
/// in increasing order). Identical in all four lanes; 0 for inactive quads,
/// whose `row` must still be readable (it is loaded). With PC (= padded / 16)
/// known at compile time every load of the row is issued before the
/// arithmetic.
template <int PC, int G = 4>
__device__ __forceinline__ float quad_dot_bf16(const uint16_t* __restrict__ row, const uint16_t* q,
                                               uint32_t padded, bool active) {
  const uint32_t g = static_cast<uint32_t>(lane_id() & 3);
  float a0 = 0.0f, a1 = 0.0f;
  const uint4* r4 = reinterpret_cast<const uint4*>(row);
  const uint4* q4 = reinterpret_cast<const uint4*>(q);
  if constexpr (PC > 0) {
    static_assert(PC % 2 == 0, "bf16 rows of whole 16-byte pieces per lane");
    constexpr int kPer = PC / 2;  // 2 PC pieces of 8 elements over 4 lanes
    uint4 x[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) x[i] = r4[g + 4 * i];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const uint4 y = q4[g + 4 * i];
      a0 = dot2_bf16(x[i].x, y.x, a0);
      a1 = dot2_bf16(x[i].y, y.y, a1);
      a0 = dot2_bf16(x[i].z, y.z, a0);
      a1 = dot2_bf16(x[i].w, y.w, a1);
    }
  } else {
    // Groups of up to G pieces per lane (G = 4: padded <= 128 in one group),
    // all loads of a group issued before its arithmetic.
    const uint32_t pieces = padded >> 3;
    for (uint32_t i0 = 0; 4 * i0 < pieces; i0 += G) {
      uint4 x[G];
#pragma unroll
      for (uint32_t i = 0; i < G; ++i) {
        const uint32_t pi = g + 4 * (i0 + i);
        if (pi < pieces) x[i] = r4[pi];
      }
#pragma unroll
      for (uint32_t i = 0; i < G; ++i) {
        const uint32_t pi = g + 4 * (i0 + i);
        if (pi >= pieces) break;
        const uint4 y = q4[pi];
        a0 = dot2_bf16(x[i].x, y.x, a0);
        a1 = dot2_bf16(x[i].y, y.y, a1);
        a0 = dot2_bf16(x[i].z, y.z, a0);
        a1 = dot2_bf16(x[i].w, y.w, a1);
      }
    }
  }
  float v = a0 + a1;
  v = v + quad_xor1(v);
  v = v + quad_xor2(v);
  return active ? v : 0.0f;
}

__device__ __forceinline__ void fp8_piece(uint4 x, const float* q16, float& a0, float& a1) {
  const float4* qp = reinterpret_cast<const float4*>(q16);
  const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float4 y = qp[j];
    const f32x2_t lo = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(w[j]), false);
    const f32x2_t hi = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(w[j]), true);
    a0 = __builtin_fmaf(lo.x, y.x, a0);
    a1 = __builtin_fmaf(lo.y, y.y, a1);
    a0 = __builtin_fmaf(hi.x, y.z, a0);
    a1 = __builtin_fmaf(hi.y, y.w, a1);
  }
}

/// Dot product of an e4m3 row (before its scale) with the fp32 query: lane g
/// of the quad takes 16-element pieces g, g+4, ...; 0 for inactive quads,
/// whose `row` must still be readable (it is loaded). With PC (= padded / 16)
/// known at compile time every load of the row is issued before the
/// arithmetic.
template <int PC>
__device__ __forceinline__ float quad_dot_fp8(const uint8_t* __restrict__ row, const float* q, uint32_t padded,
                                              bool active) {
  const uint32_t g = static_cast<uint32_t>(lane_id() & 3);
  float a0 = 0.0f, a1 = 0.0f;
  const uint4* r4 = reinterpret_cast<const uint4*>(row);
  if constexpr (PC > 0) {
    static_assert(PC % 4 == 0, "fp8 rows of whole 16-byte pieces per lane");
    constexpr int kPer = PC / 4;  // PC pieces of 16 elements over 4 lanes
    uint4 x[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) x[i] = r4[g + 4 * i];
#pragma unroll
    for (int i = 0; i < kPer; ++i) fp8_piece(x[i], q + 16 * (g + 4 * i), a0, a1);
  } else {
    // Groups of up to 8 pieces per lane (padded <= 512: one group), all loads
    // of a group issued before its arithmetic.
    const uint32_t pieces = padded >> 4;  // 16 elements per piece
    for (uint32_t i0 = 0; 4 * i0 < pieces; i0 += 8) {
      uint4 x[8];
#pragma unroll
      for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t pi = g + 4 * (i0 + i);
        if (pi < pieces) x[i] = r4[pi];
      }
#pragma unroll
      for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t pi = g + 4 * (i0 + i);
        if (pi >= pieces) break;
        fp8_piece(x[i], q + 16 * pi, a0, a1);
      }
    }
  }
  float v = a0 + a1;
  v = v + quad_xor1(v);
  v = v + quad_xor2(v);
  return active ? v : 0.0f;
}

/// The loads of one bf16 row of PC 16-element chunks (PC / 2 16-byte pieces
/// per quad lane) issued without waiting; dot() is quad_dot_bf16's sum of
/// them (the same order, bit for bit).
template <int PC>
struct BF16RowRegs {
  static_assert(PC % 2 == 0, "bf16 rows of whole 16-byte pieces per lane");
  uint4 x[PC / 2];

  __device__ __forceinline__ void load(const uint16_t* __restrict__ row) {
    const uint32_t g = static_cast<uint32_t>(lane_id() & 3);
    const uint4* r4 = reinterpret_cast<const uint4*>(row);
#pragma unroll
    for (int i = 0; i < PC / 2; ++i) x[i] = r4[g + 4 * i];
  }
  __device__ __forceinline__ float dot(const uint16_t* q) const {
    const uint32_t g = static_cast<uint32_t>(lane_id() & 3);
    const uint4* q4 = reinterpret_cast<const uint4*>(q);
    float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
    for (int i = 0; i < PC / 2; ++i) {
      const uint4 y = q4[g + 4 * i];
      a0 = dot2_bf16(x[i].x, y.x, a0);
      a1 = dot2_bf16(x[i].y, y.y, a1);
      a0 = dot2_bf16(x[i].z, y.z, a0);
      a1 = dot2_bf16(x[i].w, y.w, a1);
    }
    float v = a0 + a1;
    v = v + quad_xor1(v);
    v = v + quad_xor2(v);
    return v;
  }
};

/// The loads of one e4m3 row of PC 16-element chunks (PC / 4 pieces per quad
/// lane) issued without waiting; dot() is quad_dot_fp8's sum of them (before
/// the row's scale).
template <int PC>
struct FP8RowRegs {
  static_assert(PC % 4 == 0, "fp8 rows of whole 16-byte pieces per lane");
  uint4 x[PC / 4];

  __device__ __forceinline__ void load(const uint8_t* __restrict__ row) {
    const uint32_t g = static_cast<uint32_t>(lane_id() & 3);
    const uint4* r4 = reinterpret_cast<const uint4*>(row);
#pragma unroll
    for (int i = 0; i < PC / 4; ++i) x[i] = r4[g + 4 * i];
  }
  __device__ __forceinline__ float dot(const float* q) const {
    const uint32_t g = static_cast<uint32_t>(lane_id() & 3);
    float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
    for (int i = 0; i < PC / 4; ++i) fp8_piece(x[i], q + 16 * (g + 4 * i), a0, a1);
    float v = a0 + a1;
    v = v + quad_xor1(v);
    v = v + quad_xor2(v);
    return v;
  }
};

}  // namespace gpu
}  // namespace hnswamd
