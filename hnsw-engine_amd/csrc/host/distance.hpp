#pragma once

// Host-side fp32 distances in the fixed summation order the whole engine uses:
// 64 partial sums (lane p of accumulator q holds positions 64t + 16q + p), whole
// 16-float chunks past the last 64-float chunk folded into accumulator 0, a final
// partial chunk into accumulator 1, then (acc0 + acc1) + (acc2 + acc3) per lane
// and a halving tree over the 16 lanes (p + 8, p + 4, p + 2, p + 1).
//
// This is the order of the reference's AVX-512 kernels
// (src/kernels_avx512.cpp:10-58, reduced by GCC's _mm512_reduce_add_ps). The GPU
// kernels compute the same order with 16-lane groups and cross-lane xor
// shuffles, so the builder (which uses these functions), the GPU search and the
// reference on an AVX-512 host all see bit-identical distances. Written with
// std::fma and compiled with -ffp-contract=off; on a host with AVX2+FMA the
// per-lane loops vectorise without reassociation.

#include <cmath>
#include <cstddef>

namespace hnswamd {

enum Metric : int { kL2 = 0, kInnerProduct = 1, kCosine = 2 };

// Always inlined, so that a caller compiled for a wider ISA (the builder's
// AVX-512 clone) runs the loops in its own registers.
template <bool kSquaredDiff>
[[gnu::always_inline]] inline float ordered_sum(const float* __restrict a, const float* __restrict b, std::size_t n) noexcept {
  float acc[4][16] = {};
  std::size_t i = 0;
  for (; i + 64 <= n; i += 64) {
    for (int q = 0; q < 4; ++q) {
      for (int p = 0; p < 16; ++p) {
        const float x = a[i + 16 * q + p], y = b[i + 16 * q + p];
        if (kSquaredDiff) {
          const float d = x - y;
          acc[q][p] = std::fma(d, d, acc[q][p]);
        } else {
          acc[q][p] = std::fma(x, y, acc[q][p]);
        }
      }
    }
  }
  for (; i + 16 <= n; i += 16) {
    for (int p = 0; p < 16; ++p) {
      const float x = a[i + p], y = b[i + p];
      if (kSquaredDiff) {
        const float d = x - y;
        acc[0][p] = std::fma(d, d, acc[0][p]);
      } else {
        acc[0][p] = std::fma(x, y, acc[0][p]);
      }
    }
  }
  for (std::size_t p = 0; i + p < n; ++p) {
    const float x = a[i + p], y = b[i + p];
    if (kSquaredDiff) {
      const float d = x - y;
      acc[1][p] = std::fma(d, d, acc[1][p]);
    } else {
      acc[1][p] = std::fma(x, y, acc[1][p]);
    }
  }
  float v[16];
  for (int p = 0; p < 16; ++p) v[p] = (acc[0][p] + acc[1][p]) + (acc[2][p] + acc[3][p]);
  for (int w = 8; w >= 1; w >>= 1) {
    for (int p = 0; p < w; ++p) v[p] = v[p] + v[p + w];
  }
  return v[0];
}

inline float l2_sq(const float* a, const float* b, std::size_t n) noexcept { return ordered_sum<true>(a, b, n); }
inline float dot(const float* a, const float* b, std::size_t n) noexcept { return ordered_sum<false>(a, b, n); }

/// Squared L2 for L2; 1 - dot for inner product and cosine (src/space.hpp:14-27).
inline float distance(int metric, const float* a, const float* b, std::size_t n) noexcept {
  return metric == kL2 ? l2_sq(a, b, n) : 1.0f - dot(a, b, n);
}

/// Copy a vector, scaling it to unit length for cosine with the norm taken in
/// double; a zero vector is copied unchanged (src/space.hpp:31-40).
inline void prepare_vector(int metric, const float* src, float* dst, std::size_t n) noexcept {
  if (metric != kCosine) {
    for (std::size_t i = 0; i < n; ++i) dst[i] = src[i];
    return;
  }
  double ss = 0.0;
  for (std::size_t i = 0; i < n; ++i) ss += static_cast<double>(src[i]) * static_cast<double>(src[i]);
  const float inv = ss > 0.0 ? static_cast<float>(1.0 / std::sqrt(ss)) : 1.0f;
  for (std::size_t i = 0; i < n; ++i) dst[i] = src[i] * inv;
}

}  // namespace hnswamd
