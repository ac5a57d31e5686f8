#pragma once

// The cosine query scale of the reference's prepare_vector
// (src/space.hpp:31-40):
//
//   double norm = 0;  for i: norm += double(x[i]) * x[i];
//   inv = norm > 0 ? float(1.0 / sqrt(norm)) : 1.0f;
//
// reproduced bit for bit with integer arithmetic only, so the device needs no
// FP64. Every step is an IEEE binary64 operation rounded to nearest-even, done
// here on explicit (mantissa, exponent) pairs:
//   * x[i]^2 is exact in binary64 (a 24-bit mantissa squared has 48 bits);
//   * each addition of two non-negative values is rounded with guard and
//     sticky bits;
//   * sqrt is a digit-by-digit integer square root of a 106-bit radicand,
//     rounded by its remainder (a square root never lies exactly halfway);
//   * 1/r is a long division of 2^105 by the 53-bit mantissa, rounded by its
//     remainder (no halfway case either);
//   * the binary64 result is rounded to binary32 (subnormals and overflow
//     included).
// The sums stay inside the normal binary64 range: the smallest non-zero square
// is 2^-298 and the largest sum below 2^268 for any dimension up to 2^12.
// Inputs with NaN give 1 (norm > 0 is false), with an infinity 0, as in the
// reference.

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hnswamd {
namespace exact {

/// A non-negative binary64 value m * 2^e with m = 0 or 2^52 <= m < 2^53.
struct F64 {
  uint64_t m;
  int e;
};

__host__ __device__ inline int bit_length64(uint64_t v) {
  int n = 0;
  while (v != 0) {
    v >>= 1;
    ++n;
  }
  return n;
}

/// double(x) * x for a finite float x (exact).
__host__ __device__ inline F64 square_of(float x) {
  union {
    float f;
    uint32_t u;
  } b{x};
  const uint32_t bexp = (b.u >> 23) & 0xFFu;
  const uint32_t frac = b.u & 0x7FFFFFu;
  if (bexp == 0 && frac == 0) return F64{0, 0};
  const uint64_t mx = bexp == 0 ? frac : (frac | 0x800000u);
  const int ex = bexp == 0 ? -149 : static_cast<int>(bexp) - 150;
  const uint64_t sq = mx * mx;  // < 2^48
  const int shift = 53 - bit_length64(sq);
  return F64{sq << shift, 2 * ex - shift};
}

/// Round-to-nearest-even sum of two non-negative values.
__host__ __device__ inline F64 add(F64 a, F64 b) {
  if (a.m == 0) return b;
  if (b.m == 0) return a;
  if (a.e < b.e) {
    const F64 t = a;
    a = b;
    b = t;
  }
  const int d = a.e - b.e;
  if (d >= 54) return a;  // b < half an ulp of a
  // Ten guard bits: A in units of 2^(a.e - 10); B truncated, with a sticky bit.
  const uint64_t A = a.m << 10;
  const uint64_t Bw = b.m << 10;
  const uint64_t B = Bw >> d;
  const bool sticky = d > 0 && (Bw & ((uint64_t{1} << d) - 1)) != 0;
  const uint64_t S = A + B;  // < 2^64
  const int sh = (S >> 63) != 0 ? 11 : 10;
  uint64_t m = S >> sh;
  const uint64_t rem = S & ((uint64_t{1} << sh) - 1);
  const uint64_t half = uint64_t{1} << (sh - 1);
  if (rem > half || (rem == half && (sticky || (m & 1) != 0))) ++m;
  int e = a.e - 10 + sh;
  if (m == (uint64_t{1} << 53)) {
    m >>= 1;
    ++e;
  }
  return F64{m, e};
}

/// Correctly rounded square root of a positive normal value.
__host__ __device__ inline F64 sqrt_rn(F64 s) {
  uint64_t m = s.m;
  int e = s.e;
  if ((e & 1) != 0) {  // make the exponent even
    m <<= 1;
    --e;
  }
  // N = m * 2^52 < 2^106; floor(sqrt(N)) has 53 bits.
  const unsigned __int128 N = static_cast<unsigned __int128>(m) << 52;
  unsigned __int128 rem = N, res = 0;
  unsigned __int128 bit = static_cast<unsigned __int128>(1) << 104;
  while (bit > rem) bit >>= 2;
  while (bit != 0) {
    if (rem >= res + bit) {
      rem -= res + bit;
      res = (res >> 1) + bit;
    } else {
      res >>= 1;
    }
    bit >>= 2;
  }
  uint64_t r = static_cast<uint64_t>(res);
  // N lies above (r + 1/2)^2 = r^2 + r + 1/4 iff rem > r (no halfway case).
  if (rem > static_cast<unsigned __int128>(r)) ++r;
  int er = e / 2 - 26;
  if (r == (uint64_t{1} << 53)) {
    r >>= 1;
    ++er;
  }
  return F64{r, er};
}

/// Correctly rounded 1 / v of a positive normal value.
__host__ __device__ inline F64 recip_rn(F64 v) {
  if (v.m == (uint64_t{1} << 52)) return F64{v.m, -104 - v.e};  // a power of two
  // q = floor(2^105 / m), in (2^52, 2^53), by restoring division.
  uint64_t R = 1, q = 0;
  for (int i = 104; i >= 0; --i) {
    R <<= 1;
    if (R >= v.m) {
      R -= v.m;
      q |= uint64_t{1} << i;
    }
  }
  if (2 * R > v.m) ++q;  // no halfway case for a divisor strictly between powers of two
  int e = -105 - v.e;
  if (q == (uint64_t{1} << 53)) {
    q >>= 1;
    ++e;
  }
  return F64{q, e};
}

/// Round a positive binary64 value to binary32 (nearest-even), as bits.
__host__ __device__ inline uint32_t to_f32_bits(F64 y) {
  // y = m * 2^e with a 53-bit m; a float has 24 bits: f = mf * 2^ef.
  int ef = y.e + 29;
  int sh = 29;
  if (ef < -149) {  // subnormal range: the ulp is 2^-149
    sh += -149 - ef;
    ef = -149;
  }
  if (sh > 60) return 0u;  // below half the smallest subnormal
  uint64_t mf = y.m >> sh;
  const uint64_t rem = y.m & ((uint64_t{1} << sh) - 1);
  const uint64_t half = uint64_t{1} << (sh - 1);
  if (rem > half || (rem == half && (mf & 1) != 0)) ++mf;
  if (mf == (uint64_t{1} << 24)) {
    mf >>= 1;
    ++ef;
  }
  if (mf < (uint64_t{1} << 23)) return static_cast<uint32_t>(mf);  // subnormal (or the smallest normal)
  const int biased = ef + 150;
  if (biased >= 255) return 0x7F800000u;
  return (static_cast<uint32_t>(biased) << 23) | static_cast<uint32_t>(mf - (uint64_t{1} << 23));
}

/// float(1.0 / sqrt(sum of double(x[i])^2)), 1 for a zero vector (bits).
__host__ __device__ inline uint32_t inv_norm_bits(const float* x, uint32_t dim) {
  F64 s{0, 0};
  bool nan = false, inf = false;
  for (uint32_t i = 0; i < dim; ++i) {
    union {
      float f;
      uint32_t u;
    } b{x[i]};
    if ((b.u & 0x7F800000u) == 0x7F800000u) {
      if ((b.u & 0x7FFFFFu) != 0) {
        nan = true;
      } else {
        inf = true;
      }
      continue;
    }
    s = add(s, square_of(x[i]));
  }
  if (nan) return 0x3F800000u;  // norm is NaN: norm > 0 is false
  if (inf) return 0u;           // 1 / sqrt(inf) = 0
  if (s.m == 0) return 0x3F800000u;
  return to_f32_bits(recip_rn(sqrt_rn(s)));
}

}  // namespace exact
}  // namespace hnswamd
