#pragma once

// Device helpers shared by the HIP kernels (gfx950, wave64).

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hnswamd {
namespace gpu {

constexpr int kWave = 64;
constexpr uint64_t kKeyInf = ~0ULL;
constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;
constexpr uint32_t kOrdInf = 0xFFFFFFFFu;  // larger than the key of any non-NaN distance

// fp32 distance -> u32 whose unsigned order is the float order (no NaN).
__device__ __forceinline__ uint32_t ord_of(float d) {
  const uint32_t u = __float_as_uint(d);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float float_of(uint32_t o) {
  const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
  return __uint_as_float(u);
}
// (distance, id) packed so that u64 order is the (distance, id) order.
__device__ __forceinline__ uint64_t key_of(uint32_t ord, uint32_t id) {
  return (static_cast<uint64_t>(ord) << 32) | id;
}
__device__ __forceinline__ uint32_t key_ord(uint64_t k) { return static_cast<uint32_t>(k >> 32); }
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return static_cast<uint32_t>(k); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rdlane(uint32_t v, int lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), lane));
}
__device__ __forceinline__ uint64_t rdlane64(uint64_t v, int lane) {
  const uint32_t lo = rdlane(static_cast<uint32_t>(v), lane);
  const uint32_t hi = rdlane(static_cast<uint32_t>(v >> 32), lane);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = uniform(static_cast<uint32_t>(v));
  const uint32_t hi = uniform(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Value of lane (lane - 1) / (lane + 1) through the LDS crossbar (no memory).
__device__ __forceinline__ uint32_t from_lane(uint32_t v, int src_lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_lane << 2, static_cast<int>(v)));
}
__device__ __forceinline__ uint64_t from_lane64(uint64_t v, int src_lane) {
  const uint32_t lo = from_lane(static_cast<uint32_t>(v), src_lane);
  const uint32_t hi = from_lane(static_cast<uint32_t>(v >> 32), src_lane);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Whole-wave DPP shifts (wave_shr:1 / wave_shl:1): lane i gets the value of
// lane i - 1 / i + 1, and the lane without a source (0 / 63) gets `fill`. One
// VALU operation, where a ds_bpermute is a round trip through the LDS crossbar.
// All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v, uint32_t fill) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(static_cast<int>(fill), static_cast<int>(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ uint32_t wave_shl1(uint32_t v, uint32_t fill) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(static_cast<int>(fill), static_cast<int>(v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ uint64_t wave_shr1_64(uint64_t v, uint64_t fill) {
  const uint32_t lo = wave_shr1(static_cast<uint32_t>(v), static_cast<uint32_t>(fill));
  const uint32_t hi = wave_shr1(static_cast<uint32_t>(v >> 32), static_cast<uint32_t>(fill >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_shl1_64(uint64_t v, uint64_t fill) {
  const uint32_t lo = wave_shl1(static_cast<uint32_t>(v), static_cast<uint32_t>(fill));
  const uint32_t hi = wave_shl1(static_cast<uint32_t>(v >> 32), static_cast<uint32_t>(fill >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Quad permutes through DPP: partner lane (lane ^ 1) / (lane ^ 2) within each
// group of four lanes.
__device__ __forceinline__ float quad_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float quad_xor2(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
}

/// A sorted ascending array of u64 keys spread over the wave: element k lives in
/// row k / 64 of lane k % 64. Empty slots hold kKeyInf. Every member is called
/// by the whole wave with uniform arguments.
template <int E>
struct WaveSorted {
  uint64_t v[E];

  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = kKeyInf;
  }
  __device__ __forceinline__ uint32_t rank(uint64_t x) const {
    uint32_t c = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) c += static_cast<uint32_t>(__popcll(ballot(v[e] < x)));
    return c;
  }
  __device__ __forceinline__ uint64_t at(uint32_t k) const {
    uint64_t r = kKeyInf;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (static_cast<uint32_t>(e) == (k >> 6)) r = rdlane64(v[e], static_cast<int>(k & 63));
    }
    return r;
  }
  /// Insert x keeping the first `cap` (<= 64 * E) elements. Returns the element
  /// that no longer fits (x itself, or the old element at cap - 1), or kKeyInf.
  /// Rows wholly before x's position do not change and are skipped.
  __device__ __forceinline__ uint64_t insert(uint64_t x, uint32_t cap) {
    const uint32_t pos = rank(x);
    if (pos >= cap) return x;
    const uint64_t out = at(cap - 1);
    const int lane = lane_id();
#pragma unroll
    for (int e = E - 1; e >= 0; --e) {
      if (static_cast<uint32_t>(e) >= (pos >> 6)) {
        const uint32_t k = static_cast<uint32_t>(e * 64 + lane);
        // Lane 0 of row e takes the last element of row e - 1 (unused in row 0).
        const uint64_t up = wave_shr1_64(v[e], e > 0 ? rdlane64(v[e > 0 ? e - 1 : 0], 63) : kKeyInf);
        uint64_t nv = v[e];
        if (k == pos) nv = x;
        if (k > pos) nv = up;
        if (k >= cap) nv = kKeyInf;
        v[e] = nv;
      }
    }
    return out;
  }
  /// insert() that also moves a per-slot flag word with its entry; x gets
  /// `xflag`, slots at cap and beyond get flag 1 (the flag of an empty slot),
  /// and the flag of the element pushed out is stored in out_flag (0 when x
  /// itself does not fit).
  __device__ __forceinline__ uint64_t insert_flagged(uint64_t x, uint32_t cap, uint32_t (&fl)[E], uint32_t xflag,
                                                     uint32_t& out_flag) {
    const uint32_t pos = rank(x);
    out_flag = 0;
    if (pos >= cap) return x;
    const uint64_t out = at(cap - 1);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (static_cast<uint32_t>(e) == ((cap - 1) >> 6)) out_flag = rdlane(fl[e], static_cast<int>((cap - 1) & 63));
    }
    const int lane = lane_id();
#pragma unroll
    for (int e = E - 1; e >= 0; --e) {
      if (static_cast<uint32_t>(e) >= (pos >> 6)) {
        const uint32_t k = static_cast<uint32_t>(e * 64 + lane);
        // Lane 0 of row e takes the last entry of row e - 1 (unused in row 0).
        const int pe = e > 0 ? e - 1 : 0;
        const uint64_t up = wave_shr1_64(v[e], e > 0 ? rdlane64(v[pe], 63) : kKeyInf);
        const uint32_t fup = wave_shr1(fl[e], e > 0 ? rdlane(fl[pe], 63) : 1u);
        uint64_t nv = v[e];
        uint32_t nf = fl[e];
        if (k == pos) {
          nv = x;
          nf = xflag;
        }
        if (k > pos) {
          nv = up;
          nf = fup;
        }
        if (k >= cap) {
          nv = kKeyInf;
          nf = 1u;
        }
        v[e] = nv;
        fl[e] = nf;
      }
    }
    return out;
  }
  /// Ascending sort of all 64 * E entries (entry 64e + l in row e, lane l) by
  /// a bitonic network: partners more than 63 entries apart are in other rows
  /// of the same lane (a register compare-exchange), nearer ones in other
  /// lanes of the same row (a lane gather). E must be a power of two.
  __device__ __forceinline__ void sort_all() {
    static_assert((E & (E - 1)) == 0, "E must be a power of two");
    const uint32_t lane = static_cast<uint32_t>(lane_id());
#pragma unroll
    for (uint32_t k = 2; k <= 64u * E; k <<= 1) {
#pragma unroll
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        if (j >= 64) {
          const uint32_t jr = j >> 6;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            if ((static_cast<uint32_t>(e) & jr) == 0) {
              const int f = e | static_cast<int>(jr);
              // Direction: bit k of the lower index 64e + lane; k >= 128 here, so the row decides.
              const bool up = ((static_cast<uint32_t>(e) * 64u) & k) == 0;
              const uint64_t a = v[e], b = v[f < E ? f : e];
              const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
              v[e] = up ? lo : hi;
              v[f < E ? f : e] = up ? hi : lo;
            }
          }
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const uint64_t o = from_lane64(v[e], static_cast<int>(lane ^ j));
            const uint32_t i = static_cast<uint32_t>(e) * 64u + lane;
            const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
            v[e] = keep_min ? (o < v[e] ? o : v[e]) : (o > v[e] ? o : v[e]);
          }
        }
      }
    }
  }
  /// Remove element 0.
  __device__ __forceinline__ void pop_front() {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      // Lane 63 of row e takes the first element of row e + 1 (empty after the last row).
      const uint64_t carry = e + 1 < E ? rdlane64(v[e + 1 < E ? e + 1 : e], 0) : kKeyInf;
      v[e] = wave_shl1_64(v[e], carry);
    }
  }
};

}  // namespace gpu
}  // namespace hnswamd
