#pragma once

// Graph search on the GPU: greedy descent through the upper layers, then
// Alg. 2 (search_layer) on layer 0, one wavefront per query.
//
// The semantics are the reference's (src/search_layer.cpp:29-171):
//   * greedy: per pass, the first strict minimum in list order replaces the
//     current node if it is strictly closer; passes repeat until none does;
//   * layer 0: pop the (distance, id)-smallest candidate; stop when |R| >= ef
//     and its distance is strictly above the worst result; visit neighbours in
//     list order, admitting one iff |R| < ef or its distance is strictly below
//     the worst result; admitted nodes join C and, when eligible, R, which keeps
//     the ef (distance, id)-smallest;
//   * output: R in (distance, id) order, labels looked up, padded with
//     {~0, +inf}.
// Distances are fp32 in the engine's fixed summation order (see
// host/distance.hpp), computed by 4-lane groups that each hold 16 partial sums,
// so they are bit-identical to the host and to the reference's AVX-512 kernels.
//
// Fast kernel (one wavefront per query, persistent over an atomic ticket):
//   * R (capacity ef <= 64 * ER) and C (capacity 64 * EC) are sorted arrays of
//     (distance, id) keys held in registers across the wave;
//   * the visited set is an open-addressing hash in LDS;
//   * C keeps its 64 * EC smallest entries. The smallest entry it ever dropped
//     is remembered: the search is exact as long as that entry is never the
//     true minimum of C at a point where the exact algorithm would go on. If it
//     is, or if the visited hash reaches 3/4 load, the query is queued for the
//     general kernel.
// General kernel: the same algorithm with a visited bitmap and two binary heaps
// in HBM (unbounded C, any ef), used for queued queries and ef > 512.

//
// This header holds the device code of both kernels. The fast kernel's
// instantiations (storage x row width x filter mode x capacity class) are
// compiled in slices by fast_variant.hip, one object per (storage, row width),
// so they build in parallel; search_kernels.hip holds the general kernel, the
// small kernels and the launchers.

#include <hip/hip_runtime.h>
#include <math.h>

#include "device_common.hpp"
#include "exact_scale.hpp"
#include "launch.hpp"
#include "quad_distance.hpp"
#include "reduced.hpp"

namespace hnswamd {
namespace gpu {

namespace {


constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;
typedef __attribute__((address_space(1))) uint32_t global_u32;

/// A 16-byte LDS read that is never cached in registers (the hash words change
/// under concurrent compare-and-swaps of other lanes). The pointer must point
/// into LDS: a volatile access through a generic pointer would compile to a
/// FLAT load, whose wait also retires every global load in flight.
__device__ __forceinline__ u32x4 lds_read16(const uint32_t* p) {
  return *reinterpret_cast<const volatile lds_u32x4*>(reinterpret_cast<uintptr_t>(p));
}

/// Bit `id` of an eligibility bitmap in global memory (a global load, not a
/// FLAT one: the pointer arrives through a struct and its address space is
/// otherwise unknown to the compiler).
__device__ __forceinline__ bool elig_bit(const uint32_t* elig, uint32_t id) {
  const auto* g = reinterpret_cast<const global_u32*>(reinterpret_cast<uintptr_t>(elig));
  return (g[id >> 5] >> (id & 31)) & 1u;
}

struct WaveLds {
  uint32_t* hash;  // visited hash slots
  uint32_t* nid;   // 64 compacted neighbour ids
  uint32_t* nd;    // 64 ordered distances
  uint32_t* nel;   // 64 eligibility flags
  float* q;        // the prepared query (padded)
  uint16_t* qb;    // the query rounded to bf16 (bf16 storage)
  float* qn;       // |q^|^2 of the query as the reduced distance sees it
};

__device__ __forceinline__ WaveLds wave_lds(uint32_t* base, uint32_t log2h, uint32_t padded) {
  const uint32_t H = 1u << log2h;
  return WaveLds{base,
                 base + H,
                 base + H + 64,
                 base + H + 128,
                 reinterpret_cast<float*>(base + H + 192),
                 reinterpret_cast<uint16_t*>(base + H + 192 + padded),
                 reinterpret_cast<float*>(base + H + 192 + padded + padded / 2)};
}

/// The loads one quad lane makes for one row of a fixed width (PC chunks) of
/// storage ST, and the row's distance from them: the row itself, and for a
/// reduced copy its |x^|^2 (L2) and fp8 scale. Inactive quads load row 0.
template <int PC, int ST>
struct RowRegs;

template <int PC>
struct RowRegs<PC, kStoreF32> {
  F32RowRegs<PC> r;
  template <bool L2>
  __device__ __forceinline__ void load(const DevIndexView& ix, uint32_t id, bool) {
    r.load(ix.vecs + static_cast<size_t>(id) * (16 * PC));
  }
  template <bool L2>
  __device__ __forceinline__ float dist(const WaveLds& w, float) const {
    const float d = r.template sum<L2>(w.q);
    return L2 ? d : 1.0f - d;
  }
};

template <int PC>
struct RowRegs<PC, kStoreBF16> {
  BF16RowRegs<PC> r;
  float xn;
  template <bool L2>
  __device__ __forceinline__ void load(const DevIndexView& ix, uint32_t id, bool active) {
    xn = (L2 && active) ? ix.vnorm[id] : 0.0f;
    r.load(static_cast<const uint16_t*>(ix.vlow) + static_cast<size_t>(id) * (16 * PC));
  }
  template <bool L2>
  __device__ __forceinline__ float dist(const WaveLds& w, float qn) const {
    const float dp = r.dot(w.qb);
    return L2 ? (qn + xn) - 2.0f * dp : 1.0f - dp;
  }
};

template <int PC>
struct RowRegs<PC, kStoreFP8> {
  FP8RowRegs<PC> r;
  float xn, xs;
  template <bool L2>
  __device__ __forceinline__ void load(const DevIndexView& ix, uint32_t id, bool active) {
    xn = (L2 && active) ? ix.vnorm[id] : 0.0f;
    xs = active ? ix.vscale[id] : 0.0f;
    r.load(static_cast<const uint8_t*>(ix.vlow) + static_cast<size_t>(id) * (16 * PC));
  }
  template <bool L2>
  __device__ __forceinline__ float dist(const WaveLds& w, float qn) const {
    const float dp = r.dot(w.q) * xs;
    return L2 ? (qn + xn) - 2.0f * dp : 1.0f - dp;
  }
};

/// Ordered distances of the rows w.nid[0..U) into w.nd[0..U), 16 rows per step,
/// from the fp32 rows (ST = 0, exact) or the bf16 / fp8 copy (ST = 1 / 2).
/// Every load a row needs (its norm and scale included) is issued before its
/// arithmetic, so a step costs one memory round trip. With a fixed row width
/// (PC > 0), RT steps (16 * RT rows) share one round trip: all their loads
/// are issued before the arithmetic of the first.
template <int PC, int ST, bool L2, int G, int BG, int RT>
__device__ __forceinline__ void row_dists_metric(const DevIndexView& ix, const WaveLds& w, uint32_t U) {
  const int lane = lane_id();
  const float qn = ST == kStoreF32 ? 0.0f : w.qn[0];
  if constexpr (PC > 0) {
    for (uint32_t base = 0; base < U; base += 16 * RT) {
      RowRegs<PC, ST> rr[RT];
#pragma unroll
      for (int s = 0; s < RT; ++s) {
        if (s == 0 || base + 16u * s < U) {
          const uint32_t r = base + 16u * s + static_cast<uint32_t>(lane >> 2);
          const bool active = r < U;
          rr[s].template load<L2>(ix, active ? w.nid[r] : 0u, active);
        }
      }
#pragma unroll
      for (int s = 0; s < RT; ++s) {
        if (s == 0 || base + 16u * s < U) {
          const uint32_t r = base + 16u * s + static_cast<uint32_t>(lane >> 2);
          const float d = rr[s].template dist<L2>(w, qn);
          if (r < U && (lane & 3) == 0) w.nd[r] = ord_of(d);
        }
      }
    }
    return;
  }
  for (uint32_t base = 0; base < U; base += 16) {
    const uint32_t r = base + static_cast<uint32_t>(lane >> 2);
    const bool active = r < U;
    const uint32_t id = active ? w.nid[r] : 0u;
    float d;
    if constexpr (ST == kStoreF32) {
      const float* row = ix.vecs + static_cast<size_t>(id) * ix.padded;
      d = quad_raw_rows<PC, L2, G>(row, w.q, ix.padded >> 4, active);
      if (!L2) d = 1.0f - d;
    } else {
      const float xn = (L2 && active) ? ix.vnorm[id] : 0.0f;
      float dp;
      if constexpr (ST == kStoreBF16) {
        const uint16_t* row = static_cast<const uint16_t*>(ix.vlow) + static_cast<size_t>(id) * ix.padded;
        dp = quad_dot_bf16<PC, BG>(row, w.qb, ix.padded, active);
      } else {
        const float xs = active ? ix.vscale[id] : 0.0f;
        const uint8_t* row = static_cast<const uint8_t*>(ix.vlow) + static_cast<size_t>(id) * ix.padded;
        dp = quad_dot_fp8<PC>(row, w.q, ix.padded, active) * xs;
      }
      d = L2 ? (qn + xn) - 2.0f * dp : 1.0f - dp;
    }
    if (active && (lane & 3) == 0) w.nd[r] = ord_of(d);
  }
}

/// G / BG: group sizes of the run-time-width fp32 / bf16 row loops (chunks /
/// 16-byte pieces per lane loaded before their arithmetic; a bf16 row of up to
/// 128 elements is one group of 4 pieces, so one memory round trip).
/// RT: 16-row steps per memory round trip where the row width is fixed.
template <int PC, int ST = kStoreF32, int G = 8, int BG = 4, int RT = 1>
__device__ __forceinline__ void row_dists(const DevIndexView& ix, const WaveLds& w, uint32_t U) {
  if (ix.metric == 0) {
    row_dists_metric<PC, ST, true, G, BG, RT>(ix, w, U);
  } else {
    row_dists_metric<PC, ST, false, G, BG, RT>(ix, w, U);
  }
}


template <int ctrl, int row_mask>
__device__ __forceinline__ uint64_t dpp_min64(uint64_t v) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(static_cast<uint32_t>(v)), static_cast<int>(static_cast<uint32_t>(v)), ctrl, row_mask, 0xF, false));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(static_cast<uint32_t>(v >> 32)), static_cast<int>(static_cast<uint32_t>(v >> 32)), ctrl,
      row_mask, 0xF, false));
  const uint64_t o = (static_cast<uint64_t>(hi) << 32) | lo;
  return o < v ? o : v;
}

/// Minimum over the wave (all 64 lanes active), by DPP: quads, half rows and
/// rows of 16 by lane permutes, then rows 1 and 3 take rows 0 and 2 (row_bcast
/// of lane 15) and rows 2-3 take lane 31: lane 63 holds the minimum.
__device__ __forceinline__ uint64_t wave_min64(uint64_t v) {
  v = dpp_min64<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v = dpp_min64<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v = dpp_min64<0x141, 0xF>(v);  // row_half_mirror
  v = dpp_min64<0x140, 0xF>(v);  // row_mirror
  v = dpp_min64<0x142, 0xA>(v);  // row_bcast:15
  v = dpp_min64<0x143, 0xC>(v);  // row_bcast:31
  return rdlane64(v, 63);
}

/// Insert id into the visited hash; true iff it was not there.
__device__ __forceinline__ bool hash_visit(uint32_t* tab, uint32_t log2h, uint32_t id) {
  const uint32_t mask = (1u << log2h) - 1u;
  uint32_t h = (id * 0x9E3779B1u) >> (32u - log2h);
  for (;;) {
    const uint32_t old = atomicCAS(&tab[h], kEmptySlot, id);
    if (old == kEmptySlot) return true;
    if (old == id) return false;
    h = (h + 1u) & mask;
  }
}

/// True iff id is in the LDS hash (no insertion).
__device__ __forceinline__ bool hash_has(const uint32_t* tab, uint32_t log2h, uint32_t id) {
  const uint32_t mask = (1u << log2h) - 1u;
  uint32_t h = (id * 0x9E3779B1u) >> (32u - log2h);
  for (;;) {
    const uint32_t v = tab[h];
    if (v == id) return true;
    if (v == kEmptySlot) return false;
    h = (h + 1u) & mask;
  }
}

/// Compact LDS hash: 2^L slots of 16 bits, two per LDS word (0 = empty).
/// Ids are permuted by an odd multiplier mod 2^B (a bijection on B-bit ids);
/// the top L bits of the permuted id give the home slot and the other
/// S = B - L <= 8 bits its remainder. An entry stores 1 + (the remainder with
/// the low 7 bits of the home slot), so an id's entry is the same in every slot
/// it may occupy (linear probing, at most `maxp` <= 127 slots past home):
/// within that window the low home bits identify the home slot, so the entry
/// identifies the id exactly, and it is never 0. Probes read the aligned group
/// of 8 slots (16 bytes) holding the current slot and test all 8 against the
/// one expected entry with packed 16-bit operations: most lookups cost one LDS
/// read and a few instructions (and one compare-and-swap to insert). The table
/// must be 16-byte aligned, L >= 8.
struct Hash16 {
  uint32_t L, B, maxp;

  /// Home slot and stored entry of id.
  __device__ __forceinline__ void place(uint32_t id, uint32_t& home, uint32_t& entry) const {
    const uint32_t x = (id * 0x9E3779B1u) & ((1u << B) - 1u);
    const uint32_t S = B - L;
    home = x >> S;
    entry = (((home & 0x7Fu) << S) | (x & ((1u << S) - 1u))) + 1u;
  }
  /// min(a, b) of each 16-bit half (v_pk_min_u16). Written out because the
  /// compiler turns a packed minimum against a constant into per-half compares.
  static __device__ __forceinline__ uint32_t pk_min_u16(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
  /// Bit j set: slot j of the group holds the entry or is empty (a stop).
  /// entry2 = the entry in both halves. A half of min(w ^ entry2, w) is zero
  /// iff the slot holds the entry or is empty; min of that with 1 is 0 for a
  /// stop and 1 otherwise, which is gathered into an 8-bit mask in slot order
  /// (slot 2i: bit 0 of word i, slot 2i + 1: bit 16).
  static __device__ __forceinline__ uint32_t stop_mask(const u32x4 g, uint32_t entry2) {
    const uint32_t wd[4] = {g.x, g.y, g.z, g.w};
    uint32_t z[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) z[i] = pk_min_u16(pk_min_u16(wd[i] ^ entry2, wd[i]), 0x00010001u);
    const uint32_t x = z[0] | (z[1] << 2) | (z[2] << 4) | (z[3] << 6);  // bits 2i and 16 + 2i
    return ~(x | (x >> 15)) & 0xFFu;
  }
  /// Slots of the group that lie in the probe window: j >= o (the current
  /// slot) and at most maxp slots past home (p: probe distance of slot o).
  __device__ __forceinline__ uint32_t window(uint32_t o, uint32_t p) const {
    uint32_t m = (0xFFu << o) & 0xFFu;
    const uint32_t last = maxp - p + o;  // p <= maxp
    if (last < 7u) m &= (2u << last) - 1u;
    return m;
  }
  /// 1: inserted, 0: present, 2: no free slot within maxp probes.
  __device__ __forceinline__ int visit(uint32_t* words, uint32_t id) const {
    uint32_t home, entry;
    place(id, home, entry);
    const uint32_t entry2 = entry | (entry << 16);
    const uint32_t smask = (1u << L) - 1u;
    uint32_t p = 0;  // probe distance of the first slot to examine
    for (;;) {
      const uint32_t slot = (home + p) & smask;
      const uint32_t o = slot & 7u, gi = slot >> 3;
      const u32x4 g = lds_read16(words + 4 * gi);
      const uint32_t cand = stop_mask(g, entry2) & window(o, p);
      if (cand != 0u) {
        const uint32_t j = static_cast<uint32_t>(__builtin_ctz(cand));
        const uint32_t wi = j >> 1, sh = (j & 1u) * 16u;
        const uint32_t cur = wi == 0 ? g.x : (wi == 1 ? g.y : (wi == 2 ? g.z : g.w));
        if (((cur >> sh) & 0xFFFFu) == entry) return 0;
        const uint32_t old = atomicCAS(words + 4 * gi + wi, cur, cur | (entry << sh));  // the half is 0
        if (old == cur) return 1;
        p = p - o + j;  // the word changed under us: look at this slot again
        continue;
      }
      p += 8u - o;
      if (p > maxp) return 2;
    }
  }
  __device__ __forceinline__ bool has(const uint32_t* words, uint32_t id) const {
    uint32_t home, entry;
    place(id, home, entry);
    const uint32_t entry2 = entry | (entry << 16);
    const uint32_t smask = (1u << L) - 1u;
    uint32_t p = 0;
    for (;;) {
      const uint32_t slot = (home + p) & smask;
      const uint32_t o = slot & 7u, gi = slot >> 3;
      const u32x4 g = lds_read16(words + 4 * gi);
      const uint32_t cand = stop_mask(g, entry2) & window(o, p);
      if (cand != 0u) {
        const uint32_t j = static_cast<uint32_t>(__builtin_ctz(cand));
        const uint32_t wi = j >> 1, sh = (j & 1u) * 16u;
        const uint32_t cur = wi == 0 ? g.x : (wi == 1 ? g.y : (wi == 2 ? g.z : g.w));
        return ((cur >> sh) & 0xFFFFu) == entry;
      }
      p += 8u - o;
      if (p > maxp) return false;
    }
  }
};

// Load of the LDS visited hash (in eighths) at which a query's visited set
// moves on to its HBM table. Unfiltered searches fit the LDS table at 3/4 up
// to ef 128 and mostly at 256 (7/8 measured 2-4 % slower from longer probes,
// 5/8 6 % slower at ef 256 from more spills). A sparse filter's search visits
// several times as many nodes and spills anyway: moving earlier, at 5/8,
// spares it the long probes (10 % filter at ef 256: 6 % faster; 4/8 lost).
#ifndef HNSWGPU_LDS_LOAD_EIGHTHS
#define HNSWGPU_LDS_LOAD_EIGHTHS 6
#endif
#ifndef HNSWGPU_FILTERED_LDS_EIGHTHS
#define HNSWGPU_FILTERED_LDS_EIGHTHS 5
#endif

/// Exact visited set of one query: an open-addressing hash in LDS (16-bit
/// entries when ids fit, else 32-bit) and, once that reaches 3/4 load (5/8 in
/// the filtered search), a second one in HBM (one table per resident
/// wavefront, emptied after a query that used it). After the switch new ids
/// go to the HBM table; lookups check LDS first.
struct Visited {
  uint32_t* lds;
  uint32_t lds_log2;  // LDS words
  uint32_t* hbm;  // null: no second level
  uint32_t hbm_log2;
  uint32_t in_lds;
  uint32_t in_hbm;
  bool spilled;
  bool compact;   // 16-bit LDS entries (h16)
  Hash16 h16;
  bool overflow;  // per lane: a compact insert found no slot and went to the HBM table
  bool lost;      // per lane: ... and there is no HBM table

  __device__ __forceinline__ uint32_t lds_slots() const { return 1u << (compact ? lds_log2 + 1 : lds_log2); }
  /// Empty the LDS table (whole wave): 0 in compact slots, kEmptySlot in 32-bit ones.
  __device__ __forceinline__ void clear_lds() {
    const uint32_t e = compact ? 0u : kEmptySlot;
    for (uint32_t i = static_cast<uint32_t>(lane_id()); i < (1u << lds_log2); i += 64) lds[i] = e;
  }
  /// Room for `more` new ids? Switches to the HBM table when the LDS one
  /// reaches EIGHTHS / 8 load; false when both are full.
  template <uint32_t EIGHTHS = HNSWGPU_LDS_LOAD_EIGHTHS>
  __device__ __forceinline__ bool reserve(uint32_t more) {
    if (!spilled) {
      if (in_lds + more <= (lds_slots() >> 3) * EIGHTHS) return true;
      if (hbm == nullptr) return false;
      spilled = true;
    }
    return in_hbm + more <= ((1u << hbm_log2) >> 2) * 3;
  }
  /// Per-lane test-and-set; every lane holding an id calls it together.
  __device__ __forceinline__ bool visit(uint32_t id) {
    if (!spilled) {
      if (!compact) return hash_visit(lds, lds_log2, id);
      const int r = h16.visit(lds, id);
      if (r != 2) return r == 1;
      overflow = true;
      if (hbm == nullptr) {
        lost = true;
        return false;
      }
      return hash_visit(hbm, hbm_log2, id);
    }
    if (compact ? h16.has(lds, id) : hash_has(lds, lds_log2, id)) return false;
    return hash_visit(hbm, hbm_log2, id);
  }
  /// After a batch of visits: a compact-hash overflow moves the set to its
  /// HBM level (the overflowed ids are there). False if an id had nowhere to go.
  __device__ __forceinline__ bool settle() {
    const bool any_overflow = ballot(overflow) != 0, any_lost = ballot(lost) != 0;
    overflow = false;
    lost = false;
    if (any_overflow) spilled = true;
    return !any_lost;
  }
  __device__ __forceinline__ void count(uint32_t fresh) {
    if (spilled) {
      in_hbm += fresh;
    } else {
      in_lds += fresh;
    }
  }
  /// Empty the HBM table if this query used it.
  __device__ __forceinline__ void finish() {
    if (!spilled) return;
    uint4* t = reinterpret_cast<uint4*>(hbm);
    const uint4 e = make_uint4(kEmptySlot, kEmptySlot, kEmptySlot, kEmptySlot);
    for (uint32_t i = static_cast<uint32_t>(lane_id()); i < ((1u << hbm_log2) >> 2); i += 64) t[i] = e;
    __threadfence();
    spilled = false;
  }
};

/// Per-query work counts (distances, layer-0 list reads, upper-layer list reads).
struct Work {
  uint32_t dist = 0, exp0 = 0, expu = 0;
};

/// Greedy descent from `cur` through layers max_level..1. A list's count and
/// entries are loaded together, and each neighbour's list offset is fetched
/// with its row, so a step costs two memory round trips (list, then rows).
template <int PC, int ST, int BG = 4>
__device__ uint64_t descend(const DevIndexView& ix, const WaveLds& w, uint64_t cur, Work& work) {
  const int lane = lane_id();
  if (ix.max_level <= 0) return cur;
  uint32_t cur_off = uniform(ix.upper_off[key_id(cur)]);
  for (int L = ix.max_level; L > 0; --L) {
    for (;;) {
      const uint32_t* l = ix.upper + cur_off + static_cast<size_t>(L - 1) * (1u + ix.M);
      const uint32_t head = l[0];
      const uint32_t nb = static_cast<uint32_t>(lane) < ix.M ? l[1 + lane] : 0u;
      const uint32_t cnt = uniform(head);
      work.dist += cnt;
      ++work.expu;
      wave_sync();
      const bool has = static_cast<uint32_t>(lane) < cnt;
      if (has) w.nid[lane] = nb;
      const uint32_t nb_off = has ? ix.upper_off[nb] : 0u;
      wave_sync();
      row_dists<PC, ST, 8, BG>(ix, w, cnt);
      wave_sync();
      const uint64_t mine = has ? key_of(w.nd[lane], static_cast<uint32_t>(lane)) : kKeyInf;
      const uint64_t best = wave_min64(mine);
      if (best == kKeyInf || key_ord(best) >= key_ord(cur)) break;
      const int at = static_cast<int>(key_id(best));
      cur = key_of(key_ord(best), rdlane(nb, at));
      cur_off = rdlane(nb_off, at);
    }
  }
  return cur;
}

/// Load the prepared query into LDS and compute the entry point distance.
template <int PC, int ST = kStoreF32, int BG = 4>
__device__ uint64_t start_query(const DevIndexView& ix, const SearchScratch& s, const WaveLds& w, uint32_t qi,
                                Work& work) {
  const int lane = lane_id();
  const float* src = s.qprep + static_cast<size_t>(qi) * ix.padded;
  wave_sync();
  float sq = 0.0f;
  for (uint32_t i = static_cast<uint32_t>(lane); i < ix.padded; i += 64) {
    const float x = src[i];
    w.q[i] = x;
    if constexpr (ST == kStoreBF16) {
      const uint32_t b = bf16_bits_rne(x);
      w.qb[i] = static_cast<uint16_t>(b);
      const float xb = bf16_value(b);
      sq = __builtin_fmaf(xb, xb, sq);
    } else if constexpr (ST == kStoreFP8) {
      sq = __builtin_fmaf(x, x, sq);
    }
  }
  if constexpr (ST != kStoreF32) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o);
    if (lane == 0) w.qn[0] = sq;
  }
  if (lane == 0) w.nid[0] = ix.entry;
  wave_sync();
  row_dists<PC, ST, 8, BG>(ix, w, 1);
  wave_sync();
  const uint64_t cur = key_of(uniform(w.nd[0]), ix.entry);
  work.dist += 1;
  return descend<PC, ST, BG>(ix, w, cur, work);
}

/// Re-rank R (sR keys with distances from a reduced copy) with exact fp32
/// distances: R becomes the min(k, sR) exactly smallest keys, sorted, which is
/// all the output takes. XE register rows hold them (64 * XE >= min(k, sR)).
/// Each batch of 64 keys is scored at once and only keys below the current
/// k-th are inserted.
template <int XE, int ER, int PC, int G>
__device__ __forceinline__ WaveSorted<ER> rerank_top(const DevIndexView& ix, const WaveLds& w, WaveSorted<ER> R,
                                                     uint32_t sR, uint32_t k) {
  const int lane = lane_id();
  const uint32_t cap = k < sR ? k : sR;
  WaveSorted<XE> X;
  X.clear();
  uint64_t worst = kKeyInf;  // X's key at cap - 1, kKeyInf until X holds cap keys
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    const uint32_t base = static_cast<uint32_t>(e) * 64u;
    if (base >= sR) break;
    const uint32_t cnt = sR - base < 64u ? sR - base : 64u;
    wave_sync();
    if (static_cast<uint32_t>(lane) < cnt) w.nid[lane] = key_id(R.v[e]);
    wave_sync();
    row_dists<PC, kStoreF32, G>(ix, w, cnt);
    wave_sync();
    uint64_t m = ballot(static_cast<uint32_t>(lane) < cnt && key_of(w.nd[lane], w.nid[lane]) < worst);
    for (; m != 0; m &= m - 1) {
      const uint32_t i = static_cast<uint32_t>(__builtin_ctzll(m));
      const uint64_t x = key_of(uniform(w.nd[i]), uniform(w.nid[i]));
      if (x >= worst) continue;
      X.insert(x, cap);
      worst = X.at(cap - 1);
    }
  }
#pragma unroll
  for (int e = 0; e < ER; ++e) R.v[e] = e < XE ? X.v[e < XE ? e : 0] : kKeyInf;
  return R;
}

/// The same result in place, for the largest capacity class, where a second
/// array of register rows would cost occupancy: each key's distance is
/// replaced by its exact one, 64 rows at a time, and R is sorted, so its first
/// min(k, sR) entries are the exactly smallest keys in (distance, id) order.
/// Never inlined, so the sort's registers are not added to the search loop's.
template <int ER, int PC, int G>
__device__ __attribute__((noinline)) WaveSorted<ER> rerank_sorted(const DevIndexView ix, const WaveLds w,
                                                                  WaveSorted<ER> R, uint32_t sR) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    const uint32_t base = static_cast<uint32_t>(e) * 64u;
    if (base < sR) {
      const uint32_t cnt = sR - base < 64u ? sR - base : 64u;
      wave_sync();
      if (lane < cnt) w.nid[lane] = key_id(R.v[e]);
      wave_sync();
      row_dists<PC, kStoreF32, G>(ix, w, cnt);
      wave_sync();
      R.v[e] = lane < cnt ? key_of(w.nd[lane], w.nid[lane]) : kKeyInf;
    }
  }
  R.sort_all();
  return R;
}

/// Everything is taken (and R returned) by value: a reference would keep the
/// caller's R, index view and LDS view in scratch memory for the whole search
/// wherever this is not inlined (the largest capacity class), not just around
/// the call.
template <int ER, int PC, int G>
__device__ WaveSorted<ER> rerank_exact(const DevIndexView ix, const WaveLds w, WaveSorted<ER> R, uint32_t sR,
                                       uint32_t k) {
  if constexpr (ER >= 8) {
    return rerank_sorted<ER, PC, G>(ix, w, R, sR);
  } else {
    return rerank_top<ER, ER, PC, G>(ix, w, R, sR, k);
  }
}

/// A node's layer-0 list in registers: lane j holds entry j + 1, `cnt` the
/// length. One load per lane, so it can be issued ahead of its use.
struct ListRegs {
  uint32_t id;
  uint32_t cnt;
  uint32_t nb;
};

__device__ __forceinline__ ListRegs load_list0(const DevIndexView& ix, uint32_t id) {
  const int lane = lane_id();
  const uint32_t* l = ix.l0 + static_cast<size_t>(id) * (1u + ix.M0);
  ListRegs r;
  r.id = id;
  r.cnt = l[0];
  r.nb = static_cast<uint32_t>(lane) < ix.M0 ? l[1 + lane] : 0u;
  return r;
}

/// Neighbours in `lr` that are new to the visited set, in list order: ids into
/// w.nid, eligibility into w.nel. Returns their number; the test-and-set is
/// `visit(id)`, called by all lanes holding an id at once.
template <typename Visit>
__device__ __forceinline__ uint32_t fresh_from_list(const DevIndexView& ix, const WaveLds& w, const ListRegs& lr,
                                                    const uint32_t* elig, Visit visit) {
  const int lane = lane_id();
  const uint32_t cnt = uniform(lr.cnt);
  bool has = static_cast<uint32_t>(lane) < cnt;
  const uint32_t nb = lr.nb;
  if (ix.has_dups) {
    // A repeated id counts as visited at its later positions.
    for (uint32_t i = 0; i + 1 < cnt; ++i) {
      const uint32_t vi = rdlane(nb, static_cast<int>(i));
      if (has && static_cast<uint32_t>(lane) > i && nb == vi) has = false;
    }
  }
  // Eligibility words are loaded for the whole list before the visited probe,
  // so their latency runs under the probe instead of after it.
  uint32_t ew = 0;
  if (elig != nullptr && has) ew = reinterpret_cast<const global_u32*>(reinterpret_cast<uintptr_t>(elig))[nb >> 5];
  const bool fresh = has && visit(nb);
  const uint64_t fm = ballot(fresh);
  wave_sync();
  if (fresh) {
    const uint32_t p = lanes_below(fm);
    w.nid[p] = nb;
    if (elig != nullptr) w.nel[p] = (ew >> (nb & 31)) & 1u;
  }
  wave_sync();
  return static_cast<uint32_t>(__popcll(fm));
}

template <typename Visit>
__device__ __forceinline__ uint32_t fresh_neighbours(const DevIndexView& ix, const WaveLds& w, uint32_t cid,
                                                     const uint32_t* elig, Visit visit) {
  return fresh_from_list(ix, w, load_list0(ix, cid), elig, visit);
}

__device__ __forceinline__ void write_pad(OutRow* row, uint32_t from, uint32_t k) {
  for (uint32_t j = from + static_cast<uint32_t>(lane_id()); j < k; j += 64) {
    row[j] = OutRow{~0ULL, __uint_as_float(0x7F800000u), 0u};
  }
}

/// The parameters of query qi: from its task in a mixed launch, else from the
/// request. Wave-uniform.
struct QueryArgs {
  const float* query;
  OutRow* out;
  uint32_t* count;
  const uint32_t* elig;
  uint32_t* stats;
  uint32_t k;
  uint32_t ef;
  float scale;     // cosine: 1/|q| from the host (has_scale)
  bool has_scale;
};

template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  return reinterpret_cast<T*>(uniform64(reinterpret_cast<uint64_t>(p)));
}

__device__ __forceinline__ QueryArgs query_args(const SearchRequest& r, uint32_t qi) {
  QueryArgs a;
  if (r.tasks != nullptr) {
    const QueryTask& t = r.tasks[qi];
    a.query = uniform_ptr(t.query);
    a.out = uniform_ptr(t.out);
    a.count = uniform_ptr(t.count);
    a.elig = uniform_ptr(t.elig);
    a.k = uniform(t.k);
    a.ef = uniform(t.ef);
    a.scale = __uint_as_float(uniform(__float_as_uint(t.scale)));
    a.has_scale = uniform(t.has_scale) != 0u;
  } else {
    a.query = r.queries + static_cast<size_t>(qi) * r.dim;
    a.out = r.out + static_cast<size_t>(qi) * r.k;
    a.count = r.counts != nullptr ? r.counts + qi : nullptr;
    a.elig = r.elig;
    a.k = r.k;
    a.ef = r.ef;
    a.has_scale = r.qscale != nullptr;
    a.scale = a.has_scale ? r.qscale[qi] : 1.0f;
  }
  a.stats = r.stats != nullptr ? r.stats + 3 * static_cast<size_t>(qi) : nullptr;
  return a;
}

__device__ __forceinline__ uint32_t launch_nq(const SearchRequest& r) {
  return r.nq_dev != nullptr ? uniform(*r.nq_dev) : r.nq;
}

/// Why the fast kernel handed a query to the general kernel.
enum class Handoff : uint32_t { kNone = 0, kCandidates = 4, kVisited = 5 };

/// Results of a finished query: R's first min(k, |R|) entries with labels,
/// padding, result count and work counts.
template <int ER>
__device__ void write_results(const DevIndexView& ix, const QueryArgs& a, const WaveSorted<ER>& R, uint32_t sR,
                              const Work& work) {
  const int lane = lane_id();
  const uint32_t kout = sR < a.k ? sR : a.k;
  OutRow* row = a.out;
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    const uint32_t idx = static_cast<uint32_t>(e * 64 + lane);
    if (idx < kout) {
      const uint64_t key = R.v[e];
      row[idx] = OutRow{ix.labels[key_id(key)], float_of(key_ord(key)), 0u};
    }
  }
  write_pad(row, kout, a.k);
  if (a.count != nullptr && lane == 0) *a.count = kout;
  if (a.stats != nullptr && lane == 0) {
    a.stats[0] = work.dist;
    a.stats[1] = work.exp0;
    a.stats[2] = work.expu;
  }
}

/// The result slots not yet expanded, as one ballot mask per register row:
/// bit l of m[e] is slot 64e + l. Empty slots carry flag 1, so they never
/// appear. Computed once per expansion; the pop and the peek both read it.
template <int ER>
struct Unexpanded {
  uint64_t m[ER];

  __device__ __forceinline__ void load(const uint32_t (&fl)[ER]) {
#pragma unroll
    for (int e = 0; e < ER; ++e) m[e] = ballot(fl[e] == 0u);
  }
  /// Index of the first unexpanded slot, or ~0 if none.
  __device__ __forceinline__ uint32_t first() const {
    uint32_t slot = ~0u;
#pragma unroll
    for (int e = ER - 1; e >= 0; --e) {
      if (m[e] != 0) slot = static_cast<uint32_t>(e * 64 + __builtin_ctzll(m[e]));
    }
    return slot;
  }
  /// The first unexpanded slot after `slot`, which must be first() (the
  /// lowest bit of its row, with no bit in the rows before). Read-only, so the
  /// masks stay in scalar registers.
  __device__ __forceinline__ uint32_t after(uint32_t slot) const {
    uint32_t r = ~0u;
#pragma unroll
    for (int e = ER - 1; e >= 0; --e) {
      uint64_t mm = m[e];
      if (static_cast<uint32_t>(e) == (slot >> 6)) mm &= mm - 1;
      if (mm != 0) r = static_cast<uint32_t>(e * 64 + __builtin_ctzll(mm));
    }
    return r;
  }
};

/// Admission of a whole expansion at once while R is full (|R| = ef): R
/// becomes the ef smallest keys of R and of the candidates (lanes of `smask`,
/// keys `skey`, all below tau), flags moving with their keys (candidates
/// unexpanded). That is the sequential admission's result unless a candidate,
/// or a key pushed out of R, has the new worst distance: then the arrival order
/// decides ties, nothing is changed and false is returned. On success tau is
/// the new worst distance, which is below the old one (keys were pushed out),
/// so the caller's list of ties is empty.
template <int ER>
__device__ bool admit_merged(WaveSorted<ER>& R, uint32_t (&fl)[ER], uint32_t ef, uint32_t& tau, uint64_t smask,
                             uint64_t skey) {
  const int lane = lane_id();
  const bool in_s = ((smask >> lane) & 1u) != 0;
  // Per candidate j (uniform in its iteration): merged position p_j = keys
  // below it in R and among the candidates. Output slot k holds the candidate
  // with p_j = k, else R[k - (candidates placed before k)]. Candidates placed
  // in rows before row e count for every lane of row e (a uniform count per
  // row); only the row holding p_j needs per-lane work.
  uint32_t before[ER], rows_before[ER];
  uint64_t nv[ER];
  bool from_s[ER];
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    before[e] = 0;
    rows_before[e] = 0;
    nv[e] = kKeyInf;
    from_s[e] = false;
  }
  uint32_t s_before_ef = 0;
  uint64_t s_at_ef = kKeyInf;
  for (uint64_t m = smask; m != 0; m &= m - 1) {
    const int j = __builtin_ctzll(m);
    const uint64_t sj = rdlane64(skey, j);
    const uint32_t pj = R.rank(sj) + static_cast<uint32_t>(__popcll(ballot(in_s && skey < sj)));
    s_before_ef += pj < ef ? 1u : 0u;
    if (pj == ef) s_at_ef = sj;
    const uint32_t pr = pj >> 6, pl = pj & 63u;
#pragma unroll
    for (int e = 0; e < ER; ++e) {
      if (static_cast<uint32_t>(e) > pr) {
        ++rows_before[e];
      } else if (static_cast<uint32_t>(e) == pr) {
        before[e] += static_cast<uint32_t>(lane) > pl ? 1u : 0u;
        if (static_cast<uint32_t>(lane) == pl) {
          from_s[e] = true;
          nv[e] = sj;
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < ER; ++e) before[e] += rows_before[e];
  // The key after the last one kept (merged position ef) must be beyond the new worst distance.
  const uint64_t first_out = s_at_ef != kKeyInf ? s_at_ef : R.at(ef - s_before_ef);
  uint32_t nf[ER];
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    const uint32_t k = static_cast<uint32_t>(e * 64 + lane);
    // At most 64 candidates: the source is in row e or row e - 1.
    const uint32_t src = from_s[e] ? k : k - before[e];
    const int sl = static_cast<int>(src & 63u);
    uint64_t rv = from_lane64(R.v[e], sl);
    uint32_t rf = from_lane(fl[e], sl);
    if (e > 0) {
      const uint64_t pv = from_lane64(R.v[e > 0 ? e - 1 : 0], sl);
      const uint32_t pf = from_lane(fl[e > 0 ? e - 1 : 0], sl);
      if ((src >> 6) != static_cast<uint32_t>(e)) {
        rv = pv;
        rf = pf;
      }
    }
    nf[e] = from_s[e] ? 0u : rf;
    if (!from_s[e]) nv[e] = rv;
    if (k >= ef) {
      nv[e] = kKeyInf;
      nf[e] = 1u;
    }
  }
  uint64_t worst = kKeyInf;
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    if (static_cast<uint32_t>(e) == ((ef - 1) >> 6)) worst = rdlane64(nv[e], static_cast<int>((ef - 1) & 63));
  }
  const uint32_t t = key_ord(worst);
  if (key_ord(first_out) <= t || ballot(in_s && key_ord(skey) == t) != 0) return false;
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    R.v[e] = nv[e];
    fl[e] = nf[e];
  }
  tau = t;
  return true;
}

/// Layer-0 search without filter or deletions. Here every admitted node joins
/// both C and R, so C needs no array of its own: its live entries are the
/// entries of R not yet expanded (a flag per slot) plus the entries pushed out
/// of R whose distance ties the new worst result (`ties`, lane i holding the
/// i-th smallest id; all have distance tau). Every entry of R orders before
/// every tie, and an entry pushed out with a larger distance can never be
/// expanded again, so popping "first unexpanded slot of R, else the smallest
/// tie" is the exact Alg. 2 order, and the loop ends exactly when Alg. 2 would
/// stop (the smallest remaining candidate is further than the worst result).
///
/// DEL: the eligibility bitmap a.elig excludes a few nodes (soft deletions). An
/// admitted ineligible node joins C only: it goes to a sorted side list D of up
/// to 64 keys, and the next candidate is the smallest of R's first unexpanded
/// slot, the smallest tie and D's head. An entry of D further than the worst
/// result of a full R is dead (never expanded); D's head being dead ends the
/// search, as Alg. 2's stop rule does. A live key pushed out of a full D hands
/// the query to the general kernel.
#ifndef HNSWGPU_MERGE_MAX_ER
#define HNSWGPU_MERGE_MAX_ER 8
#endif
template <int ER, int PC, int ST, bool DEL = false, int RT = 1>
__device__ Handoff plain_query(const DevIndexView& ix, const QueryArgs& a, const SearchScratch& s,
                               const WaveLds& w, Visited& vis, uint32_t qi) {
  const int lane = lane_id();
  const uint32_t ef = a.ef;

  vis.clear_lds();
  vis.in_lds = 0;
  vis.in_hbm = 0;
  Work work;
  const uint64_t ep = start_query<PC, ST, 4>(ix, s, w, qi, work);

  WaveSorted<ER> R;
  R.clear();
  uint32_t fl[ER];  // 0: an entry not yet expanded; 1: expanded, or an empty slot
#pragma unroll
  for (int e = 0; e < ER; ++e) fl[e] = 1u;
  uint32_t ties = 0u, n_ties = 0u;
  WaveSorted<1> D;
  D.clear();
  wave_sync();
  if (lane == 0) (void)vis.visit(key_id(ep));
  vis.in_lds = 1;
  wave_sync();
  uint32_t sR = 0, tau = kOrdInf;
  if (!DEL || elig_bit(a.elig, key_id(ep))) {
    uint32_t dummy;
    R.insert_flagged(ep, ef, fl, 0u, dummy);
    sR = 1;
    tau = key_ord(ep);
  } else {
    D.insert(ep, 64u);
  }
  ListRegs next = load_list0(ix, key_id(ep));

  Handoff result = Handoff::kNone;
  for (;;) {
    uint64_t c;
    Unexpanded<ER> ux;
    ux.load(fl);
    const uint32_t slot = ux.first();
    bool popped = false;  // c is R's slot `slot`
    const uint64_t dmin = DEL ? rdlane64(D.v[0], 0) : kKeyInf;
    if (DEL && dmin < (slot != ~0u ? R.at(slot) : n_ties > 0 ? key_of(tau, rdlane(ties, 0)) : kKeyInf)) {
      c = dmin;
      if (sR >= ef && key_ord(c) > tau) break;
      D.pop_front();
    } else if (slot != ~0u) {
      c = R.at(slot);
#pragma unroll
      for (int e = 0; e < ER; ++e) {
        if (static_cast<uint32_t>(e) == (slot >> 6) && static_cast<uint32_t>(lane) == (slot & 63)) fl[e] = 1u;
      }
      popped = true;
    } else if (n_ties > 0) {
      c = key_of(tau, rdlane(ties, 0));
      ties = wave_shl1(ties, rdlane(ties, 63));
      --n_ties;
    } else {
      break;
    }
    if (!vis.reserve(ix.M0)) {
      result = Handoff::kVisited;
      break;
    }
    const uint32_t cid = key_id(c);
    const ListRegs cur = uniform(next.id) == cid ? next : load_list0(ix, cid);
    const uint32_t peek = popped ? ux.after(slot) : slot;
    uint64_t peek_key = peek != ~0u ? R.at(peek) : kKeyInf;
    if constexpr (DEL) {
      const uint64_t d0 = rdlane64(D.v[0], 0);
      peek_key = d0 < peek_key ? d0 : peek_key;
    }
    if (peek_key != kKeyInf) next = load_list0(ix, key_id(peek_key));

    const uint32_t U =
        fresh_from_list(ix, w, cur, DEL ? a.elig : nullptr, [&](uint32_t id) { return vis.visit(id); });
    if (vis.compact && !vis.settle()) {
      result = Handoff::kVisited;
      break;
    }
    vis.count(U);
    work.dist += U;
    ++work.exp0;
    row_dists<PC, ST, 8, 4, RT>(ix, w, U);
    wave_sync();
    const uint32_t od_u = static_cast<uint32_t>(lane) < U ? w.nd[lane] : kOrdInf;
    const uint32_t id_u = w.nid[lane];
    const uint32_t el_u = DEL ? w.nel[lane] : 1u;
    uint64_t cand = ballot(static_cast<uint32_t>(lane) < U && (sR < ef || od_u < tau));
    // Two or more candidates into a full R: one merge, unless ties at the new
    // worst distance need the sequential order. (At ef <= 64 the merge's
    // registers would cost a wavefront per SIMD: the sequential loop is kept.)
    // With deletions only when every candidate is eligible.
    if constexpr (ER >= 2 && ER <= HNSWGPU_MERGE_MAX_ER) {
      if (sR == ef && (cand & (cand - 1)) != 0 && (!DEL || (cand & ~ballot(el_u != 0u)) == 0) &&
          admit_merged<ER>(R, fl, ef, tau, cand, key_of(od_u, id_u))) {
        n_ties = 0;
        cand = 0;
      }
    }
    while (cand != 0) {
      const int i = __builtin_ctzll(cand);
      cand &= cand - 1;
      const uint32_t od = rdlane(od_u, i);
      if (sR >= ef && od >= tau) continue;
      const uint64_t key = key_of(od, rdlane(id_u, i));
      if constexpr (DEL) {
        if (rdlane(el_u, i) == 0u) {
          const uint64_t dout = D.insert(key, 64u);
          if (dout != kKeyInf && (sR < ef || key_ord(dout) <= tau)) {
            result = Handoff::kCandidates;
            break;
          }
          continue;
        }
      }
      uint32_t out_flag;
      const uint64_t out = R.insert_flagged(key, ef, fl, 0u, out_flag);
      if (sR < ef) {
        ++sR;
        tau = key_ord(R.at(sR - 1));
        continue;
      }
      tau = key_ord(R.at(ef - 1));
      if (key_ord(out) > tau) {
        n_ties = 0;  // the old ties had the old, larger worst distance
      } else if (out_flag == 0u) {
        if (n_ties == 64) {
          result = Handoff::kCandidates;
          break;
        }
        const uint32_t tid = key_id(out);
        const uint32_t pos =
            static_cast<uint32_t>(__popcll(ballot(static_cast<uint32_t>(lane) < n_ties && ties < tid)));
        const uint32_t up = wave_shr1(ties, rdlane(ties, 0));
        if (static_cast<uint32_t>(lane) > pos) ties = up;
        if (static_cast<uint32_t>(lane) == pos) ties = tid;
        ++n_ties;
      }
    }
    if (result != Handoff::kNone) break;
  }
  if (vis.spilled && lane == 0) atomicAdd(&s.counters[3], 1u);
  vis.finish();
  if (result != Handoff::kNone) return result;
  if constexpr (ST != kStoreF32) R = rerank_exact<ER, PC, (ER >= 4 && ST == kStoreBF16 ? 4 : 8)>(ix, w, R, sR, a.k);
  write_results<ER>(ix, a, R, sR, work);
  return Handoff::kNone;
}

/// Ascending sort of 64 keys, one per lane (bitonic network over the wave).
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t o = from_lane64(v, lane ^ j);
      const bool up = (lane & k) == 0;
      const bool low = (lane & j) == 0;
      v = (low == up) ? (o < v ? o : v) : (o > v ? o : v);
    }
  }
  return v;
}

/// Candidates the register array C could not hold: an unsorted row of up to
/// 64 (entry i in lane i) and up to kRunSlots sorted runs of 64 in this
/// wavefront's HBM region, each consumed from a head position. C and the
/// overflow together hold every admitted candidate not yet expanded, except
/// runs found dead (head distance above the worst result of a full R: the
/// search stops before reaching them), whose slots are reused.
struct Overflow {
  uint64_t* runs;   // kRunSlots x 64 keys; null = no overflow (hand the query off instead)
  uint64_t buf;
  uint32_t nbuf;
  uint64_t bufmin;
  uint64_t head[2];  // lane l: head key of runs l and 64 + l (kKeyInf = free)
  uint32_t pos[2];
  uint64_t runmin;

  __device__ __forceinline__ void clear() {
    buf = kKeyInf;
    nbuf = 0;
    bufmin = kKeyInf;
    head[0] = head[1] = kKeyInf;
    pos[0] = pos[1] = 0;
    runmin = kKeyInf;
  }
  __device__ __forceinline__ uint64_t min() const { return bufmin < runmin ? bufmin : runmin; }
  __device__ __forceinline__ void update_runmin() { runmin = wave_min64(head[0] < head[1] ? head[0] : head[1]); }

  /// Sort the row and store it as a run. False when every slot is taken.
  __device__ bool spill_row(bool r_full, uint32_t tau) {
    const int lane = lane_id();
    const uint64_t v = wave_sort64(static_cast<uint32_t>(lane) < nbuf ? buf : kKeyInf);
    const uint64_t f0 = ballot(head[0] == kKeyInf || (r_full && key_ord(head[0]) > tau));
    const uint64_t f1 = ballot(head[1] == kKeyInf || (r_full && key_ord(head[1]) > tau));
    if (f0 == 0 && f1 == 0) return false;
    const uint32_t slot = f0 != 0 ? static_cast<uint32_t>(__builtin_ctzll(f0))
                                  : 64u + static_cast<uint32_t>(__builtin_ctzll(f1));
    runs[static_cast<size_t>(slot) * 64 + lane] = v;
    wave_sync();
    const uint64_t h = rdlane64(v, 0);
    if (static_cast<uint32_t>(lane) == (slot & 63)) {
      if (slot < 64) {
        head[0] = h;
        pos[0] = 0;
      } else {
        head[1] = h;
        pos[1] = 0;
      }
    }
    buf = kKeyInf;
    nbuf = 0;
    bufmin = kKeyInf;
    update_runmin();
    return true;
  }
  /// Append one key (uniform). False when the overflow is full.
  __device__ __forceinline__ bool push(uint64_t x, bool r_full, uint32_t tau) {
    if (static_cast<uint32_t>(lane_id()) == nbuf) buf = x;
    ++nbuf;
    bufmin = x < bufmin ? x : bufmin;
    if (nbuf == 64) return spill_row(r_full, tau);
    return true;
  }
  /// The overflow's smallest key precedes C's: move back into C, from the run
  /// holding it, every key below the smallest key held elsewhere (at least
  /// that one). Keys C pushes out go back to the overflow. False when full.
  template <int EC>
  __device__ bool refill(WaveSorted<EC>& C, uint32_t ccap, uint32_t& sC, bool r_full, uint32_t tau) {
    const int lane = lane_id();
    if (nbuf > 0 && bufmin < runmin && !spill_row(r_full, tau)) return false;
    const uint64_t m0 = ballot(head[0] == runmin), m1 = ballot(head[1] == runmin);
    const uint32_t slot = m0 != 0 ? static_cast<uint32_t>(__builtin_ctzll(m0))
                                  : 64u + static_cast<uint32_t>(__builtin_ctzll(m1));
    const bool mine = static_cast<uint32_t>(lane) == (slot & 63);
    const uint64_t o0 = (mine && slot < 64) ? kKeyInf : head[0];
    const uint64_t o1 = (mine && slot >= 64) ? kKeyInf : head[1];
    const uint64_t other = wave_min64(o0 < o1 ? o0 : o1);
    const uint64_t bound = other < bufmin ? other : bufmin;
    const uint32_t p = rdlane(slot < 64 ? pos[0] : pos[1], static_cast<int>(slot & 63));
    const uint64_t v = runs[static_cast<size_t>(slot) * 64 + lane];
    const uint32_t cnt = static_cast<uint32_t>(__popcll(ballot(static_cast<uint32_t>(lane) >= p && v < bound)));
    const uint32_t np = p + cnt;
    const uint64_t nh = np < 64 ? rdlane64(v, static_cast<int>(np)) : kKeyInf;
    if (mine) {
      if (slot < 64) {
        head[0] = nh;
        pos[0] = np;
      } else {
        head[1] = nh;
        pos[1] = np;
      }
    }
    update_runmin();
    for (uint32_t i = 0; i < cnt; ++i) {
      const uint64_t d = C.insert(rdlane64(v, static_cast<int>(p + i)), ccap);
      if (d == kKeyInf) {
        ++sC;
      } else if (!push(d, r_full, tau)) {
        return false;
      }
    }
    return true;
  }
};

/// Layer-0 search with an eligibility bitmap (filter and/or deletions): C and R
/// are separate sorted arrays; C keeps its 64 * EC smallest entries in
/// registers and the rest in the overflow, from which entries come back before
/// they are needed, so candidates are expanded in the exact (distance, id)
/// order. Without an overflow region, or when it fills, the query goes to the
/// general kernel.
template <int ER, int EC, int PC, int ST, int RT = 1>
__device__ Handoff filtered_query(const DevIndexView& ix, const QueryArgs& a, const SearchScratch& s,
                                  const WaveLds& w, Visited& vis, uint32_t qi, uint64_t* runs) {
  const int lane = lane_id();
  const uint32_t ccap = 64u * EC;
  const uint32_t ef = a.ef;

  vis.clear_lds();
  vis.in_lds = 0;
  vis.in_hbm = 0;
  Work work;
  const uint64_t ep = start_query<PC, ST, 4>(ix, s, w, qi, work);

  WaveSorted<ER> R;
  WaveSorted<EC> C;
  R.clear();
  C.clear();
  uint32_t sR = 0, sC = 0, tau = kOrdInf;
  Overflow ov;
  ov.runs = runs;
  ov.clear();
  wave_sync();
  if (lane == 0) (void)vis.visit(key_id(ep));
  vis.in_lds = 1;
  wave_sync();
  C.insert(ep, ccap);
  sC = 1;
  if (elig_bit(a.elig, key_id(ep))) {
    R.insert(ep, ef);
    sR = 1;
    tau = key_ord(ep);
  }
  ListRegs next = load_list0(ix, key_id(ep));

  Handoff result = Handoff::kNone;
  for (;;) {
    const uint64_t c = C.at(0);
    const uint64_t omin = ov.min();
    if (omin < c) {
      if (sR >= ef && key_ord(omin) > tau) break;
      if (runs == nullptr || !ov.refill<EC>(C, ccap, sC, sR >= ef, tau)) {
        result = Handoff::kCandidates;
        break;
      }
      continue;
    }
    if (sC == 0) break;
    if (sR >= ef && key_ord(c) > tau) break;
    C.pop_front();
    --sC;
    if (!vis.reserve<HNSWGPU_FILTERED_LDS_EIGHTHS>(ix.M0)) {
      result = Handoff::kVisited;
      break;
    }
    const uint32_t cid = key_id(c);
    const ListRegs cur = uniform(next.id) == cid ? next : load_list0(ix, cid);
    const uint64_t peek = C.at(0);
    if (peek != kKeyInf) next = load_list0(ix, key_id(peek));

    const uint32_t U = fresh_from_list(ix, w, cur, a.elig, [&](uint32_t id) { return vis.visit(id); });
    if (vis.compact && !vis.settle()) {
      result = Handoff::kVisited;
      break;
    }
    vis.count(U);
    work.dist += U;
    ++work.exp0;
    row_dists<PC, ST, 8, 4, RT>(ix, w, U);
    wave_sync();
    const uint32_t od_u = static_cast<uint32_t>(lane) < U ? w.nd[lane] : kOrdInf;
    const uint32_t id_u = w.nid[lane];
    const uint32_t el_u = w.nel[lane];
    uint64_t cand = ballot(static_cast<uint32_t>(lane) < U && (sR < ef || od_u < tau));
    while (cand != 0) {
      const int i = __builtin_ctzll(cand);
      cand &= cand - 1;
      const uint32_t od = rdlane(od_u, i);
      if (sR >= ef && od >= tau) continue;
      const uint64_t key = key_of(od, rdlane(id_u, i));
      if (rdlane(el_u, i) != 0u) {
        R.insert(key, ef);
        if (sR < ef) ++sR;
        tau = key_ord(R.at(sR - 1));
      }
      const uint64_t dropped = C.insert(key, ccap);
      if (dropped == kKeyInf) {
        ++sC;
      } else if (runs != nullptr) {
        if (!ov.push(dropped, sR >= ef, tau)) {
          result = Handoff::kCandidates;
          break;
        }
      } else {
        // No overflow: remember that a candidate was lost; exact while it is never needed.
        const uint64_t b = ov.bufmin;
        ov.bufmin = dropped < b ? dropped : b;
      }
    }
    if (result != Handoff::kNone) break;
  }
  if (vis.spilled && lane == 0) atomicAdd(&s.counters[3], 1u);
  vis.finish();
  if (result != Handoff::kNone) return result;
  if constexpr (ST != kStoreF32) R = rerank_exact<ER, PC, (ER >= 4 && ST == kStoreBF16 ? 4 : 8)>(ix, w, R, sR, a.k);
  write_results<ER>(ix, a, R, sR, work);
  return Handoff::kNone;
}

/// Filter modes of the fast kernel: every query unfiltered, every query
/// filtered, per query (mixed launches), or every query with an eligibility
/// bitmap that excludes only a few deleted nodes (one launch, or a mixed one).
enum FilterMode : int { kPlain = 0, kFiltered = 1, kPerQuery = 2, kDeleted = 3, kPerQueryDeleted = 4 };

/// One query of a uniform launch (one filter mode and capacity class).
template <int ER, int EC, int PC, int FM, int ST, int RT>
__device__ Handoff run_query(const DevIndexView& ix, const QueryArgs& a, const SearchScratch& s, const WaveLds& w,
                             Visited& vis, uint32_t qi, uint64_t* runs) {
  if constexpr (FM == kPlain) {
    return plain_query<ER, PC, ST, false, RT>(ix, a, s, w, vis, qi);
  } else if constexpr (FM == kFiltered) {
    return filtered_query<ER, EC, PC, ST, RT>(ix, a, s, w, vis, qi, runs);
  } else {
    return plain_query<ER, PC, ST, true, RT>(ix, a, s, w, vis, qi);
  }
}

/// One query of a mixed launch (kPerQuery, kPerQueryDeleted): the kernel is
/// built for the largest capacity class of the batch and a query with a
/// smaller ef runs the code of its own class. Every
/// class is inlined into the kernel: a call would take the index view,
/// query arguments, scratch and LDS views and visited state by address, which
/// keeps them in scratch memory, read from the search loops of every class.
template <int ER, int EC, int PC, int FM, int ST, int RT>
__device__ __forceinline__ Handoff run_mixed(const DevIndexView& ix, const QueryArgs& a, const SearchScratch& s,
                                             const WaveLds& w, Visited& vis, uint32_t qi, uint64_t* runs) {
  if constexpr (ER > 1) {
    if (a.ef <= 32u * ER) return run_mixed<ER / 2, EC / 2, PC, FM, ST, RT>(ix, a, s, w, vis, qi, runs);
  }
  if constexpr (FM == kPerQueryDeleted) {
    [[clang::always_inline]] return plain_query<ER, PC, ST, true, RT>(ix, a, s, w, vis, qi);
  } else {
    if (a.elig == nullptr) {
      [[clang::always_inline]] return plain_query<ER, PC, ST>(ix, a, s, w, vis, qi);
    }
    [[clang::always_inline]] return filtered_query<ER, EC, PC, ST>(ix, a, s, w, vis, qi, runs);
  }
}

template <int ER, int EC, int PC, int FM, int ST, int RT>
__device__ __forceinline__ void fast_search_body(uint32_t* lds, const DevIndexView& ix, const SearchRequest& r,
                                                 const SearchScratch& s) {
  const int lane = lane_id();
  const uint32_t per_wave = lds_words_per_wave(r.hash_log2, ix.padded);
  uint32_t* base = lds + static_cast<size_t>(threadIdx.x >> 6) * per_wave;
  const WaveLds w = wave_lds(base, r.hash_log2, ix.padded);
  const size_t wave_slot = static_cast<size_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  // 16-bit LDS entries when an entry can hold the id bits beyond the home
  // slot (B - L <= 8) beside 7 bits of the home slot; else 32-bit entries.
  const uint32_t L = r.hash_log2 + 1u;
  uint32_t B = ix.count > 1u ? 32u - static_cast<uint32_t>(__builtin_clz(ix.count - 1u)) : 1u;
  B = B < L ? L : B;
  const bool compact = r.wide_hash == 0 && B <= L + 8u;
  Visited vis{w.hash, r.hash_log2,
              s.visit_tables != nullptr ? s.visit_tables + (wave_slot << s.visit_log2) : nullptr,
              s.visit_log2, 0u, 0u, false, compact, Hash16{L, B, 127u}, false, false};
  uint64_t* runs = s.cand_runs != nullptr ? s.cand_runs + wave_slot * (static_cast<size_t>(kRunSlots) * 64) : nullptr;
  const uint32_t nq = launch_nq(r);
  for (;;) {
    uint32_t qi = 0;
    if (lane == 0) qi = atomicAdd(&s.counters[0], 1u);
    qi = uniform(qi);
    if (qi >= nq) break;
    Handoff h = Handoff::kCandidates;
    if (!r.force_spill) {
      const QueryArgs a = query_args(r, qi);
      // A task whose ef exceeds the register capacity goes to the general kernel.
      if (a.ef <= 64u * ER) {
        if constexpr (FM == kPerQuery || FM == kPerQueryDeleted) {
          h = run_mixed<ER, EC, PC, FM, ST, RT>(ix, a, s, w, vis, qi, runs);
        } else {
          h = run_query<ER, EC, PC, FM, ST, RT>(ix, a, s, w, vis, qi, runs);
        }
      }
    }
    if (h != Handoff::kNone && lane == 0) {
      const uint32_t slot = atomicAdd(&s.counters[1], 1u);
      s.overflow[slot] = qi;
      if (!r.force_spill) atomicAdd(&s.counters[static_cast<uint32_t>(h)], 1u);
    }
  }
}


template <int ER, int EC, int PC, int FM, int ST, int RT>
__global__ void __launch_bounds__(kWavesPerBlock * 64) fast_search(DevIndexView ix, SearchRequest r,
                                                                   SearchScratch s) {
  extern __shared__ __align__(16) uint32_t lds[];
  fast_search_body<ER, EC, PC, FM, ST, RT>(lds, ix, r, s);
}

/// The same kernel for filtered launches up to ef 256. The LDS already caps
/// every fast kernel at 4 wavefronts per SIMD; holding the compiler to that
/// (128 VGPRs, a few bytes of spill) gives these kernels their 4th wavefront,
/// which their candidate array otherwise costs (131-153 VGPRs): 10 % and 50 %
/// filters 5-12 % faster. The unfiltered and deletion kernels already fit and
/// only lose from the tighter schedule (3 %); the largest class would spill
/// hundreds of bytes.
template <int ER, int EC, int PC, int FM, int ST, int RT>
__global__ void __launch_bounds__(kWavesPerBlock * 64) __attribute__((amdgpu_waves_per_eu(4, 4)))
fast_search_w4(DevIndexView ix, SearchRequest r, SearchScratch s) {
  extern __shared__ __align__(16) uint32_t lds[];
  fast_search_body<ER, EC, PC, FM, ST, RT>(lds, ix, r, s);
}

template <int ER, int EC, int PC, int FM, int ST, int RT>
hipError_t run_fast(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st) {
  const size_t lds = static_cast<size_t>(r.waves_per_block) * lds_words_per_wave(r.hash_log2, ix.padded) * 4u;
  void (*kernel)(DevIndexView, SearchRequest, SearchScratch);
  if constexpr (FM == kFiltered && ER <= 4) {
    kernel = &fast_search_w4<ER, EC, PC, FM, ST, RT>;
  } else {
    kernel = &fast_search<ER, EC, PC, FM, ST, RT>;
  }
  static bool configured = false;
  if (!configured) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024);
    if (e != hipSuccess) return e;
    configured = true;
  }
  hipLaunchKernelGGL(kernel, dim3(r.blocks), dim3(r.waves_per_block * 64), lds, st, ix, r, s);
  return hipGetLastError();
}

/// Capacity class: the smallest whose register arrays hold ef, or a larger
/// one the request asks for (min_class, a tuner dimension: 0 = ef <= 64,
/// 1 = <= 128, 2 = <= 256, 3 = <= 512).
template <int PC, int FM, int ST, int RT>
hipError_t pick_class(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st) {
  const uint32_t own = r.ef <= 64 ? 0u : r.ef <= 128 ? 1u : r.ef <= 256 ? 2u : 3u;
  const uint32_t cls = own > r.min_class ? own : (r.min_class < 3u ? r.min_class : 3u);
  if (cls == 0) return run_fast<1, 2, PC, FM, ST, RT>(ix, r, s, st);
  if (cls == 1) return run_fast<2, 4, PC, FM, ST, RT>(ix, r, s, st);
  if (cls == 2) return run_fast<4, 8, PC, FM, ST, RT>(ix, r, s, st);
  return run_fast<8, 16, PC, FM, ST, RT>(ix, r, s, st);
}

/// Rows per memory round trip in the layer-0 expansion (16 * row_trip; a
/// tuner dimension): two 16-row steps share a trip only where the row width
/// is fixed, and not in mixed launches, whose kernels hold several capacity
/// classes and are the tightest on registers.
template <int PC, int FM, int ST>
hipError_t pick_capacity(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st) {
  if constexpr (PC > 0 && FM != kPerQuery && FM != kPerQueryDeleted) {
    if (r.row_trip == 2) return pick_class<PC, FM, ST, 2>(ix, r, s, st);
  }
  return pick_class<PC, FM, ST, 1>(ix, r, s, st);
}

template <int PC, int ST>
hipError_t pick_filter(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st) {
  if (r.tasks != nullptr) {
    return r.deleted_only ? pick_capacity<PC, kPerQueryDeleted, ST>(ix, r, s, st)
                          : pick_capacity<PC, kPerQuery, ST>(ix, r, s, st);
  }
  if (r.elig != nullptr && r.deleted_only) return pick_capacity<PC, kDeleted, ST>(ix, r, s, st);
  if (r.elig != nullptr) return pick_capacity<PC, kFiltered, ST>(ix, r, s, st);
  return pick_capacity<PC, kPlain, ST>(ix, r, s, st);
}

}  // namespace

/// The fast kernel for storage ST (kStoreF32 / kStoreBF16 / kStoreFP8) and row
/// width PC (8: padded dim 128 with the fixed-width row loop; 0: any width):
/// picks the request's filter mode and capacity class and launches. Defined
/// in fast_variant.hip, one (ST, PC) per object file.
template <int ST, int PC>
hipError_t launch_fast(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st);

}  // namespace gpu
}  // namespace hnswamd
