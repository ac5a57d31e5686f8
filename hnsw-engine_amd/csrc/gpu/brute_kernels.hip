// Exact k-NN (BruteForceIndex, reference src/brute_force.cpp:34-69): every
// allowed row is scored and the k best by (distance, label) are kept, on the
// GPU, with distances in the engine's fixed summation order over the unpadded
// dimension (a partial last chunk counts as the reference's masked tail).
//
// Scan kernel: one wavefront per query, 16 rows per step (one row per quad);
// the running top-k is a register-resident sorted array of (distance, rank)
// keys, where rank orders rows by (label, row), so key order is the
// reference's (distance, label) order.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../error.hpp"
#include "brute.hpp"
#include "device_common.hpp"
#include "quad_distance.hpp"

namespace hnswamd {
namespace gpu {

struct BruteSet {
  uint32_t n = 0, dim = 0, padded = 0;
  int metric = 0;
  float* rows = nullptr;         // n x padded, zero-padded
  float* xnorm = nullptr;        // n: |x|^2 (fp32, for the GEMM form and its error bound)
  float xnorm_max = 0.0f;
  uint32_t* rank = nullptr;      // n: position of (label, row) in sorted order
  uint64_t* rank_label = nullptr;  // n: label at each rank
  hipStream_t stream = nullptr;
  // scratch
  float* qs = nullptr;
  size_t qs_bytes = 0;
  hnswgpu_neighbor* out = nullptr;
  size_t out_bytes = 0;
  uint32_t* counts = nullptr;
  size_t counts_bytes = 0;
  uint64_t* allow = nullptr;
  size_t allow_bytes = 0;
  float* qprep = nullptr;        // MFMA path: prepared queries, nq x padded
  size_t qprep_bytes = 0;
  float* qnorm = nullptr;
  size_t qnorm_bytes = 0;
  uint64_t* cand = nullptr;      // MFMA path: per (query, chunk) lists
  size_t cand_bytes = 0;
  uint32_t* redo = nullptr;      // queries the lists could not certify (count, then ids)
  size_t redo_bytes = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // start, GEMM start, GEMM end, end
  // options
  int force_scan = 0;
  int kernel_timing = 0;
  uint32_t chunk_rows = 0;       // 0: chosen per call
  int mfma_variant = 1;          // 0: register-staged tiles, 1: global_load_lds tiles
  // last call
  int last_mfma = 0;
  uint32_t last_redo = 0;
  uint32_t last_chunks = 0;
  double last_gemm_ms = 0.0, last_total_ms = 0.0;
};

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T* grow(T*& p, size_t& have, size_t need) {
  if (need <= have && p != nullptr) return p;
  if (p != nullptr) (void)hipFree(p);
  p = nullptr;
  have = 0;
  check(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(need, 256)), "hipMalloc");
  have = std::max<size_t>(need, 256);
  return p;
}

struct ScanArgs {
  const float* rows;
  const uint32_t* rank;
  const uint64_t* rank_label;
  const uint64_t* allow;   // by row, or null
  const float* qs;         // nq x dim (raw)
  hnswgpu_neighbor* out;
  uint32_t* counts;
  uint32_t n, dim, padded, nq, k;
  int metric;
  const uint32_t* qlist;   // when set: qlist[0] queries, ids at qlist[1..]
};

template <int E>
__global__ void __launch_bounds__(256) brute_scan(ScanArgs a) {
  extern __shared__ float qsh[];
  const int lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  float* q = qsh + static_cast<size_t>(wave) * a.padded;
  uint32_t qi = blockIdx.x * (blockDim.x >> 6) + wave;
  if (a.qlist != nullptr) {
    if (qi >= a.qlist[0]) return;
    qi = a.qlist[1 + qi];
  }
  if (qi >= a.nq) return;
  const float* src = a.qs + static_cast<size_t>(qi) * a.dim;
  // Query preparation: cosine scales by float(1 / sqrt(sum of squares in double)).
  float inv = 1.0f;
  if (a.metric == 2) {
    double ss = 0.0;
    if (lane == 0) {
      for (uint32_t i = 0; i < a.dim; ++i) {
        const double x = static_cast<double>(src[i]);
        ss = ss + x * x;
      }
    }
    ss = __shfl(ss, 0);
    inv = ss > 0.0 ? static_cast<float>(1.0 / __builtin_sqrt(ss)) : 1.0f;
  }
  for (uint32_t i = static_cast<uint32_t>(lane); i < a.padded; i += 64) {
    q[i] = i < a.dim ? (a.metric == 2 ? src[i] * inv : src[i]) : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");

  const bool l2 = a.metric == 0;
  const uint32_t whole = a.dim >> 4;
  const bool partial = (a.dim & 15u) != 0u;
  WaveSorted<E> top;
  top.clear();
  uint32_t held = 0;
  uint64_t thr = kKeyInf;  // current k-th key once k rows are held
  for (uint32_t base = 0; base < a.n; base += 16) {
    const uint32_t r = base + static_cast<uint32_t>(lane >> 2);
    bool active = r < a.n;
    if (active && a.allow != nullptr) active = ((a.allow[r >> 6] >> (r & 63)) & 1ULL) != 0ULL;
    float d = quad_raw<0>(l2, a.rows + static_cast<size_t>(active ? r : 0) * a.padded, q, whole, partial, active);
    if (!l2) d = 1.0f - d;
    const uint64_t key = active ? key_of(ord_of(d), a.rank[r]) : kKeyInf;
    uint64_t want = ballot(((lane & 3) == 0) && key < thr);
    while (want != 0) {
      const int src_lane = __builtin_ctzll(want);
      want &= want - 1;
      const uint64_t x = rdlane64(key, src_lane);
      if (x < thr) {
        top.insert(x, a.k);
        if (held < a.k) ++held;
        if (held == a.k) thr = top.at(a.k - 1);
      }
    }
  }
  hnswgpu_neighbor* row = a.out + static_cast<size_t>(qi) * a.k;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t idx = static_cast<uint32_t>(e * 64 + lane);
    if (idx < held) {
      const uint64_t key = top.v[e];
      row[idx] = hnswgpu_neighbor{a.rank_label[key_id(key)], float_of(key_ord(key)), 0u};
    }
  }
  for (uint32_t j = held + static_cast<uint32_t>(lane); j < a.k; j += 64) {
    row[j] = hnswgpu_neighbor{~0ULL, __uint_as_float(0x7F800000u), 0u};
  }
  if (a.counts != nullptr && lane == 0) a.counts[qi] = held;
}

// ---------------------------------------------------------------------------
// MFMA path. fp32 matrix cores (v_mfma_f32_16x16x4f32) compute, for tiles of
// 64 rows x 16 queries per wavefront, the approximate value
//   v = |x|^2 - 2 q.x   (L2)      or      v = -q.x   (inner product, cosine),
// so that distance = |q|^2 + v (L2) or 1 + v. Rows are split over the row
// chunks of the grid and, inside a chunk, over the four lane groups of the
// MFMA output; every (query, chunk, lane group) keeps its 8 smallest v in
// registers. The merge kernel takes theta = the smallest 8th value over all
// lists: every row outside the lists has v >= theta. It re-ranks the listed
// rows with v < theta by the exact distance (the scan's summation order) and
// certifies the result when the k-th exact distance is below
// theta + |q|^2 - eps (L2; 1 + theta - eps otherwise), where eps bounds the
// fp32 rounding of both forms. Queries that cannot be certified are re-run by
// the scan kernel, so results always equal the scan's.
//
// The k dimension is permuted inside each 16-element block (lane group g
// supplies elements 4g..4g+3 over four MFMA steps), which lets every lane
// read its A operand with one 16-byte LDS load per block and leaves the dot
// product unchanged.
// ---------------------------------------------------------------------------
constexpr uint32_t kTileRows = 64;
constexpr uint32_t kMfmaWaves = 8;                 // 16 queries each
constexpr uint32_t kMfmaQueries = 16 * kMfmaWaves;  // per block
constexpr uint32_t kLaneKeep = 8;
constexpr uint32_t kChunkList = 4 * kLaneKeep;      // entries per (query, chunk)
constexpr uint32_t kMfmaMaxPadded = 256;
constexpr uint32_t kMfmaMaxK = 32;

typedef float f32x4_t __attribute__((ext_vector_type(4)));

struct MfmaArgs {
  const float* rows;      // n x P
  const float* xnorm;     // n
  const uint64_t* allow;  // bitmap by row, or null
  const float* qprep;     // nq x P
  uint64_t* cand;         // nq x nchunks x kChunkList keys (approximate ord, row)
  uint32_t n, nq, nchunks, rows_per_chunk, qblocks;
  int metric;
};

__device__ __forceinline__ void lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int P16>
__global__ void __launch_bounds__(512, P16 <= 8 ? 4 : 2) brute_mfma(MfmaArgs a) {
  constexpr uint32_t P = 16u * P16;
  constexpr uint32_t S = P + 4u;                     // LDS row stride (floats)
  constexpr uint32_t kLoads = (kTileRows * P / 4 + 511) / 512;  // float4 per thread per tile
  __shared__ float xs[kTileRows * S];
  __shared__ float xn[kTileRows];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t j = lane & 15, g = lane >> 4;
  // Blocks of one chunk are spread across the grid in order, so that the
  // blocks resident at one time read the same tiles.
  const uint32_t qb = blockIdx.x % a.qblocks, chunk = blockIdx.x / a.qblocks;
  const uint32_t q = qb * kMfmaQueries + 16u * wave + j;
  const uint32_t r0 = chunk * a.rows_per_chunk;
  const uint32_t r1 = min(a.n, r0 + a.rows_per_chunk);

  // B operand: this lane's query, elements 16b + 4g + t.
  float4 bq[P16];
#pragma unroll
  for (int b = 0; b < P16; ++b) {
    bq[b] = q < a.nq ? *reinterpret_cast<const float4*>(a.qprep + static_cast<size_t>(q) * P + 16 * b + 4 * g)
                     : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  uint32_t lo[kLaneKeep], li[kLaneKeep];
#pragma unroll
  for (int s = 0; s < static_cast<int>(kLaneKeep); ++s) {
    lo[s] = kOrdInf;
    li[s] = 0;
  }
  const bool l2 = a.metric == 0;

  float4 pre[kLoads];
  float pren = 0.0f;
  auto fetch = [&](uint32_t rb) {
#pragma unroll
    for (uint32_t i = 0; i < kLoads; ++i) {
      const uint32_t f = (tid + 512 * i) * 4;  // float index within the tile
      const uint32_t row = f / P, c = f % P;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (f < kTileRows * P && rb + row < r1) v = *reinterpret_cast<const float4*>(a.rows + static_cast<size_t>(rb + row) * P + c);
      pre[i] = v;
    }
    if (tid < kTileRows) pren = rb + tid < r1 ? a.xnorm[rb + tid] : 0.0f;
  };
  auto stage = [&]() {
#pragma unroll
    for (uint32_t i = 0; i < kLoads; ++i) {
      const uint32_t f = (tid + 512 * i) * 4;
      if (f < kTileRows * P) {
        const uint32_t row = f / P, c = f % P;
        *reinterpret_cast<float4*>(xs + row * S + c) = pre[i];
      }
    }
    if (tid < kTileRows) xn[tid] = pren;
  };

  if (r0 < r1) {
    fetch(r0);
    stage();
  }
  __syncthreads();
  for (uint32_t rb = r0; rb < r1; rb += kTileRows) {
    const bool more = rb + kTileRows < r1;
    if (more) fetch(rb + kTileRows);
    f32x4_t acc[4];
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) acc[jt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < P16; ++b) {
#pragma unroll
      for (int jt = 0; jt < 4; ++jt) {
        const float4 av = *reinterpret_cast<const float4*>(xs + (16 * jt + j) * S + 16 * b + 4 * g);
        acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bq[b].x, acc[jt], 0, 0, 0);
        acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bq[b].y, acc[jt], 0, 0, 0);
        acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bq[b].z, acc[jt], 0, 0, 0);
        acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bq[b].w, acc[jt], 0, 0, 0);
      }
    }
    // Valid rows of this tile (in range and allowed), one bit per row.
    uint64_t valid = r1 - rb >= 64 ? ~0ULL : ((1ULL << (r1 - rb)) - 1ULL);
    if (a.allow != nullptr) valid &= a.allow[rb >> 6];  // rb is a multiple of 64
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) {
      const float4 nv = *reinterpret_cast<const float4*>(xn + 16 * jt + 4 * g);
      const float norms[4] = {nv.x, nv.y, nv.z, nv.w};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t i = 16u * jt + 4u * g + static_cast<uint32_t>(r);
        const float dp = acc[jt][r];
        const float v = l2 ? norms[r] - 2.0f * dp : -dp;
        const uint32_t o = ((valid >> i) & 1ULL) ? ord_of(v) : kOrdInf;
        const bool want = o < lo[kLaneKeep - 1];
        if (ballot(want) != 0ULL) {
          if (want) {
            uint32_t co = o, ci = rb + i;
#pragma unroll
            for (int s = 0; s < static_cast<int>(kLaneKeep); ++s) {
              const bool lt = co < lo[s];
              const uint32_t to = lt ? lo[s] : co, ti = lt ? li[s] : ci;
              lo[s] = lt ? co : lo[s];
              li[s] = lt ? ci : li[s];
              co = to;
              ci = ti;
            }
          }
        }
      }
    }
    __syncthreads();
    if (more) stage();
    __syncthreads();
  }
  if (q < a.nq) {
    uint64_t* dst = a.cand + (static_cast<size_t>(q) * a.nchunks + chunk) * kChunkList + kLaneKeep * g;
#pragma unroll
    for (int s = 0; s < static_cast<int>(kLaneKeep); ++s) dst[s] = key_of(lo[s], li[s]);
  }
}

/// Column piece stored in LDS slot s of tile row r (and the slot holding
/// column piece s): 16-byte pieces are XOR-swizzled inside complete groups of
/// eight, so the wavefront's 16-byte reads of one piece over 8 rows hit 8
/// different bank groups.
template <uint32_t PP>
__device__ __forceinline__ uint32_t swz(uint32_t s, uint32_t r) {
  return (s | 7u) < PP ? (s ^ (r & 7u)) : s;
}

/// As brute_mfma, with the row tiles copied HBM -> LDS by global_load_lds
/// (no register staging) into two LDS buffers: the copy of tile t + 1 runs
/// while tile t is multiplied.
template <int P16>
__global__ void __launch_bounds__(512, P16 <= 8 ? 4 : 2) brute_mfma_dma(MfmaArgs a) {
  constexpr uint32_t P = 16u * P16;
  constexpr uint32_t PP = P / 4;  // 16-byte pieces per row; one wave instruction moves 64 pieces
  constexpr uint32_t kIssues = (PP + kMfmaWaves - 1) / kMfmaWaves;
  // Two distinct LDS objects (not one array of two), so that the compiler
  // sees that reads of one buffer do not depend on copies into the other and
  // does not wait for those copies before the reads.
  __shared__ float4 xsA[kTileRows * PP];
  __shared__ float4 xsB[kTileRows * PP];
  __shared__ float xnA[kTileRows];
  __shared__ float xnB[kTileRows];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t j = lane & 15, g = lane >> 4;
  const uint32_t qb = blockIdx.x % a.qblocks, chunk = blockIdx.x / a.qblocks;
  const uint32_t q = qb * kMfmaQueries + 16u * wave + j;
  const uint32_t r0 = chunk * a.rows_per_chunk;
  const uint32_t r1 = min(a.n, r0 + a.rows_per_chunk);

  float4 bq[P16];
#pragma unroll
  for (int b = 0; b < P16; ++b) {
    bq[b] = q < a.nq ? *reinterpret_cast<const float4*>(a.qprep + static_cast<size_t>(q) * P + 16 * b + 4 * g)
                     : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  uint32_t lo[kLaneKeep], li[kLaneKeep];
#pragma unroll
  for (int s = 0; s < static_cast<int>(kLaneKeep); ++s) {
    lo[s] = kOrdInf;
    li[s] = 0;
  }
  const bool l2 = a.metric == 0;

  // Rows past the chunk end read the chunk's last row; the epilogue masks them.
  // Addresses are 32-bit byte offsets from the chunk's first row.
  const char* cbase = reinterpret_cast<const char*>(a.rows + static_cast<size_t>(r0) * P);
  const uint32_t last = r1 - 1u - r0;
  auto issue = [&](uint32_t rb, float4* dst) {
    // The lane index is re-read here so that the per-lane offsets are
    // recomputed at each copy rather than held in registers across the loop.
    uint32_t ln = lane;
    asm volatile("" : "+v"(ln));
#pragma unroll
    for (uint32_t i = 0; i < kIssues; ++i) {
      const uint32_t inst = i * kMfmaWaves + wave;
      if (inst < PP) {
        const uint32_t p = inst * 64u + ln;
        const uint32_t r = p / PP, c = swz<PP>(p % PP, r);
        const uint32_t row = min(rb - r0 + r, last);
        const uint32_t off = (row * P + 4u * c) * 4u;
        typedef const __attribute__((address_space(1))) float* gptr_t;
        typedef __attribute__((address_space(3))) float* lptr_t;
        __builtin_amdgcn_global_load_lds((gptr_t)(cbase + off), (lptr_t)(dst + inst * 64u), 16, 0, 0);
      }
    }
  };
  // Next tile's norms (wave 0, one per lane) and allow word, loaded with its copy.
  float xreg = 0.0f;
  uint64_t wreg = ~0ULL;
  auto side = [&](uint32_t rb) {
    if (wave == 0) xreg = rb + lane < r1 ? a.xnorm[rb + lane] : 0.0f;
    if (a.allow != nullptr) wreg = a.allow[rb >> 6];
  };
  // One tile: wait for its copy, start the next tile's copy into the other
  // buffer, multiply, and keep the 8 best per lane.
  const uint32_t slot0 = g ^ (j & 7u), slot1 = (4u + g) ^ (j & 7u);
  auto tile = [&](uint32_t rb, const float4* X, float* xn, float4* Xnext, float* xnnext) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's copies of tile rb have landed
    if (wave == 0) xn[lane] = xreg;
    uint64_t valid = r1 - rb >= 64 ? ~0ULL : ((1ULL << (r1 - rb)) - 1ULL);
    valid &= wreg;
    __syncthreads();                     // every wave's copies landed; the other buffer is free
    if (rb + kTileRows < r1) {
      issue(rb + kTileRows, Xnext);
      side(rb + kTileRows);
    }
    (void)xnnext;
    f32x4_t acc[4];
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) acc[jt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < P16; ++b) {
      // With whole groups of 8 pieces, piece 4b + g of row j sits at
      // 8 * (b / 2) + ((4 * (b % 2) + g) ^ (j % 8)): two per-lane bases and
      // constant offsets, so the reads need no per-b address registers.
      const uint32_t slot = PP % 8 == 0 ? (b % 2 ? slot1 : slot0) + 8u * (b / 2) : swz<PP>(4u * b + g, j);
#pragma unroll
      for (int jt = 0; jt < 4; ++jt) {
        const float4 av = X[(16u * jt + j) * PP + slot];
        acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bq[b].x, acc[jt], 0, 0, 0);
        acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bq[b].y, acc[jt], 0, 0, 0);
        acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bq[b].z, acc[jt], 0, 0, 0);
        acc[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bq[b].w, acc[jt], 0, 0, 0);
      }
    }
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) {
      const float4 nv = *reinterpret_cast<const float4*>(&xn[16 * jt + 4 * g]);
      const float norms[4] = {nv.x, nv.y, nv.z, nv.w};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t i = 16u * jt + 4u * g + static_cast<uint32_t>(r);
        const float dp = acc[jt][r];
        const float v = l2 ? norms[r] - 2.0f * dp : -dp;
        const uint32_t o = ((valid >> i) & 1ULL) ? ord_of(v) : kOrdInf;
        const bool want = o < lo[kLaneKeep - 1];
        if (ballot(want) != 0ULL) {
          if (want) {
            uint32_t co = o, ci = rb + i;
#pragma unroll
            for (int s = 0; s < static_cast<int>(kLaneKeep); ++s) {
              const bool lt = co < lo[s];
              const uint32_t to = lt ? lo[s] : co, ti = lt ? li[s] : ci;
              lo[s] = lt ? co : lo[s];
              li[s] = lt ? ci : li[s];
              co = to;
              ci = ti;
            }
          }
        }
      }
    }
  };
  if (r0 < r1) {
    issue(r0, xsA);
    side(r0);
  }
  for (uint32_t rb = r0; rb < r1; rb += 2 * kTileRows) {
    tile(rb, xsA, xnA, xsB, xnB);
    if (rb + kTileRows < r1) tile(rb + kTileRows, xsB, xnB, xsA, xnA);
  }
  if (q < a.nq) {
    uint64_t* dst = a.cand + (static_cast<size_t>(q) * a.nchunks + chunk) * kChunkList + kLaneKeep * g;
#pragma unroll
    for (int s = 0; s < static_cast<int>(kLaneKeep); ++s) dst[s] = key_of(lo[s], li[s]);
  }
}

/// One wavefront per query: the padded, prepared query (cosine: scaled by
/// float(1 / sqrt(sum of squares in double)), summed in order) and |q|^2.
__global__ void __launch_bounds__(256) brute_prepare(const float* qs, uint32_t nq, uint32_t dim, uint32_t P,
                                                     int metric, float* qprep, float* qnorm) {
  const int lane = lane_id();
  const uint32_t qi = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (qi >= nq) return;
  const float* src = qs + static_cast<size_t>(qi) * dim;
  float inv = 1.0f;
  if (metric == 2) {
    double ss = 0.0;
    if (lane == 0) {
      for (uint32_t i = 0; i < dim; ++i) {
        const double x = static_cast<double>(src[i]);
        ss = ss + x * x;
      }
    }
    ss = __shfl(ss, 0);
    inv = ss > 0.0 ? static_cast<float>(1.0 / __builtin_sqrt(ss)) : 1.0f;
  }
  float sq = 0.0f;
  for (uint32_t i = static_cast<uint32_t>(lane); i < P; i += 64) {
    const float v = i < dim ? (metric == 2 ? src[i] * inv : src[i]) : 0.0f;
    qprep[static_cast<size_t>(qi) * P + i] = v;
    sq = __builtin_fmaf(v, v, sq);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o);
  if (lane == 0) qnorm[qi] = sq;
}

__global__ void __launch_bounds__(256) brute_norms(const float* rows, uint32_t n, uint32_t P, float* xnorm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float sq = 0.0f;
  for (uint32_t c = 0; c < P; ++c) {
    const float v = rows[static_cast<size_t>(i) * P + c];
    sq = __builtin_fmaf(v, v, sq);
  }
  xnorm[i] = sq;
}

struct MergeArgs {
  const float* rows;
  const uint32_t* rank;
  const uint64_t* rank_label;
  const float* qprep;
  const float* qnorm;
  const uint64_t* cand;
  hnswgpu_neighbor* out;
  uint32_t* counts;
  uint32_t* redo;      // [0] count, then query ids
  uint32_t nq, dim, P, k, nchunks;
  int metric;
  float eps_scale;     // eps = eps_scale * (|q|^2 + max |x|^2 + 1)
  float xnorm_max;
};

template <int E>
__global__ void __launch_bounds__(256) brute_merge(MergeArgs a) {
  extern __shared__ float msh[];
  const int lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  float* q = msh + static_cast<size_t>(wave) * (a.P + 64);
  uint32_t* pick = reinterpret_cast<uint32_t*>(q + a.P);
  const uint32_t qi = blockIdx.x * (blockDim.x >> 6) + wave;
  if (qi >= a.nq) return;
  for (uint32_t i = static_cast<uint32_t>(lane); i < a.P; i += 64) q[i] = a.qprep[static_cast<size_t>(qi) * a.P + i];
  const uint64_t* list = a.cand + static_cast<size_t>(qi) * a.nchunks * kChunkList;
  const uint32_t total = a.nchunks * kChunkList;
  // theta: the smallest last entry over all (chunk, lane group) lists.
  uint32_t theta = kOrdInf;
  for (uint32_t e = static_cast<uint32_t>(lane); e < total; e += 64) {
    if ((e & (kLaneKeep - 1)) == kLaneKeep - 1) {
      const uint32_t o = key_ord(list[e]);
      theta = o < theta ? o : theta;
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const uint32_t t = static_cast<uint32_t>(__shfl_xor(static_cast<int>(theta), o));
    theta = t < theta ? t : theta;
  }
  lds_fence();
  const bool l2 = a.metric == 0;
  const uint32_t whole = a.dim >> 4;
  const bool partial = (a.dim & 15u) != 0u;
  WaveSorted<E> top;
  top.clear();
  uint32_t held = 0;
  uint64_t kth = kKeyInf;
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t e = base + static_cast<uint32_t>(lane);
    const uint64_t key = e < total ? list[e] : kKeyInf;
    const bool take = key_ord(key) < theta;
    const uint64_t m = ballot(take);
    const uint32_t nt = static_cast<uint32_t>(__popcll(m));
    if (take) pick[__popcll(m & ((1ULL << lane) - 1ULL))] = key_id(key);
    lds_fence();
    for (uint32_t s = 0; s < nt; s += 16) {
      const uint32_t jj = s + static_cast<uint32_t>(lane >> 2);
      const bool active = jj < nt;
      const uint32_t row = active ? pick[jj] : 0u;
      float d = quad_raw<0>(l2, a.rows + static_cast<size_t>(row) * a.P, q, whole, partial, active);
      if (!l2) d = 1.0f - d;
      const uint64_t x = active ? key_of(ord_of(d), a.rank[row]) : kKeyInf;
      uint64_t want = ballot(((lane & 3) == 0) && x < kth);
      while (want != 0) {
        const int src = __builtin_ctzll(want);
        want &= want - 1;
        const uint64_t y = rdlane64(x, src);
        if (y < kth) {
          top.insert(y, a.k);
          if (held < a.k) ++held;
          if (held == a.k) kth = top.at(a.k - 1);
        }
      }
    }
    lds_fence();
  }
  // Certificate: every row left out has exact distance >= bound > k-th distance.
  bool exact = theta == kOrdInf;
  if (!exact && held == a.k) {
    const float qn = a.qnorm[qi];
    const float eps = a.eps_scale * ((qn + a.xnorm_max) + 1.0f);
    const float th = float_of(theta);
    const float bound = (l2 ? th + qn : 1.0f + th) - eps;
    exact = float_of(key_ord(kth)) < bound;
  }
  if (!exact) {
    if (lane == 0) {
      const uint32_t slot = atomicAdd(&a.redo[0], 1u);
      a.redo[1 + slot] = qi;
    }
    return;
  }
  hnswgpu_neighbor* row = a.out + static_cast<size_t>(qi) * a.k;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t idx = static_cast<uint32_t>(e * 64 + lane);
    if (idx < held) {
      const uint64_t key = top.v[e];
      row[idx] = hnswgpu_neighbor{a.rank_label[key_id(key)], float_of(key_ord(key)), 0u};
    }
  }
  for (uint32_t jj = held + static_cast<uint32_t>(lane); jj < a.k; jj += 64) {
    row[jj] = hnswgpu_neighbor{~0ULL, __uint_as_float(0x7F800000u), 0u};
  }
  if (a.counts != nullptr && lane == 0) a.counts[qi] = held;
}

__global__ void zero_word(uint32_t* p) {
  if (threadIdx.x == 0) p[0] = 0u;
}

template <int P16>
void launch_mfma_t(const MfmaArgs& a, int variant, hipStream_t st) {
  if (variant == 1) {
    hipLaunchKernelGGL((brute_mfma_dma<P16>), dim3(a.qblocks * a.nchunks), dim3(512), 0, st, a);
  } else {
    hipLaunchKernelGGL((brute_mfma<P16>), dim3(a.qblocks * a.nchunks), dim3(512), 0, st, a);
  }
}

void launch_mfma(uint32_t P, const MfmaArgs& a, int variant, hipStream_t st) {
  switch (P / 16) {
    case 1: launch_mfma_t<1>(a, variant, st); break;
    case 2: launch_mfma_t<2>(a, variant, st); break;
    case 3: launch_mfma_t<3>(a, variant, st); break;
    case 4: launch_mfma_t<4>(a, variant, st); break;
    case 5: launch_mfma_t<5>(a, variant, st); break;
    case 6: launch_mfma_t<6>(a, variant, st); break;
    case 7: launch_mfma_t<7>(a, variant, st); break;
    case 8: launch_mfma_t<8>(a, variant, st); break;
    case 9: launch_mfma_t<9>(a, variant, st); break;
    case 10: launch_mfma_t<10>(a, variant, st); break;
    case 11: launch_mfma_t<11>(a, variant, st); break;
    case 12: launch_mfma_t<12>(a, variant, st); break;
    case 13: launch_mfma_t<13>(a, variant, st); break;
    case 14: launch_mfma_t<14>(a, variant, st); break;
    case 15: launch_mfma_t<15>(a, variant, st); break;
    case 16: launch_mfma_t<16>(a, variant, st); break;
    default: throw std::logic_error("brute_mfma: unsupported padded dimension");
  }
  check(hipGetLastError(), "brute mfma launch");
}

template <int E>
void launch_scan(const ScanArgs& a, hipStream_t st) {
  const uint32_t wpb = 4;
  const size_t lds = static_cast<size_t>(wpb) * a.padded * sizeof(float);
  // With a query list the grid still covers nq waves; waves past the list's length exit at once.
  hipLaunchKernelGGL((brute_scan<E>), dim3((a.nq + wpb - 1) / wpb), dim3(wpb * 64), lds, st, a);
  check(hipGetLastError(), "brute scan launch");
}

void scan_dispatch(const ScanArgs& a, hipStream_t st) {
  if (a.k <= 64) {
    launch_scan<1>(a, st);
  } else if (a.k <= 128) {
    launch_scan<2>(a, st);
  } else if (a.k <= 256) {
    launch_scan<4>(a, st);
  } else {
    launch_scan<8>(a, st);
  }
}

}  // namespace

BruteSet* brute_create(std::size_t dim, int metric, const float* rows, const uint64_t* labels, std::size_t n) {
  if (dim == 0 || dim > 4096) throw std::invalid_argument("BruteForceIndex: GPU path supports 1 <= dim <= 4096");
  if (n >= 0xFFFFFFFFull) throw std::invalid_argument("BruteForceIndex: too many rows");
  auto* s = new BruteSet();
  try {
    s->n = static_cast<uint32_t>(n);
    s->dim = static_cast<uint32_t>(dim);
    s->padded = static_cast<uint32_t>((dim + 15) / 16 * 16);
    s->metric = metric;
    check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    std::vector<float> padded(static_cast<size_t>(n) * s->padded, 0.0f);
    for (size_t i = 0; i < n; ++i) std::copy(rows + i * dim, rows + (i + 1) * dim, padded.data() + i * s->padded);
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return labels[x] < labels[y]; });
    std::vector<uint32_t> rank(n);
    std::vector<uint64_t> rank_label(n);
    for (size_t p = 0; p < n; ++p) {
      rank[order[p]] = static_cast<uint32_t>(p);
      rank_label[p] = labels[order[p]];
    }
    // Upper bound of max |x|^2 for the certificate (double sum, then one part in 2^20 added).
    double xmax = 0.0;
    for (size_t i = 0; i < n; ++i) {
      double sq = 0.0;
      for (size_t c = 0; c < dim; ++c) sq += static_cast<double>(rows[i * dim + c]) * rows[i * dim + c];
      xmax = std::max(xmax, sq);
    }
    s->xnorm_max = static_cast<float>(xmax * (1.0 + 0x1p-20));
    check(hipMalloc(reinterpret_cast<void**>(&s->rows), std::max<size_t>(padded.size() * 4, 256)), "hipMalloc");
    check(hipMalloc(reinterpret_cast<void**>(&s->xnorm), std::max<size_t>(n * 4, 256)), "hipMalloc");
    check(hipMalloc(reinterpret_cast<void**>(&s->rank), std::max<size_t>(n * 4, 256)), "hipMalloc");
    check(hipMalloc(reinterpret_cast<void**>(&s->rank_label), std::max<size_t>(n * 8, 256)), "hipMalloc");
    check(hipMemcpy(s->rows, padded.data(), padded.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemcpy(s->rank, rank.data(), n * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemcpy(s->rank_label, rank_label.data(), n * 8, hipMemcpyHostToDevice), "hipMemcpy");
    if (n > 0) {
      hipLaunchKernelGGL(brute_norms, dim3((static_cast<uint32_t>(n) + 255) / 256), dim3(256), 0, s->stream, s->rows,
                         s->n, s->padded, s->xnorm);
      check(hipGetLastError(), "brute norms launch");
    }
    for (hipEvent_t& e : s->ev) check(hipEventCreate(&e), "hipEventCreate");
    check(hipStreamSynchronize(s->stream), "brute create");
  } catch (...) {
    brute_free(s);
    throw;
  }
  return s;
}

void brute_free(BruteSet* s) {
  if (s == nullptr) return;
  for (void* p : {static_cast<void*>(s->rows), static_cast<void*>(s->xnorm), static_cast<void*>(s->rank),
                  static_cast<void*>(s->rank_label), static_cast<void*>(s->qs), static_cast<void*>(s->out),
                  static_cast<void*>(s->counts), static_cast<void*>(s->allow), static_cast<void*>(s->qprep),
                  static_cast<void*>(s->qnorm), static_cast<void*>(s->cand), static_cast<void*>(s->redo)}) {
    if (p != nullptr) (void)hipFree(p);
  }
  for (hipEvent_t e : s->ev) {
    if (e != nullptr) (void)hipEventDestroy(e);
  }
  if (s->stream != nullptr) (void)hipStreamDestroy(s->stream);
  delete s;
}

void brute_search(BruteSet* s, const float* qs, std::size_t nq, std::size_t k, const uint64_t* allow_by_row,
                  hnswgpu_neighbor* out, uint32_t* counts) {
  if (k > 512) throw std::invalid_argument("BruteForceIndex: GPU path supports k <= 512");
  if (nq == 0) return;
  hipStream_t st = s->stream;
  ScanArgs a{};
  a.rows = s->rows;
  a.rank = s->rank;
  a.rank_label = s->rank_label;
  a.n = s->n;
  a.dim = s->dim;
  a.padded = s->padded;
  a.nq = static_cast<uint32_t>(nq);
  a.k = static_cast<uint32_t>(k);
  a.metric = s->metric;
  a.qs = grow(s->qs, s->qs_bytes, nq * s->dim * sizeof(float));
  a.out = grow(s->out, s->out_bytes, nq * k * sizeof(hnswgpu_neighbor));
  a.counts = grow(s->counts, s->counts_bytes, nq * sizeof(uint32_t));
  check(hipMemcpyAsync(const_cast<float*>(a.qs), qs, nq * s->dim * sizeof(float), hipMemcpyHostToDevice, st),
        "hipMemcpy");
  a.allow = nullptr;
  if (allow_by_row != nullptr) {
    const size_t bytes = ((s->n + 63) / 64) * sizeof(uint64_t);
    a.allow = grow(s->allow, s->allow_bytes, bytes);
    check(hipMemcpyAsync(const_cast<uint64_t*>(a.allow), allow_by_row, bytes, hipMemcpyHostToDevice, st),
          "hipMemcpy");
  }
  a.qlist = nullptr;
  const bool mfma = s->force_scan == 0 && k <= kMfmaMaxK && s->padded <= kMfmaMaxPadded && s->n > 0;
  s->last_mfma = mfma ? 1 : 0;
  s->last_redo = 0;
  s->last_chunks = 0;
  if (s->kernel_timing) check(hipEventRecord(s->ev[0], st), "hipEventRecord");
  if (mfma) {
    const uint32_t P = s->padded, n32 = s->n, q32 = static_cast<uint32_t>(nq);
    const uint32_t qblocks = (q32 + kMfmaQueries - 1) / kMfmaQueries;
    uint32_t rows = s->chunk_rows;
    if (rows == 0) {
      // Enough blocks to fill the chip several times over, chunks of 2k..16k rows.
      const uint32_t want = std::max<uint32_t>(1, (4096 + qblocks - 1) / qblocks);
      rows = std::clamp<uint32_t>((n32 + want - 1) / want, 2048, 16384);
    }
    rows = (rows + kTileRows - 1) / kTileRows * kTileRows;
    const uint32_t nchunks = (n32 + rows - 1) / rows;
    s->last_chunks = nchunks;
    MfmaArgs m{};
    m.rows = s->rows;
    m.xnorm = s->xnorm;
    m.allow = a.allow;
    m.qprep = grow(s->qprep, s->qprep_bytes, nq * P * sizeof(float));
    m.cand = grow(s->cand, s->cand_bytes, nq * nchunks * kChunkList * sizeof(uint64_t));
    m.n = n32;
    m.nq = q32;
    m.nchunks = nchunks;
    m.rows_per_chunk = rows;
    m.qblocks = qblocks;
    m.metric = s->metric;
    float* qnorm = grow(s->qnorm, s->qnorm_bytes, nq * sizeof(float));
    uint32_t* redo = grow(s->redo, s->redo_bytes, (nq + 1) * sizeof(uint32_t));
    hipLaunchKernelGGL(brute_prepare, dim3((q32 + 3) / 4), dim3(256), 0, st, a.qs, q32, s->dim, P, s->metric,
                       const_cast<float*>(m.qprep), qnorm);
    check(hipGetLastError(), "brute prepare launch");
    hipLaunchKernelGGL(zero_word, dim3(1), dim3(64), 0, st, redo);
    check(hipGetLastError(), "zero launch");
    if (s->kernel_timing) check(hipEventRecord(s->ev[1], st), "hipEventRecord");
    launch_mfma(P, m, s->mfma_variant, st);
    if (s->kernel_timing) check(hipEventRecord(s->ev[2], st), "hipEventRecord");
    MergeArgs g{};
    g.rows = s->rows;
    g.rank = s->rank;
    g.rank_label = s->rank_label;
    g.qprep = m.qprep;
    g.qnorm = qnorm;
    g.cand = m.cand;
    g.out = a.out;
    g.counts = a.counts;
    g.redo = redo;
    g.nq = q32;
    g.dim = s->dim;
    g.P = P;
    g.k = static_cast<uint32_t>(k);
    g.nchunks = nchunks;
    g.metric = s->metric;
    g.eps_scale = 8.0f * static_cast<float>(P + 4) * 0x1p-24f;
    g.xnorm_max = s->xnorm_max;
    const size_t lds = 4 * (static_cast<size_t>(P) + 64) * sizeof(float);
    hipLaunchKernelGGL((brute_merge<1>), dim3((q32 + 3) / 4), dim3(256), lds, st, g);
    check(hipGetLastError(), "brute merge launch");
    // Queries without a certificate: the scan, on those queries only.
    a.qlist = redo;
  }
  scan_dispatch(a, st);
  if (s->kernel_timing) check(hipEventRecord(s->ev[3], st), "hipEventRecord");
  check(hipMemcpyAsync(out, a.out, nq * k * sizeof(hnswgpu_neighbor), hipMemcpyDeviceToHost, st), "hipMemcpy");
  if (counts != nullptr) {
    check(hipMemcpyAsync(counts, a.counts, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "hipMemcpy");
  }
  check(hipStreamSynchronize(st), "brute search");
  if (mfma) {
    check(hipMemcpy(&s->last_redo, s->redo, sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy");
  }
  if (s->kernel_timing) {
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, s->ev[0], s->ev[3]), "hipEventElapsedTime");
    s->last_total_ms = ms;
    s->last_gemm_ms = 0.0;
    if (mfma) {
      check(hipEventElapsedTime(&ms, s->ev[1], s->ev[2]), "hipEventElapsedTime");
      s->last_gemm_ms = ms;
    }
  }
}

void brute_set_option(BruteSet* s, const std::string& name, int64_t value) {
  if (name == "force_scan") {
    s->force_scan = value != 0;
  } else if (name == "kernel_timing") {
    s->kernel_timing = value != 0;
  } else if (name == "mfma_variant") {
    if (value < 0 || value > 1) throw std::invalid_argument("mfma_variant must be 0 or 1");
    s->mfma_variant = static_cast<int>(value);
  } else if (name == "chunk_rows") {
    if (value < 0 || value > (1 << 30)) throw std::invalid_argument("chunk_rows out of range");
    s->chunk_rows = static_cast<uint32_t>(value);
  } else {
    throw std::invalid_argument("unknown brute-force option: " + name);
  }
}

int64_t brute_get_option(const BruteSet* s, const std::string& name) {
  if (name == "force_scan") return s->force_scan;
  if (name == "kernel_timing") return s->kernel_timing;
  if (name == "chunk_rows") return s->chunk_rows;
  if (name == "mfma_variant") return s->mfma_variant;
  if (name == "last_mfma") return s->last_mfma;
  if (name == "last_redo") return s->last_redo;
  if (name == "last_chunks") return s->last_chunks;
  if (name == "last_gemm_ns") return static_cast<int64_t>(s->last_gemm_ms * 1e6);
  if (name == "last_total_ns") return static_cast<int64_t>(s->last_total_ms * 1e6);
  throw std::invalid_argument("unknown brute-force option: " + name);
}

}  // namespace gpu
}  // namespace hnswamd
