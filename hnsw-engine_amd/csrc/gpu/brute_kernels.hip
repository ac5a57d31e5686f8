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
};

template <int E>
__global__ void __launch_bounds__(256) brute_scan(ScanArgs a) {
  extern __shared__ float qsh[];
  const int lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  float* q = qsh + static_cast<size_t>(wave) * a.padded;
  const uint32_t qi = blockIdx.x * (blockDim.x >> 6) + wave;
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

template <int E>
void launch_scan(const ScanArgs& a, hipStream_t st) {
  const uint32_t wpb = 4;
  const size_t lds = static_cast<size_t>(wpb) * a.padded * sizeof(float);
  hipLaunchKernelGGL((brute_scan<E>), dim3((a.nq + wpb - 1) / wpb), dim3(wpb * 64), lds, st, a);
  check(hipGetLastError(), "brute scan launch");
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
    check(hipMalloc(reinterpret_cast<void**>(&s->rows), std::max<size_t>(padded.size() * 4, 256)), "hipMalloc");
    check(hipMalloc(reinterpret_cast<void**>(&s->rank), std::max<size_t>(n * 4, 256)), "hipMalloc");
    check(hipMalloc(reinterpret_cast<void**>(&s->rank_label), std::max<size_t>(n * 8, 256)), "hipMalloc");
    check(hipMemcpy(s->rows, padded.data(), padded.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemcpy(s->rank, rank.data(), n * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemcpy(s->rank_label, rank_label.data(), n * 8, hipMemcpyHostToDevice), "hipMemcpy");
  } catch (...) {
    brute_free(s);
    throw;
  }
  return s;
}

void brute_free(BruteSet* s) {
  if (s == nullptr) return;
  for (void* p : {static_cast<void*>(s->rows), static_cast<void*>(s->rank), static_cast<void*>(s->rank_label),
                  static_cast<void*>(s->qs), static_cast<void*>(s->out), static_cast<void*>(s->counts),
                  static_cast<void*>(s->allow)}) {
    if (p != nullptr) (void)hipFree(p);
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
  if (k <= 64) {
    launch_scan<1>(a, st);
  } else if (k <= 128) {
    launch_scan<2>(a, st);
  } else if (k <= 256) {
    launch_scan<4>(a, st);
  } else {
    launch_scan<8>(a, st);
  }
  check(hipMemcpyAsync(out, a.out, nq * k * sizeof(hnswgpu_neighbor), hipMemcpyDeviceToHost, st), "hipMemcpy");
  if (counts != nullptr) {
    check(hipMemcpyAsync(counts, a.counts, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "hipMemcpy");
  }
  check(hipStreamSynchronize(st), "brute search");
}

}  // namespace gpu
}  // namespace hnswamd
