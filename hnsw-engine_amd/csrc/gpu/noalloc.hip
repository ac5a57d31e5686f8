// hnswgpu_search_batch_device_noalloc: hnswgpu_search_batch_device for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc). It runs the same kernels;
// their scratch comes from one file-scope device array per GPU instead of the
// handle's context pool, and the call makes no hipMalloc / hipFree and no
// stream or event creation.
//
// Layout of the array (kNoAllocBytes per GPU, held until the process ends):
//   [0, kVisitRegion)   HBM visited tables, 2^kVisitLog2 slots per resident
//                       wavefront (at most kMaxWaves wavefronts); a fill
//                       kernel empties the part a call uses, on the call's
//                       stream, before its search (a kernel, not a memset, so
//                       the call can be captured into a graph);
//   [kVisitRegion, ...) per call, 256-byte aligned: counters, the hand-off
//                       queue (nq), prepared queries (nq x padded_dim),
//                       the eligibility bitmap (filter or deletions), and the
//                       general kernel's bitmaps and heaps for 1..8 wavefronts
//                       (O(count) each).
// A call whose layout does not fit is refused with EINVAL and the sizes. The
// filtered kernel gets no HBM candidate runs here: a query whose candidates
// outgrow its registers finishes in the general kernel (same rows, slower).
//
// Every call on one GPU uses that GPU's array: calls that overlap on a GPU
// (other streams, other threads, other handles) race and must be ordered by
// the caller. The first call on each device resolves the array's address
// (hipGetSymbolAddress, with that device current) and may load this module:
// make it outside graph capture.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>

#include "../../../include/hnswgpu.h"
#include "../error.hpp"
#include "index_state.hpp"
#include "launch.hpp"

namespace {

using hnswamd::dev::check;

constexpr size_t kNoAllocBytes = size_t{96} << 20;
constexpr size_t kVisitRegion = size_t{16} << 20;
constexpr uint32_t kVisitLog2 = 11;
constexpr uint32_t kMaxWaves = static_cast<uint32_t>(kVisitRegion / ((size_t{1} << kVisitLog2) * sizeof(uint32_t)));
constexpr uint32_t kMaxSpillWaves = 8;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

/// The array's device address per device (resolved once, with that device current).
struct DeviceScratch {
  std::once_flag once;
  uint8_t* ptr = nullptr;
  hipError_t err = hipSuccess;
};
DeviceScratch g_scratch[kMaxDevices];

__global__ void __launch_bounds__(256) fill_words(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    p[i] = v;
  }
}

uint8_t* scratch_on_current_device() {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw hnswamd::DeviceError("search_noalloc: device ordinal out of range");
  DeviceScratch& d = g_scratch[dev];
  std::call_once(d.once, [&] {
    void* p = nullptr;
    d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (d.err == hipSuccess) d.ptr = static_cast<uint8_t*>(p);
  });
  if (d.err != hipSuccess || d.ptr == nullptr) {
    throw hnswamd::DeviceError(std::string("search_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(d.err));
  }
  return d.ptr;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return HNSWGPU_OK;
  } catch (...) {
    return hnswamd::status_from_current_exception();
  }
}

}  // namespace

extern "C" int hnswgpu_search_batch_device_noalloc(hnswgpu_index* h, const float* d_qs, uint64_t nq, uint64_t k,
                                                   uint64_t ef, const uint64_t* d_filter_words, uint64_t filter_size,
                                                   hnswgpu_neighbor* d_out, uint32_t* d_counts, void* stream) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("search_batch: null handle");
    if (k == 0) throw std::invalid_argument("search_batch: k must be > 0");
    if (nq == 0) return;
    if (d_qs == nullptr || d_out == nullptr) throw std::invalid_argument("search_batch: null pointer");
    std::shared_lock<std::shared_mutex> lk(h->mu);
    const hnswamd::gpu::DevIndexView& ix = h->view;
    if (ix.count == 0) throw std::invalid_argument("search_batch: empty index on the device path");
    const uint64_t ef_eff = hnswamd::dev::effective_ef(*h, k, ef);
    if (ef_eff > 0xFFFFFFFFull || k > 0xFFFFFFFFull || nq > 0xFFFFFFFFull) {
      throw std::invalid_argument("search_batch: size out of range");
    }
    const auto st = static_cast<hipStream_t>(stream);

    hnswamd::gpu::SearchRequest r{};
    r.queries = d_qs;
    r.dim = static_cast<uint32_t>(h->dim);
    r.nq = static_cast<uint32_t>(nq);
    r.k = static_cast<uint32_t>(k);
    r.ef = static_cast<uint32_t>(ef_eff);
    r.out = reinterpret_cast<hnswamd::gpu::OutRow*>(d_out);
    r.counts = d_counts;
    r.force_spill = h->force_spill;
    hnswamd::dev::plan_request(*h, r, r.nq);
    // The persistent grid may be smaller than the handle's plan: every
    // wavefront pulls queries from the ticket until none are left.
    const uint32_t max_blocks = std::max<uint32_t>(1, kMaxWaves / r.waves_per_block);
    r.blocks = std::min(r.blocks, max_blocks);

    // Per-call layout after the visited tables.
    const size_t n = ix.count;
    const bool filtered = d_filter_words != nullptr || h->num_deleted > 0;
    const size_t words = (n + 31) / 32;
    size_t off = kVisitRegion;
    const size_t counters_off = off;
    off += 256;
    const size_t overflow_off = off;
    off += align256(nq * sizeof(uint32_t));
    const size_t qprep_off = off;
    off += align256(nq * ix.padded * sizeof(float));
    const size_t elig_off = off;
    if (filtered) off += align256(words * sizeof(uint32_t));
    const size_t spill_bitmap_bytes = align256(words * sizeof(uint32_t));
    const size_t spill_heap_bytes = align256((2 * n + 4) * sizeof(uint64_t));
    const size_t per_spill = spill_bitmap_bytes + spill_heap_bytes;
    const size_t room = kNoAllocBytes > off ? kNoAllocBytes - off : 0;
    const uint32_t spill_waves = static_cast<uint32_t>(std::min<size_t>(kMaxSpillWaves, room / per_spill));
    if (spill_waves == 0) {
      throw std::invalid_argument("search_noalloc: this call needs " + std::to_string(off + per_spill) +
                                  " bytes of scratch (" + std::to_string(nq) + " queries of padded dim " +
                                  std::to_string(ix.padded) + ", " + std::to_string(n) +
                                  " nodes); the static scratch holds " + std::to_string(kNoAllocBytes));
    }

    uint8_t* base = scratch_on_current_device();
    hnswamd::gpu::SearchScratch s{};
    s.counters = reinterpret_cast<uint32_t*>(base + counters_off);
    s.overflow = reinterpret_cast<uint32_t*>(base + overflow_off);
    s.qprep = reinterpret_cast<float*>(base + qprep_off);
    s.visit_tables = reinterpret_cast<uint32_t*>(base);
    s.visit_log2 = kVisitLog2;
    s.spill_waves = spill_waves;
    s.spill_bitmap_words = static_cast<uint32_t>(words);
    s.spill_words = reinterpret_cast<uint32_t*>(base + off);
    s.spill_heaps = reinterpret_cast<uint64_t*>(base + off + spill_waves * spill_bitmap_bytes);
    s.spill_heap_stride = spill_heap_bytes / sizeof(uint64_t);
    s.cand_runs = nullptr;

    // Empty the visited tables of the wavefronts this launch has (the kernels
    // keep them empty, but the array is shared by every call on the GPU).
    const size_t table_words = static_cast<size_t>(r.blocks) * r.waves_per_block << kVisitLog2;
    hipLaunchKernelGGL(fill_words, dim3(static_cast<unsigned>(std::min<size_t>((table_words + 255) / 256, 4096))),
                       dim3(256), 0, st, s.visit_tables, table_words, 0xFFFFFFFFu);
    check(hipGetLastError(), "fill kernel");

    if (filtered) {
      auto* e = reinterpret_cast<uint32_t*>(base + elig_off);
      check(hnswamd::gpu::launch_eligibility(ix, static_cast<const uint32_t*>(h->deleted.p), d_filter_words,
                                             filter_size, 0, e, st), "eligibility kernel");
      r.elig = e;
      r.deleted_only = d_filter_words == nullptr ? static_cast<int>(hnswamd::dev::deleted_side_list(*h)) : 0;
    }
    check(hnswamd::gpu::launch_prepare(ix, r, s, st), "prepare kernel");
    check(hnswamd::gpu::launch_search(ix, r, s, st), "search kernel");
  });
}
