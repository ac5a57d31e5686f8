#pragma once

// Host-side state of a resident index (the hnswgpu_index handle) and the
// per-launch scratch, shared by the C-ABI layer (device_index.hip) and the
// serving loop (server.hip).
//
// HBM layout of one index (all arrays separate allocations, 256-byte aligned):
//   vecs       count x padded fp32       rows of padded*4 bytes (a multiple of 64)
//   l0         count x (1 + 2M) u32      [count][ids...] per node
//   upper      upper_used u32            per node of level L: L blocks of (1 + M)
//   upper_off  count u32                 prefix sum of level * (1 + M)
//   labels     count u64
//   deleted    ceil(count / 32) u32      bit set = tombstone
//   reduced    per built storage: count x padded bf16 (2 B) or e4m3 (1 B), scales, |x^|^2
// Scratch (one set per launch in flight): counters, overflow queue, prepared
// queries, eligibility bitmaps, HBM visited tables (one per resident
// wavefront), general-kernel bitmaps and heaps.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../../include/hnswgpu.h"
#include "../error.hpp"
#include "launch.hpp"

namespace hnswamd {
namespace dev {

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

/// A device allocation that grows (discarding contents) when asked for more.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void* need(size_t n) {
    if (n <= bytes && p != nullptr) return p;
    release();
    const size_t alloc = std::max<size_t>(n, 256);
    check(hipMalloc(&p, alloc), "hipMalloc");
    bytes = alloc;
    return p;
  }
  void release() {
    if (p != nullptr) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
};

/// Pinned host memory (hipHostMalloc) that grows when asked for more.
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void* need(size_t n) {
    if (n <= bytes && p != nullptr) return p;
    release();
    const size_t alloc = std::max<size_t>(n, 256);
    check(hipHostMalloc(&p, alloc, hipHostMallocDefault), "hipHostMalloc");
    bytes = alloc;
    return p;
  }
  void release() {
    if (p != nullptr) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
};

/// Scratch of one launch in flight. Launches that may overlap on the device
/// need separate Scratch objects (each holds the visited tables of the
/// wavefronts of its kernel).
struct Scratch {
  DevBuf counters, overflow, qprep, elig, spill_bits, spill_heaps, visit_tables, stats, cand_runs;
  size_t visit_tables_cleared = 0;
  uint64_t spill_alloc_count = 0;
  uint32_t spill_alloc_waves = 0;
};

/// A reduced-precision traversal copy of the rows.
struct Reduced {
  DevBuf low, scale, norm;
  bool built = false;
};

}  // namespace dev
}  // namespace hnswamd

namespace hnswamd {
namespace dev {

/// The resources of one search call in flight: its scratch, device staging for
/// host-buffer calls, a stream of its own (host-buffer calls run there) and an
/// event marking the end of the last work it enqueued. A call takes a context
/// from the handle's pool and gives it back when its work is enqueued (device
/// buffers) or finished (host buffers); the next user of the context orders
/// itself after that event when it runs on another stream.
struct CallCtx {
  Scratch sc;
  DevBuf qs, out, counts, filter, qscale, task_dev;
  PinnedBuf task_host;
  PinnedBuf stage_in, stage_out;  // host-buffer searches of small batches: queries + scales, rows + counts
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipStream_t last_stream = nullptr;
  bool used = false;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;  // fast-kernel timing (option kernel_timing)
  bool timing_pending = false;
  uint64_t last_nq = 0;   // queries of the last search (work counts)
  bool stats_on = false;  // the last search recorded work counts

  CallCtx() = default;
  CallCtx(const CallCtx&) = delete;
  CallCtx& operator=(const CallCtx&) = delete;
  ~CallCtx() {
    if (ev_begin != nullptr) (void)hipEventDestroy(ev_begin);
    if (ev_end != nullptr) (void)hipEventDestroy(ev_end);
    if (done != nullptr) (void)hipEventDestroy(done);
    if (stream != nullptr) (void)hipStreamDestroy(stream);
  }
};

}  // namespace dev
}  // namespace hnswamd

/// A resident index. Searches hold `mu` shared for their whole call, so
/// concurrent callers overlap on the device, each with a context of its own;
/// uploads, deletions, options and tuning hold it exclusively.
struct hnswgpu_index {
  std::shared_mutex mu;
  hnswamd::gpu::DevIndexView view{};  // view.storage: the copy search_batch traverses
  uint64_t dim = 0;
  uint64_t ef_search = 10;
  uint64_t num_deleted = 0;
  uint64_t upper_used = 0;
  std::vector<uint8_t> deleted_host;  // mirror of the device bits, for counting

  hnswamd::dev::DevBuf vecs, l0, upper, upper_off, labels, deleted;
  hnswamd::dev::Reduced reduced[3];  // [1] bf16, [2] fp8 e4m3 + scale

  // Call contexts: at most max_ctxs, created on demand; `last` is the one the
  // most recent search used (read by the last_* options and read_stats).
  std::mutex pool_mu;
  std::condition_variable pool_cv;
  std::vector<std::unique_ptr<hnswamd::dev::CallCtx>> ctxs;
  std::vector<hnswamd::dev::CallCtx*> idle;
  hnswamd::dev::CallCtx* last = nullptr;
  uint32_t max_ctxs = 4;

  // Servers attached to this index (hnswgpu_server_create); while any is,
  // the index cannot be re-uploaded or freed.
  std::atomic<uint32_t> servers{0};

  // Launch shapes chosen by the autotuner, per ef class (64, 128, 256, 512),
  // used while use_tuned is set.
  struct Tuned {
    bool valid = false;
    uint32_t waves_per_cu = 16, hash_log2 = 11;
    int generic_gather = 0;
    int row_trip = 1;         // 16-row steps per memory round trip
    uint32_t min_class = 0;   // capacity class the ef class runs in (>= its own)
    double ms = 0.0;
  };
  Tuned tuned[4];
  int use_tuned = 0;

  // options
  int generic_gather = 0;
  int wide_hash = 0;
  int row_trip = 0;           // 0: 2 for a reduced copy, 1 for fp32 rows (measured defaults)
  uint32_t min_class = 0;
  uint32_t spill_waves = 64;
  uint32_t visit_log2 = 15;
  uint32_t waves_per_cu = 16;
  uint32_t hash_log2 = 0;
  int force_spill = 0;
  int deleted_path = 1;
  int collect_stats = 0;
  int kernel_timing = 0;

  hipStream_t stream = nullptr;  // uploads, deletions and reduced-copy builds

  ~hnswgpu_index() {
    ctxs.clear();
    if (stream != nullptr) (void)hipStreamDestroy(stream);
  }
};

namespace hnswamd {
namespace dev {

/// The index as the kernels see it, traversing `storage` (0 fp32, 1 bf16, 2
/// fp8); the reduced copy must be built.
gpu::DevIndexView view_for(const hnswgpu_index& h, int storage);
/// Replace the index's arrays in place (after an add): the handle stays valid
/// and keeps its options. Refused while a server is attached.
void reupload(hnswgpu_index& h, const hnswgpu_index_view& v);
/// Build the reduced copy for `storage` if it is not built yet (synchronous).
void build_reduced(hnswgpu_index& h, int storage);
/// Launch shape for r.ef and nq queries: hash size, wavefronts per block and a
/// persistent grid of the resident wavefronts.
void plan_request(const hnswgpu_index& h, gpu::SearchRequest& r, uint32_t nq);
/// Scratch for a launch planned by plan_request with room for nq queries.
gpu::SearchScratch bind_scratch(hnswgpu_index& h, Scratch& sc, const gpu::SearchRequest& r, uint32_t nq,
                                uint32_t spill_waves, hipStream_t st);
/// Index of the capacity class of ef: 0 (<= 64), 1 (<= 128), 2 (<= 256), 3.
uint32_t ef_class_index(uint32_t ef);
/// Effective ef of a request: max(ef or ef_search, k).
uint64_t effective_ef(const hnswgpu_index& h, uint64_t k, uint64_t ef);
/// 1 when searches whose only exclusions are soft deletions run the unfiltered
/// traversal with a side list of deleted nodes (option deleted_path).
uint32_t deleted_side_list(const hnswgpu_index& h);
/// The same for a filter allowing the fraction `allowed` of labels (< 0:
/// unknown): the side list when at least 15 nodes in 16 are eligible.
uint32_t side_list_for(const hnswgpu_index& h, double allowed);
/// Take a call context from the pool (waits while max_ctxs are in use) and
/// give it back; the one given back last is the handle's most recent. A call
/// that enqueues on the caller's stream `st` (non-null) is given, in this
/// order: an idle context last used on st, one whose work has finished, a
/// new one while fewer than max_ctxs exist, else the most recent (st then
/// waits for its work). So calls on different streams overlap on the device,
/// and calls on one stream reuse one context.
CallCtx* acquire_ctx(hnswgpu_index& h, hipStream_t st = nullptr);
void release_ctx(hnswgpu_index& h, CallCtx* c);
/// Before enqueuing on st: order st after the context's earlier work. After:
/// record the context's event on st.
void ctx_begin(CallCtx& c, hipStream_t st);
void ctx_end(CallCtx& c, hipStream_t st);
/// A context held for the scope of one call.
struct CtxLease {
  hnswgpu_index& h;
  CallCtx* c;
  explicit CtxLease(hnswgpu_index& hh, hipStream_t st = nullptr) : h(hh), c(acquire_ctx(hh, st)) {}
  ~CtxLease() { release_ctx(h, c); }
  CtxLease(const CtxLease&) = delete;
  CtxLease& operator=(const CtxLease&) = delete;
};
/// Cosine scale of each of nq host queries (dim floats each, row-major) as the
/// reference's prepare_vector computes it (src/space.hpp:31-40): the sum of
/// squares in double, in order, then float(1 / sqrt), 1 for a zero vector.
void host_query_scales(const float* qs, uint64_t nq, uint64_t dim, float* out);
/// Mixed launch of n tasks on `st`: eligibility bitmaps for the filters the
/// tasks use, the task array staged in pinned `stage` and copied to
/// `d_tasks`, then prepare + fast + general kernels with scratch `sc`.
void launch_tasks(hnswgpu_index& h, Scratch& sc, const hnswgpu_task* tasks, uint64_t n,
                  const hnswgpu_filter* filters, uint64_t nf, int storage, gpu::QueryTask* stage,
                  gpu::QueryTask* d_tasks, uint32_t spill_waves, hipStream_t st);

}  // namespace dev
}  // namespace hnswamd
