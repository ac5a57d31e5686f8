// The device layer of the C ABI (include/hnswgpu.h, hnswgpu_*): uploads a
// loaded index into HBM once and serves search_batch from it. The HBM layout
// is described in index_state.hpp.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/hnswgpu.h"
#include "../error.hpp"
#include "exact_scale.hpp"
#include "index_state.hpp"
#include "launch.hpp"

using hnswamd::dev::check;
using hnswamd::dev::DevBuf;
using hnswamd::gpu::DevIndexView;
using hnswamd::gpu::OutRow;
using hnswamd::gpu::QueryTask;
using hnswamd::gpu::SearchRequest;
using hnswamd::gpu::SearchScratch;

namespace {

uint32_t next_log2(uint64_t v) {
  uint32_t l = 0;
  while ((1ULL << l) < v) ++l;
  return l;
}

}  // namespace

namespace {

void upload_view(hnswgpu_index& h, const hnswgpu_index_view& v) {
  if (v.dim == 0) throw std::invalid_argument("upload: dim must be > 0");
  if (v.padded_dim != (v.dim + 15) / 16 * 16) throw std::invalid_argument("upload: padded_dim must be dim rounded up to 16");
  if (v.padded_dim > 4096) throw std::invalid_argument("upload: GPU search supports padded_dim <= 4096");
  if (v.M < 2 || 2 * v.M > hnswamd::gpu::kMaxDegree0) {
    throw std::invalid_argument("upload: GPU search supports 2 <= M <= 32");
  }
  if (v.metric < 0 || v.metric > 2) throw std::invalid_argument("upload: unknown metric");
  if (v.count >= 0xFFFFFFFFull) throw std::invalid_argument("upload: count exceeds the 32-bit id space");
  if (v.upper_used >= 0xFFFFFFFFull) throw std::invalid_argument("upload: upper section too large");
  const uint64_t n = v.count;
  const uint64_t M0 = 2 * v.M;
  if (n > 0) {
    if (v.vectors == nullptr || v.level0 == nullptr || v.labels == nullptr || v.levels == nullptr) {
      throw std::invalid_argument("upload: null array");
    }
    if (v.entry_point >= n || v.max_level < 0 || v.levels[v.entry_point] != v.max_level) {
      throw std::invalid_argument("upload: invalid entry point / max level");
    }
  }
  // Graph checks that keep every device access in bounds, and the duplicate scan.
  std::vector<uint32_t> off(n);
  uint64_t pos = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (v.levels[i] > v.max_level) throw std::invalid_argument("upload: node level exceeds max level");
    off[i] = static_cast<uint32_t>(pos);
    pos += static_cast<uint64_t>(v.levels[i]) * (1 + v.M);
  }
  if (pos != v.upper_used) throw std::invalid_argument("upload: upper size does not match levels");
  if (pos > 0 && v.upper == nullptr) throw std::invalid_argument("upload: null upper array");
  uint32_t has_dups = 0;
  std::vector<uint32_t> seen;
  for (uint64_t i = 0; i < n; ++i) {
    for (int L = 0; L <= v.levels[i]; ++L) {
      const uint32_t* l = L == 0 ? v.level0 + i * (1 + M0) : v.upper + off[i] + static_cast<uint64_t>(L - 1) * (1 + v.M);
      const uint64_t cap = L == 0 ? M0 : v.M;
      if (l[0] > cap) throw std::invalid_argument("upload: neighbor count exceeds maximum degree");
      for (uint32_t j = 0; j < l[0]; ++j) {
        const uint32_t nb = l[1 + j];
        if (nb >= n || v.levels[nb] < L) throw std::invalid_argument("upload: neighbor id out of range");
      }
      if (L == 0 && !has_dups) {
        seen.assign(l + 1, l + 1 + l[0]);
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) has_dups = 1;
      }
    }
  }

  const uint64_t words = (n + 31) / 32;
  std::vector<uint32_t> del_bits(words, 0u);
  h.deleted_host.assign(n, 0);
  h.num_deleted = 0;
  if (v.deleted != nullptr) {
    for (uint64_t i = 0; i < n; ++i) {
      if (v.deleted[i] != 0) {
        del_bits[i >> 5] |= 1u << (i & 31);
        h.deleted_host[i] = 1;
        ++h.num_deleted;
      }
    }
  }
  auto put = [&](DevBuf& b, const void* src, size_t bytes) {
    void* d = b.need(bytes);
    if (bytes > 0) check(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice), "hipMemcpy upload");
    return d;
  };
  h.view.vecs = static_cast<const float*>(put(h.vecs, v.vectors, n * v.padded_dim * sizeof(float)));
  h.view.l0 = static_cast<const uint32_t*>(put(h.l0, v.level0, n * (1 + M0) * sizeof(uint32_t)));
  h.view.upper = static_cast<const uint32_t*>(put(h.upper, v.upper, v.upper_used * sizeof(uint32_t)));
  h.view.upper_off = static_cast<const uint32_t*>(put(h.upper_off, off.data(), n * sizeof(uint32_t)));
  h.view.labels = static_cast<const uint64_t*>(put(h.labels, v.labels, n * sizeof(uint64_t)));
  put(h.deleted, del_bits.data(), words * sizeof(uint32_t));
  h.view.count = static_cast<uint32_t>(n);
  h.view.padded = static_cast<uint32_t>(v.padded_dim);
  h.view.M = static_cast<uint32_t>(v.M);
  h.view.M0 = static_cast<uint32_t>(M0);
  h.view.entry = v.entry_point;
  h.view.max_level = v.max_level;
  h.view.metric = v.metric;
  h.view.has_dups = has_dups;
  h.dim = v.dim;
  h.ef_search = v.ef_search == 0 ? 10 : v.ef_search;
  h.upper_used = v.upper_used;
}

/// LDS visited-hash words: about 16 slots per unit of ef, within [2^11, 2^13]
/// words; queries that visit more continue in the HBM table. With 16-bit
/// entries (two slots per word; ids of at most 2^(log2 slots + 8) nodes) that
/// is 8 words per unit of ef.
uint32_t auto_hash_log2(uint32_t ef, uint32_t count, bool wide) {
  auto clamp = [](uint32_t l) { return std::min<uint32_t>(13, std::max<uint32_t>(11, l)); };
  if (!wide) {
    const uint32_t l = clamp(next_log2(static_cast<uint64_t>(ef) * 8));
    const uint32_t id_bits = count > 1 ? next_log2(count) : 1;
    if (id_bits <= l + 1 + 8) return l;
  }
  return clamp(next_log2(static_cast<uint64_t>(ef) * 16));
}

/// Wavefronts resident at once for a kernel using `lds_per_wave` bytes of LDS
/// and at most `waves_per_cu` waves per CU by registers (MI355X: 256 CUs,
/// 160 KiB LDS each).
uint32_t resident_waves(size_t lds_per_wave, uint32_t wpb, uint32_t waves_per_cu) {
  const size_t per_block = lds_per_wave * wpb;
  const uint32_t by_lds = static_cast<uint32_t>((160 * 1024) / per_block) * wpb;
  return 256u * std::max<uint32_t>(wpb, std::min(by_lds, waves_per_cu));
}


/// Device-side search on device buffers with context c (caller holds h.mu
/// shared). qscale: cosine scales computed by the host, or null.
void search_device(hnswgpu_index& h, hnswamd::dev::CallCtx& c, const float* d_qs, const float* d_qscale,
                   uint64_t nq, uint64_t k, uint64_t ef, const uint64_t* d_fw, uint64_t fsize, int by_id,
                   hnswgpu_neighbor* d_out, uint32_t* d_counts, hipStream_t st, double allowed = -1.0) {
  const uint64_t ef_eff = hnswamd::dev::effective_ef(h, k, ef);
  if (ef_eff > 0xFFFFFFFFull || k > 0xFFFFFFFFull || nq > 0xFFFFFFFFull) {
    throw std::invalid_argument("search_batch: size out of range");
  }
  const DevIndexView& ix = h.view;
  SearchRequest r{};
  r.queries = d_qs;
  r.qscale = d_qscale;
  r.dim = static_cast<uint32_t>(h.dim);
  r.nq = static_cast<uint32_t>(nq);
  r.k = static_cast<uint32_t>(k);
  r.ef = static_cast<uint32_t>(ef_eff);
  r.out = reinterpret_cast<OutRow*>(d_out);
  r.counts = d_counts;
  r.stats = h.collect_stats ? static_cast<uint32_t*>(c.sc.stats.need(nq * 3 * sizeof(uint32_t))) : nullptr;
  c.last_nq = nq;
  c.stats_on = h.collect_stats != 0;
  r.force_spill = h.force_spill;
  hnswamd::dev::plan_request(h, r, r.nq);
  SearchScratch s = hnswamd::dev::bind_scratch(h, c.sc, r, r.nq, h.spill_waves, st);

  const bool filtered = d_fw != nullptr || h.num_deleted > 0;
  r.elig = nullptr;
  if (filtered) {
    auto* e = static_cast<uint32_t*>(c.sc.elig.need(((ix.count + 31) / 32) * sizeof(uint32_t)));
    check(hnswamd::gpu::launch_eligibility(ix, static_cast<const uint32_t*>(h.deleted.p), d_fw, fsize, by_id, e, st),
          "eligibility kernel");
    r.elig = e;
    // Few nodes excluded: the unfiltered traversal with a side list for the ineligible nodes it admits.
    r.deleted_only = static_cast<int>(hnswamd::dev::side_list_for(h, d_fw == nullptr ? 1.0 : allowed));
  }
  check(hnswamd::gpu::launch_prepare(ix, r, s, st), "prepare kernel");
  hipEvent_t b = nullptr, e = nullptr;
  if (h.kernel_timing) {
    if (c.ev_begin == nullptr) check(hipEventCreate(&c.ev_begin), "hipEventCreate");
    if (c.ev_end == nullptr) check(hipEventCreate(&c.ev_end), "hipEventCreate");
    b = c.ev_begin;
    e = c.ev_end;
  }
  check(hnswamd::gpu::launch_search(ix, r, s, st, b, e), "search kernel");
  c.timing_pending = b != nullptr;
}

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

namespace hnswamd {
namespace dev {

uint64_t effective_ef(const hnswgpu_index& h, uint64_t k, uint64_t ef) {
  return std::max<uint64_t>(ef == 0 ? h.ef_search : ef, k);
}

uint32_t deleted_side_list(const hnswgpu_index& h) { return side_list_for(h, 1.0); }

uint32_t side_list_for(const hnswgpu_index& h, double allowed) {
  // The side list holds the admitted ineligible nodes (64 keys in registers;
  // a live key pushed out hands the query on), so it suits bitmaps that
  // exclude few nodes: at most 1 in 16, measured against the filtered
  // traversal at 1 %, 10 %, 50 % and 75 % excluded on the 1M x 128 mixture.
  // Sparser filters run the filtered traversal. allowed < 0: the filter's
  // density is not known on the host (device words): the filtered traversal.
  if (h.deleted_path == 0 || allowed < 0.0) return h.deleted_path == 2 ? 1u : 0u;
  if (h.deleted_path == 2) return 1u;
  const double kept = h.view.count > 0 ? 1.0 - static_cast<double>(h.num_deleted) / h.view.count : 1.0;
  return allowed * kept * 16.0 >= 15.0 ? 1u : 0u;
}

CallCtx* acquire_ctx(hnswgpu_index& h, hipStream_t st) {
  std::unique_lock<std::mutex> lk(h.pool_mu);
  h.pool_cv.wait(lk, [&] { return !h.idle.empty() || h.ctxs.size() < h.max_ctxs; });
  auto take = [&](size_t i) {
    CallCtx* c = h.idle[i];
    h.idle.erase(h.idle.begin() + static_cast<std::ptrdiff_t>(i));
    return c;
  };
  if (st != nullptr) {
    for (size_t i = h.idle.size(); i-- > 0;) {
      if (h.idle[i]->used && h.idle[i]->last_stream == st) return take(i);
    }
    for (size_t i = h.idle.size(); i-- > 0;) {
      if (!h.idle[i]->used || hipEventQuery(h.idle[i]->done) == hipSuccess) return take(i);
    }
    if (h.ctxs.size() >= h.max_ctxs) return take(h.idle.size() - 1);
  } else if (!h.idle.empty()) {
    return take(h.idle.size() - 1);
  }
  auto c = std::make_unique<CallCtx>();
  check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  check(hipEventCreateWithFlags(&c->done, hipEventDisableTiming), "hipEventCreate");
  CallCtx* p = c.get();
  h.ctxs.push_back(std::move(c));
  return p;
}

void release_ctx(hnswgpu_index& h, CallCtx* c) {
  {
    std::lock_guard<std::mutex> lk(h.pool_mu);
    h.idle.push_back(c);
    h.last = c;
  }
  h.pool_cv.notify_one();
}

void ctx_begin(CallCtx& c, hipStream_t st) {
  // Also on the stream last used: a caller may have destroyed that stream with
  // work in flight and been given a new one with the same handle.
  if (c.used) check(hipStreamWaitEvent(st, c.done, 0), "hipStreamWaitEvent");
}

void ctx_end(CallCtx& c, hipStream_t st) {
  check(hipEventRecord(c.done, st), "hipEventRecord");
  c.last_stream = st;
  c.used = true;
}

void host_query_scales(const float* qs, uint64_t nq, uint64_t dim, float* out) {
  for (uint64_t i = 0; i < nq; ++i) {
    const float* x = qs + i * dim;
    double norm = 0.0;
    for (uint64_t j = 0; j < dim; ++j) norm += static_cast<double>(x[j]) * x[j];
    out[i] = norm > 0.0 ? static_cast<float>(1.0 / std::sqrt(norm)) : 1.0f;
  }
}

gpu::DevIndexView view_for(const hnswgpu_index& h, int storage) {
  gpu::DevIndexView v = h.view;
  v.storage = 0;
  v.vlow = nullptr;
  v.vscale = nullptr;
  v.vnorm = nullptr;
  if (storage == 0 || v.count == 0) return v;
  if (storage < 0 || storage > 2 || !h.reduced[storage].built) {
    throw std::invalid_argument("storage " + std::to_string(storage) + " is not built on this index");
  }
  v.storage = storage;
  v.vlow = h.reduced[storage].low.p;
  v.vscale = static_cast<const float*>(h.reduced[storage].scale.p);
  v.vnorm = static_cast<const float*>(h.reduced[storage].norm.p);
  return v;
}

void build_reduced(hnswgpu_index& h, int storage) {
  if (storage < 1 || storage > 2) return;
  Reduced& red = h.reduced[storage];
  if (red.built || h.view.count == 0) return;
  const size_t n = h.view.count, p = h.view.padded;
  void* low = red.low.need(n * p * (storage == 1 ? 2 : 1));
  auto* scale = static_cast<float*>(red.scale.need(n * sizeof(float)));
  auto* norm = static_cast<float*>(red.norm.need(n * sizeof(float)));
  gpu::DevIndexView base = h.view;
  check(gpu::launch_build_storage(base, storage, low, scale, norm, h.stream), "build storage");
  check(hipStreamSynchronize(h.stream), "build storage");
  red.built = true;
}

/// Replace the index's arrays in place (after add_batch): the handle stays
/// valid, options and tuned shapes are kept, and the traversal copy in use is
/// rebuilt from the new rows. Refused while a server is attached (its slots
/// are sized for the old index).
void reupload(hnswgpu_index& h, const hnswgpu_index_view& v) {
  std::unique_lock<std::shared_mutex> lk(h.mu);
  if (h.servers.load() != 0) throw std::invalid_argument("upload: a server is attached to this index");
  const int storage = h.view.storage;
  upload_view(h, v);
  for (Reduced& red : h.reduced) red.built = false;
  h.view.storage = 0;
  h.view.vlow = nullptr;
  h.view.vscale = nullptr;
  h.view.vnorm = nullptr;
  if (storage != 0 && h.view.count > 0) {
    build_reduced(h, storage);
    h.view = view_for(h, storage);
  }
}

uint32_t ef_class_index(uint32_t ef) {
  return ef <= 64 ? 0u : ef <= 128 ? 1u : ef <= 256 ? 2u : 3u;
}

void plan_request(const hnswgpu_index& h, gpu::SearchRequest& r, uint32_t nq) {
  const uint32_t ef = std::min<uint32_t>(r.ef, gpu::kFastMaxEf);
  uint32_t waves_per_cu = h.waves_per_cu;
  r.hash_log2 = h.hash_log2 != 0 ? h.hash_log2 : auto_hash_log2(ef, h.view.count, h.wide_hash != 0);
  r.generic_gather = h.generic_gather;
  r.wide_hash = h.wide_hash;
  // Reduced rows take half (bf16) or a quarter (fp8) of the registers of fp32
  // ones, so two 16-row steps share a memory round trip there at no cost in
  // occupancy; fp32 rows would drop a wavefront per SIMD (measured slower).
  r.row_trip = h.row_trip != 0 ? h.row_trip : (h.view.storage != 0 ? 2 : 1);
  r.min_class = h.min_class;
  const hnswgpu_index::Tuned& t = h.tuned[ef_class_index(ef)];
  if (h.use_tuned && t.valid) {
    waves_per_cu = t.waves_per_cu;
    r.hash_log2 = t.hash_log2;
    r.generic_gather = t.generic_gather;
    r.row_trip = t.row_trip;
    r.min_class = t.min_class;
  }
  const size_t per_wave = static_cast<size_t>(gpu::lds_words_per_wave(r.hash_log2, h.view.padded)) * 4u;
  r.waves_per_block = static_cast<uint32_t>(std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / per_wave)));
  // A persistent grid of the resident wavefronts (each owns an HBM visited table).
  const uint32_t waves = std::min<uint32_t>(resident_waves(per_wave, r.waves_per_block, waves_per_cu),
                                            (nq + r.waves_per_block - 1) / r.waves_per_block * r.waves_per_block);
  r.blocks = std::max<uint32_t>(1, waves / r.waves_per_block);
}

gpu::SearchScratch bind_scratch(hnswgpu_index& h, Scratch& sc, const gpu::SearchRequest& r, uint32_t nq,
                                uint32_t spill_waves, hipStream_t st) {
  const gpu::DevIndexView& ix = h.view;
  gpu::SearchScratch s{};
  s.counters = static_cast<uint32_t*>(sc.counters.need(8 * sizeof(uint32_t)));
  s.overflow = static_cast<uint32_t*>(sc.overflow.need(static_cast<size_t>(nq) * sizeof(uint32_t)));
  s.qprep = static_cast<float*>(sc.qprep.need(static_cast<size_t>(nq) * ix.padded * sizeof(float)));
  const uint64_t n = ix.count;
  if (sc.spill_alloc_count != n || sc.spill_alloc_waves != spill_waves || sc.spill_bits.p == nullptr) {
    sc.spill_bits.need(static_cast<size_t>(spill_waves) * ((n + 31) / 32) * sizeof(uint32_t));
    sc.spill_heaps.need(static_cast<size_t>(spill_waves) * (2 * n + 4) * sizeof(uint64_t));
    sc.spill_alloc_count = n;
    sc.spill_alloc_waves = spill_waves;
  }
  s.spill_words = static_cast<uint32_t*>(sc.spill_bits.p);
  s.spill_heaps = static_cast<uint64_t*>(sc.spill_heaps.p);
  s.spill_heap_stride = 2ULL * n + 4;
  s.spill_waves = spill_waves;
  s.spill_bitmap_words = static_cast<uint32_t>((n + 31) / 32);
  const size_t table_words = static_cast<size_t>(1) << h.visit_log2;
  s.visit_log2 = h.visit_log2;
  const size_t want = static_cast<size_t>(r.blocks) * r.waves_per_block * table_words * sizeof(uint32_t);
  if (sc.visit_tables.bytes < want) sc.visit_tables_cleared = 0;
  s.visit_tables = static_cast<uint32_t*>(sc.visit_tables.need(want));
  // Candidate overflow of the filtered search: kRunSlots runs of 64 keys per wavefront.
  s.cand_runs = static_cast<uint64_t*>(sc.cand_runs.need(static_cast<size_t>(r.blocks) * r.waves_per_block *
                                                          gpu::kRunSlots * 64 * sizeof(uint64_t)));
  if (sc.visit_tables_cleared < sc.visit_tables.bytes) {
    // Once per allocation; the kernels leave every table empty after each query.
    check(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(sc.visit_tables.p), 0xFFFFFFFFu,
                            sc.visit_tables.bytes / sizeof(uint32_t), st), "clear visited tables");
    sc.visit_tables_cleared = sc.visit_tables.bytes;
  }
  return s;
}

void launch_tasks(hnswgpu_index& h, Scratch& sc, const hnswgpu_task* tasks, uint64_t n,
                  const hnswgpu_filter* filters, uint64_t nf, int storage, gpu::QueryTask* stage,
                  gpu::QueryTask* d_tasks, uint32_t spill_waves, hipStream_t st) {
  if (n == 0) return;
  if (n > 0x7FFFFFFFull) throw std::invalid_argument("search_tasks: too many tasks");
  if (tasks == nullptr || (nf > 0 && filters == nullptr)) throw std::invalid_argument("search_tasks: null pointer");
  const gpu::DevIndexView ix = view_for(h, storage);
  if (ix.count == 0) throw std::invalid_argument("search_tasks: empty index on the device path");
  const size_t words = (ix.count + 31) / 32;
  // Eligibility bitmaps: one per filter that a task uses, and one for
  // deletions alone when there are deletions.
  std::vector<int> used(nf, 0);
  for (uint64_t i = 0; i < n; ++i) {
    const hnswgpu_task& t = tasks[i];
    if (t.k == 0) throw std::invalid_argument("search_tasks: k must be > 0");
    if (t.query == nullptr || t.out == nullptr) throw std::invalid_argument("search_tasks: null pointer");
    if (t.filter < -1 || (t.filter >= 0 && static_cast<uint64_t>(t.filter) >= nf)) {
      throw std::invalid_argument("search_tasks: filter index out of range");
    }
    if (t.filter >= 0) used[t.filter] = 1;
  }
  auto* elig = static_cast<uint32_t*>(sc.elig.need((nf + 1) * words * sizeof(uint32_t)));
  const auto* del = static_cast<const uint32_t*>(h.deleted.p);
  for (uint64_t f = 0; f < nf; ++f) {
    if (!used[f]) continue;
    check(gpu::launch_eligibility(ix, del, filters[f].words, filters[f].size, 0, elig + f * words, st),
          "eligibility kernel");
  }
  const uint32_t* del_only = nullptr;
  if (h.num_deleted > 0) {
    check(gpu::launch_eligibility(ix, del, nullptr, 0, 0, elig + nf * words, st), "eligibility kernel");
    del_only = elig + nf * words;
  }
  uint64_t max_ef = 1;
  for (uint64_t i = 0; i < n; ++i) {
    const hnswgpu_task& t = tasks[i];
    const uint64_t ef = effective_ef(h, t.k, t.ef);
    if (ef > 0xFFFFFFFFull) throw std::invalid_argument("search_tasks: ef out of range");
    max_ef = std::max(max_ef, ef);
    gpu::QueryTask q{};
    q.query = t.query;
    q.out = reinterpret_cast<gpu::OutRow*>(t.out);
    q.count = t.count;
    q.elig = t.filter >= 0 ? elig + static_cast<size_t>(t.filter) * words : del_only;
    q.k = t.k;
    q.ef = static_cast<uint32_t>(ef);
    q.has_scale = 0;  // device queries: the cosine scale is computed on the device
    stage[i] = q;
  }
  check(hipMemcpyAsync(d_tasks, stage, n * sizeof(gpu::QueryTask), hipMemcpyHostToDevice, st), "hipMemcpy tasks");
  gpu::SearchRequest r{};
  r.tasks = d_tasks;
  r.dim = static_cast<uint32_t>(h.dim);
  r.nq = static_cast<uint32_t>(n);
  r.ef = static_cast<uint32_t>(std::min<uint64_t>(max_ef, 0xFFFFFFFFull));
  r.force_spill = h.force_spill;
  // Deletions and no filter in the whole launch: the side-list kernel.
  bool any_filter = false;
  for (uint64_t f = 0; f < nf; ++f) any_filter = any_filter || used[f] != 0;
  r.deleted_only = del_only != nullptr && !any_filter ? static_cast<int>(deleted_side_list(h)) : 0;
  plan_request(h, r, r.nq);
  gpu::SearchScratch s = bind_scratch(h, sc, r, r.nq, spill_waves, st);
  check(gpu::launch_prepare(ix, r, s, st), "prepare kernel");
  check(gpu::launch_search(ix, r, s, st), "search kernel");
}

}  // namespace dev
}  // namespace hnswamd

extern "C" {

int hnswgpu_upload(const hnswgpu_index_view* view, hnswgpu_index** out) {
  return guarded([&] {
    if (view == nullptr || out == nullptr) throw std::invalid_argument("upload: null pointer");
    auto* h = new hnswgpu_index();
    try {
      check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
      upload_view(*h, *view);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int hnswgpu_reupload(hnswgpu_index* h, const hnswgpu_index_view* view) {
  return guarded([&] {
    if (h == nullptr || view == nullptr) throw std::invalid_argument("upload: null pointer");
    hnswamd::dev::reupload(*h, *view);
  });
}

int hnswgpu_free(hnswgpu_index* h) {
  return guarded([&] {
    if (h != nullptr && h->servers.load() != 0) throw std::invalid_argument("free: a server is attached to this index");
    if (h != nullptr) (void)hipDeviceSynchronize();  // device-buffer searches may still use its contexts
    delete h;
  });
}

int hnswgpu_set_deleted(hnswgpu_index* h, uint32_t id, uint8_t deleted) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("set_deleted: null handle");
    std::unique_lock<std::shared_mutex> lk(h->mu);
    if (id >= h->view.count) throw std::out_of_range("set_deleted: id out of range");
    const uint8_t want = deleted != 0 ? 1 : 0;
    if (h->deleted_host[id] == want) return;
    h->deleted_host[id] = want;
    h->num_deleted += want ? 1 : static_cast<uint64_t>(-1);
    const uint32_t word_idx = id >> 5;
    uint32_t word = 0;
    for (uint32_t b = 0; b < 32; ++b) {
      const uint64_t j = static_cast<uint64_t>(word_idx) * 32 + b;
      if (j < h->view.count && h->deleted_host[j]) word |= 1u << b;
    }
    auto* d = static_cast<uint32_t*>(h->deleted.p) + word_idx;
    check(hipMemcpyAsync(d, &word, sizeof(word), hipMemcpyHostToDevice, h->stream), "hipMemcpy deleted");
    check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  });
}

int hnswgpu_set_ef(hnswgpu_index* h, uint64_t ef) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("set_ef: null handle");
    if (ef == 0) throw std::invalid_argument("set_ef: ef must be > 0");
    std::unique_lock<std::shared_mutex> lk(h->mu);
    h->ef_search = ef;
  });
}

}  // extern "C"

namespace hnswamd {

/// Largest batch (bytes of queries and scales in, or of rows and counts out)
/// that a host-buffer search stages through its context's pinned buffers.
constexpr size_t kStageBytes = size_t{1} << 20;

/// Host-buffer search used by the C ABI and the C++ facade. Filter words are
/// label-indexed (a BitsetFilter) or, when by_id, indexed by internal id
/// (any other Filter, evaluated on the host). Concurrent calls on one handle
/// run on contexts of their own and overlap on the device.
void device_search_host(hnswgpu_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                        const uint64_t* filter_words, uint64_t filter_size, int by_id, hnswgpu_neighbor* out,
                        uint32_t* counts) {
  if (h == nullptr) throw std::invalid_argument("search_batch: null handle");
  if (k == 0) throw std::invalid_argument("search_batch: k must be > 0");
  if (nq == 0) return;
  if (qs == nullptr || out == nullptr) throw std::invalid_argument("search_batch: null pointer");
  std::shared_lock<std::shared_mutex> lk(h->mu);
  if (h->view.count == 0) {
    for (uint64_t i = 0; i < nq * k; ++i) out[i] = hnswgpu_neighbor{~0ULL, __builtin_inff(), 0u};
    if (counts != nullptr) std::fill(counts, counts + nq, 0u);
    return;
  }
  dev::CtxLease lease(*h);
  dev::CallCtx& c = *lease.c;
  hipStream_t st = c.stream;
  dev::ctx_begin(c, st);
  // A call that ends in an exception waits for the copies it queued before
  // the context (and its pinned staging buffers) goes to the next caller.
  struct DrainOnError {
    hipStream_t st;
    bool armed = true;
    ~DrainOnError() {
      if (armed) (void)hipStreamSynchronize(st);
    }
  } drain{st};
  auto* d_qs = static_cast<float*>(c.qs.need(nq * h->dim * sizeof(float)));
  auto* d_out = static_cast<hnswgpu_neighbor*>(c.out.need(nq * k * sizeof(hnswgpu_neighbor)));
  auto* d_counts = static_cast<uint32_t*>(c.counts.need(nq * sizeof(uint32_t)));
  // Small batches go through the context's pinned staging buffers: a copy
  // from pageable memory costs the runtime a staging pass of its own, which
  // is most of the copy time of a single query.
  const size_t q_bytes = nq * h->dim * sizeof(float), s_bytes = nq * sizeof(float);
  const size_t o_bytes = nq * k * sizeof(hnswgpu_neighbor), c_bytes = nq * sizeof(uint32_t);
  const bool staged = q_bytes + s_bytes <= kStageBytes && o_bytes + c_bytes <= kStageBytes;
  uint8_t* in = staged ? static_cast<uint8_t*>(c.stage_in.need(q_bytes + s_bytes)) : nullptr;
  uint8_t* back = staged ? static_cast<uint8_t*>(c.stage_out.need(o_bytes + c_bytes)) : nullptr;
  if (staged) std::memcpy(in, qs, q_bytes);
  check(hipMemcpyAsync(d_qs, staged ? static_cast<const void*>(in) : qs, q_bytes, hipMemcpyHostToDevice, st),
        "hipMemcpy queries");
  const float* d_scale = nullptr;
  std::vector<float> scales;
  if (h->view.metric == 2) {
    // Cosine: the query scale in double on the host, as the reference computes it.
    if (!staged) scales.resize(nq);
    float* sc = staged ? reinterpret_cast<float*>(in + q_bytes) : scales.data();
    dev::host_query_scales(qs, nq, h->dim, sc);
    auto* f = static_cast<float*>(c.qscale.need(s_bytes));
    check(hipMemcpyAsync(f, sc, s_bytes, hipMemcpyHostToDevice, st), "hipMemcpy scales");
    d_scale = f;
  }
  const uint64_t* d_fw = nullptr;
  if (filter_words != nullptr) {
    const uint64_t fbits = by_id ? h->view.count : filter_size;
    const size_t fbytes = ((fbits + 63) / 64) * sizeof(uint64_t);
    auto* f = static_cast<uint64_t*>(c.filter.need(fbytes));
    if (fbytes > 0) check(hipMemcpyAsync(f, filter_words, fbytes, hipMemcpyHostToDevice, st), "hipMemcpy filter");
    d_fw = f;
  }
  // Fraction of nodes (by_id) or labels the filter allows.
  double allowed = 1.0;
  if (filter_words != nullptr) {
    const uint64_t bits = by_id ? h->view.count : filter_size;
    uint64_t ones = 0;
    for (uint64_t i = 0; i < bits / 64; ++i) ones += static_cast<uint64_t>(__builtin_popcountll(filter_words[i]));
    if (bits % 64 != 0) ones += static_cast<uint64_t>(__builtin_popcountll(filter_words[bits / 64] & ((1ULL << (bits % 64)) - 1)));
    allowed = bits > 0 ? static_cast<double>(ones) / static_cast<double>(bits) : 0.0;
  }
  search_device(*h, c, d_qs, d_scale, nq, k, ef, d_fw, filter_size, by_id, d_out, d_counts, st, allowed);
  check(hipMemcpyAsync(staged ? static_cast<void*>(back) : out, d_out, o_bytes, hipMemcpyDeviceToHost, st),
        "hipMemcpy results");
  if (counts != nullptr) {
    check(hipMemcpyAsync(staged ? static_cast<void*>(back + o_bytes) : counts, d_counts, c_bytes,
                         hipMemcpyDeviceToHost, st), "hipMemcpy counts");
  }
  dev::ctx_end(c, st);
  check(hipStreamSynchronize(st), "search_batch");
  drain.armed = false;
  if (staged) {
    std::memcpy(out, back, o_bytes);
    if (counts != nullptr) std::memcpy(counts, back + o_bytes, c_bytes);
  }
}

/// Device-buffer search (asynchronous on `stream`).
void device_search_dev(hnswgpu_index* h, const float* d_qs, uint64_t nq, uint64_t k, uint64_t ef,
                       const uint64_t* d_words, uint64_t filter_size, int by_id, hnswgpu_neighbor* d_out,
                       uint32_t* d_counts, void* stream) {
  if (h == nullptr) throw std::invalid_argument("search_batch: null handle");
  if (k == 0) throw std::invalid_argument("search_batch: k must be > 0");
  if (nq == 0) return;
  if (d_qs == nullptr || d_out == nullptr) throw std::invalid_argument("search_batch: null pointer");
  std::shared_lock<std::shared_mutex> lk(h->mu);
  if (h->view.count == 0) throw std::invalid_argument("search_batch: empty index on the device path");
  const auto st = static_cast<hipStream_t>(stream);
  dev::CtxLease lease(*h, st);
  dev::ctx_begin(*lease.c, st);
  search_device(*h, *lease.c, d_qs, nullptr, nq, k, ef, d_words, filter_size, by_id, d_out, d_counts, st);
  dev::ctx_end(*lease.c, st);
}

}  // namespace hnswamd

extern "C" {

int hnswgpu_search_batch(hnswgpu_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                         const uint64_t* filter_words, uint64_t filter_size, hnswgpu_neighbor* out,
                         uint32_t* counts, uint64_t /*num_threads*/) {
  return guarded([&] {
    hnswamd::device_search_host(h, qs, nq, k, ef, filter_words, filter_size, 0, out, counts);
  });
}

int hnswgpu_search_batch_ids(hnswgpu_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                             const uint64_t* id_words, hnswgpu_neighbor* out, uint32_t* counts,
                             uint64_t /*num_threads*/) {
  return guarded([&] {
    hnswamd::device_search_host(h, qs, nq, k, ef, id_words, 0, 1, out, counts);
  });
}

int hnswgpu_search_batch_device(hnswgpu_index* h, const float* d_qs, uint64_t nq, uint64_t k, uint64_t ef,
                                const uint64_t* d_filter_words, uint64_t filter_size,
                                hnswgpu_neighbor* d_out, uint32_t* d_counts, void* stream) {
  return guarded([&] {
    hnswamd::device_search_dev(h, d_qs, nq, k, ef, d_filter_words, filter_size, 0, d_out, d_counts, stream);
  });
}

int hnswgpu_search_batch_ids_device(hnswgpu_index* h, const float* d_qs, uint64_t nq, uint64_t k, uint64_t ef,
                                    const uint64_t* d_id_words, hnswgpu_neighbor* d_out, uint32_t* d_counts,
                                    void* stream) {
  return guarded([&] {
    hnswamd::device_search_dev(h, d_qs, nq, k, ef, d_id_words, 0, 1, d_out, d_counts, stream);
  });
}

int hnswgpu_search_tasks_device(hnswgpu_index* h, const hnswgpu_task* tasks, uint64_t ntasks,
                                const hnswgpu_filter* filters, uint64_t nfilters, int32_t storage, void* stream) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("search_tasks: null handle");
    if (ntasks == 0) return;
    std::shared_lock<std::shared_mutex> lk(h->mu);
    hnswamd::dev::CtxLease lease(*h);
    hnswamd::dev::CallCtx& c = *lease.c;
    // The pinned task staging of this context's previous call may still be in its copy.
    if (c.used) check(hipEventSynchronize(c.done), "hipEventSynchronize");
    auto* stage = static_cast<QueryTask*>(c.task_host.need(ntasks * sizeof(QueryTask)));
    auto* d_tasks = static_cast<QueryTask*>(c.task_dev.need(ntasks * sizeof(QueryTask)));
    const auto st = static_cast<hipStream_t>(stream);
    hnswamd::dev::ctx_begin(c, st);
    hnswamd::dev::launch_tasks(*h, c.sc, tasks, ntasks, filters, nfilters, storage < 0 ? h->view.storage : storage,
                               stage, d_tasks, h->spill_waves, st);
    hnswamd::dev::ctx_end(c, st);
  });
}

int hnswgpu_query_scale(const float* x, uint64_t dim, float* out) {
  return guarded([&] {
    if (x == nullptr || out == nullptr) throw std::invalid_argument("query_scale: null pointer");
    if (dim > 4096) throw std::invalid_argument("query_scale: dim must be <= 4096");
    const uint32_t bits = hnswamd::exact::inv_norm_bits(x, static_cast<uint32_t>(dim));
    std::memcpy(out, &bits, sizeof(bits));
  });
}

int hnswgpu_autotune(hnswgpu_index* h, const float* qs, uint64_t nq, uint64_t k, const uint64_t* efs, uint64_t nefs,
                     char* report, uint64_t report_len) {
  return guarded([&] {
    if (h == nullptr || qs == nullptr || efs == nullptr) throw std::invalid_argument("autotune: null pointer");
    if (nq == 0 || k == 0 || nefs == 0) throw std::invalid_argument("autotune: empty sample");
    std::unique_lock<std::shared_mutex> lk(h->mu);
    if (h->view.count == 0) throw std::invalid_argument("autotune: empty index");
    if (nq * h->dim > (1ull << 31) || nq * k > (1ull << 31)) throw std::invalid_argument("autotune: sample too large");
    hnswamd::dev::CtxLease lease(*h);
    hnswamd::dev::CallCtx& c = *lease.c;
    hipStream_t st = c.stream;
    hnswamd::dev::ctx_begin(c, st);
    auto* d_qs = static_cast<float*>(c.qs.need(nq * h->dim * sizeof(float)));
    auto* d_out = static_cast<hnswgpu_neighbor*>(c.out.need(nq * k * sizeof(hnswgpu_neighbor)));
    check(hipMemcpyAsync(d_qs, qs, nq * h->dim * sizeof(float), hipMemcpyHostToDevice, st), "hipMemcpy queries");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    const auto saved_use = h->use_tuned;
    const auto saved_timing = h->kernel_timing;
    h->kernel_timing = 0;
    std::string text;
    try {
      for (uint64_t i = 0; i < nefs; ++i) {
        const uint64_t ef = std::max<uint64_t>(efs[i], k);
        if (ef > hnswamd::gpu::kFastMaxEf) continue;  // the general kernel has no launch shape to tune
        const uint32_t cls = hnswamd::dev::ef_class_index(static_cast<uint32_t>(ef));
        hnswgpu_index::Tuned best;
        auto time_shape = [&](hnswgpu_index::Tuned cand) {
          cand.valid = true;
          h->tuned[cls] = cand;
          h->use_tuned = 1;
          search_device(*h, c, d_qs, nullptr, nq, k, ef, nullptr, 0, 0, d_out, nullptr, st);  // warm-up
          check(hipEventRecord(e0, st), "hipEventRecord");
          search_device(*h, c, d_qs, nullptr, nq, k, ef, nullptr, 0, 0, d_out, nullptr, st);
          check(hipEventRecord(e1, st), "hipEventRecord");
          check(hipEventSynchronize(e1), "hipEventSynchronize");
          float ms = 0.0f;
          check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
          cand.ms = ms;
          if (!best.valid || cand.ms < best.ms) best = cand;
          text += "ef=" + std::to_string(ef) + " waves_per_cu=" + std::to_string(cand.waves_per_cu) + " hash_log2=" +
                  std::to_string(cand.hash_log2) + " generic_gather=" + std::to_string(cand.generic_gather) +
                  " row_trip=" + std::to_string(cand.row_trip) + " min_class=" + std::to_string(cand.min_class) +
                  " ms=" + std::to_string(ms) + "\n";
        };
        // Stage 1: occupancy (wavefronts per CU), LDS visited-hash size and row loop, at the
        // storage's default rows per round trip and the ef's own capacity class.
        const int default_trip = h->row_trip != 0 ? h->row_trip : (h->view.storage != 0 ? 2 : 1);
        for (uint32_t wpc : {8u, 12u, 16u, 20u, 24u}) {
          for (uint32_t hl : {10u, 11u, 12u, 13u}) {
            for (int gather : {0, 1}) {
              if (gather == 1 && h->view.padded != 128) continue;  // no fixed-width loop to compare with
              hnswgpu_index::Tuned cand;
              cand.waves_per_cu = wpc;
              cand.hash_log2 = hl;
              cand.generic_gather = gather;
              cand.row_trip = default_trip;
              cand.min_class = cls;
              time_shape(cand);
            }
          }
        }
        // Stage 2, around the best shape: rows per memory round trip (the gather width:
        // 16 or 32 rows, fixed-width row loop only) and the capacity class (the ef's own,
        // or the next larger one: more register rows for the result and candidate arrays).
        const hnswgpu_index::Tuned base = best;
        if (h->view.padded == 128 && base.generic_gather == 0) {
          hnswgpu_index::Tuned cand = base;
          cand.row_trip = base.row_trip == 1 ? 2 : 1;
          time_shape(cand);
        }
        if (cls < 3) {
          hnswgpu_index::Tuned cand = best;
          cand.min_class = cls + 1;
          time_shape(cand);
        }
        h->tuned[cls] = best;
        text += "best ef=" + std::to_string(ef) + " waves_per_cu=" + std::to_string(best.waves_per_cu) +
                " hash_log2=" + std::to_string(best.hash_log2) + " generic_gather=" +
                std::to_string(best.generic_gather) + " row_trip=" + std::to_string(best.row_trip) +
                " min_class=" + std::to_string(best.min_class) + " ms=" + std::to_string(best.ms) + "\n";
      }
    } catch (...) {
      h->use_tuned = saved_use;
      h->kernel_timing = saved_timing;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    hnswamd::dev::ctx_end(c, st);
    h->use_tuned = 1;
    h->kernel_timing = saved_timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (report != nullptr && report_len > 0) {
      const size_t n = std::min<size_t>(text.size(), report_len - 1);
      std::memcpy(report, text.data(), n);
      report[n] = '\0';
    }
  });
}

int hnswgpu_set_option(hnswgpu_index* h, const char* name, int64_t value) {
  return guarded([&] {
    if (h == nullptr || name == nullptr) throw std::invalid_argument("set_option: null pointer");
    std::unique_lock<std::shared_mutex> lk(h->mu);
    const std::string n(name);
    if (n == "hash_log2") {
      if (value != 0 && (value < 8 || value > 15)) throw std::invalid_argument("hash_log2 must be 0 or 8..15");
      h->hash_log2 = static_cast<uint32_t>(value);
    } else if (n == "force_spill") {
      h->force_spill = value != 0 ? 1 : 0;
    } else if (n == "deleted_path") {
      if (value < 0 || value > 2) throw std::invalid_argument("deleted_path must be 0, 1 or 2");
      h->deleted_path = static_cast<int>(value);
    } else if (n == "storage") {
      // 0 = traverse the fp32 rows; 1 = a bf16 copy; 2 = an fp8 e4m3 copy with
      // per-vector scales. Reduced copies re-rank the final candidates in fp32.
      if (value < 0 || value > 2) throw std::invalid_argument("storage must be 0 (fp32), 1 (bf16) or 2 (fp8)");
      hnswamd::dev::build_reduced(*h, static_cast<int>(value));
      h->view = hnswamd::dev::view_for(*h, static_cast<int>(value));
    } else if (n == "visit_log2") {
      if (value < 8 || value > 24) throw std::invalid_argument("visit_log2 must be 8..24");
      h->visit_log2 = static_cast<uint32_t>(value);
      std::lock_guard<std::mutex> pk(h->pool_mu);
      for (auto& c : h->ctxs) {
        c->sc.visit_tables.release();
        c->sc.visit_tables_cleared = 0;
      }
    } else if (n == "waves_per_cu") {
      if (value < 1 || value > 32) throw std::invalid_argument("waves_per_cu must be 1..32");
      h->waves_per_cu = static_cast<uint32_t>(value);
    } else if (n == "generic_gather") {
      h->generic_gather = value != 0 ? 1 : 0;
    } else if (n == "row_trip") {
      if (value < 0 || value > 2) throw std::invalid_argument("row_trip must be 0 (default), 1 or 2");
      h->row_trip = static_cast<int>(value);
    } else if (n == "min_class") {
      if (value < 0 || value > 3) throw std::invalid_argument("min_class must be 0..3");
      h->min_class = static_cast<uint32_t>(value);
    } else if (n == "use_tuned") {
      h->use_tuned = value != 0 ? 1 : 0;
    } else if (n == "wide_hash") {
      h->wide_hash = value != 0 ? 1 : 0;
    } else if (n == "collect_stats") {
      h->collect_stats = value != 0 ? 1 : 0;
    } else if (n == "kernel_timing") {
      h->kernel_timing = value != 0 ? 1 : 0;
    } else if (n == "spill_waves") {
      if (value < 1 || value > 4096) throw std::invalid_argument("spill_waves must be 1..4096");
      h->spill_waves = static_cast<uint32_t>(value);
    } else if (n == "max_contexts") {
      if (value < 1 || value > 64) throw std::invalid_argument("max_contexts must be 1..64");
      std::lock_guard<std::mutex> pk(h->pool_mu);
      h->max_ctxs = static_cast<uint32_t>(value);
    } else {
      throw std::invalid_argument("unknown option " + n);
    }
  });
}

int hnswgpu_get_option(hnswgpu_index* h, const char* name, int64_t* value) {
  return guarded([&] {
    if (h == nullptr || name == nullptr || value == nullptr) throw std::invalid_argument("get_option: null pointer");
    std::shared_lock<std::shared_mutex> lk(h->mu);
    const std::string n(name);
    hnswamd::dev::CallCtx* last = nullptr;
    {
      std::lock_guard<std::mutex> pk(h->pool_mu);
      last = h->last;
    }
    if (n.rfind("last_", 0) == 0 && n != "last_kernel_ns") {
      // Counters of the most recent search on this handle (waits for it).
      uint32_t ctr[8] = {};
      if (last != nullptr && last->sc.counters.p != nullptr) {
        check(hipEventSynchronize(last->done), "hipEventSynchronize");
        check(hipMemcpy(ctr, last->sc.counters.p, sizeof(ctr), hipMemcpyDeviceToHost), "hipMemcpy counters");
      }
      if (n == "last_overflow") {
        *value = ctr[1];
      } else if (n == "last_hbm_visited") {
        *value = ctr[3];
      } else if (n == "last_handoff_candidates") {
        *value = ctr[4];
      } else if (n == "last_handoff_visited") {
        *value = ctr[5];
      } else {
        throw std::invalid_argument("unknown option " + n);
      }
    } else if (n == "hash_log2") {
      *value = h->hash_log2;
    } else if (n == "force_spill") {
      *value = h->force_spill;
    } else if (n == "deleted_path") {
      *value = h->deleted_path;
    } else if (n == "spill_waves") {
      *value = h->spill_waves;
    } else if (n == "visit_log2") {
      *value = h->visit_log2;
    } else if (n == "waves_per_cu") {
      *value = h->waves_per_cu;
    } else if (n == "storage") {
      *value = h->view.storage;
    } else if (n == "collect_stats") {
      *value = h->collect_stats;
    } else if (n == "kernel_timing") {
      *value = h->kernel_timing;
    } else if (n == "max_contexts") {
      *value = h->max_ctxs;
    } else if (n == "servers") {
      *value = h->servers.load();
    } else if (n == "contexts") {
      std::lock_guard<std::mutex> pk(h->pool_mu);
      *value = static_cast<int64_t>(h->ctxs.size());
    } else if (n == "last_kernel_ns") {
      // Fast-kernel time of the last search, from events recorded on its stream.
      if (last == nullptr || !last->timing_pending) {
        throw std::invalid_argument("last_kernel_ns: enable kernel_timing and search first");
      }
      check(hipEventSynchronize(last->ev_end), "hipEventSynchronize");
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, last->ev_begin, last->ev_end), "hipEventElapsedTime");
      *value = static_cast<int64_t>(static_cast<double>(ms) * 1e6);
    } else if (n == "generic_gather") {
      *value = h->generic_gather;
    } else if (n == "row_trip") {
      *value = h->row_trip;
    } else if (n == "min_class") {
      *value = h->min_class;
    } else if (n == "use_tuned") {
      *value = h->use_tuned;
    } else if (n == "wide_hash") {
      *value = h->wide_hash;
    } else if (n.rfind("tuned_", 0) == 0) {
      // tuned_<field>_<ef class>: the autotuner's choice, or -1 if that class was not tuned.
      const size_t cut = n.rfind('_');
      const std::string field = n.substr(6, cut - 6);
      const int cls = std::atoi(n.c_str() + cut + 1);
      const int idx = cls == 64 ? 0 : cls == 128 ? 1 : cls == 256 ? 2 : cls == 512 ? 3 : -1;
      if (idx < 0) throw std::invalid_argument("unknown option " + n);
      const hnswgpu_index::Tuned& t = h->tuned[idx];
      if (field == "waves_per_cu") {
        *value = t.valid ? static_cast<int64_t>(t.waves_per_cu) : -1;
      } else if (field == "hash_log2") {
        *value = t.valid ? static_cast<int64_t>(t.hash_log2) : -1;
      } else if (field == "generic_gather") {
        *value = t.valid ? static_cast<int64_t>(t.generic_gather) : -1;
      } else if (field == "row_trip") {
        *value = t.valid ? static_cast<int64_t>(t.row_trip) : -1;
      } else if (field == "min_class") {
        *value = t.valid ? static_cast<int64_t>(t.min_class) : -1;
      } else if (field == "us") {
        *value = t.valid ? static_cast<int64_t>(t.ms * 1e3) : -1;
      } else {
        throw std::invalid_argument("unknown option " + n);
      }
    } else if (n == "has_dups") {
      *value = h->view.has_dups;
    } else if (n == "num_deleted") {
      *value = static_cast<int64_t>(h->num_deleted);
    } else {
      throw std::invalid_argument("unknown option " + n);
    }
  });
}

}  // extern "C"

extern "C" int hnswgpu_read_stats(hnswgpu_index* h, uint32_t* out, uint64_t nq) {
  return guarded([&] {
    if (h == nullptr || out == nullptr) throw std::invalid_argument("read_stats: null pointer");
    std::shared_lock<std::shared_mutex> lk(h->mu);
    hnswamd::dev::CallCtx* last = nullptr;
    {
      std::lock_guard<std::mutex> pk(h->pool_mu);
      last = h->last;
    }
    if (last == nullptr || !last->stats_on || last->sc.stats.p == nullptr) {
      throw std::invalid_argument("read_stats: collect_stats is off");
    }
    if (nq > last->last_nq) throw std::invalid_argument("read_stats: more queries than the last search had");
    check(hipEventSynchronize(last->done), "hipEventSynchronize");
    check(hipMemcpy(out, last->sc.stats.p, nq * 3 * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy stats");
  });
}

namespace {
__global__ void copy_words(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    dst[i] = src[i];
  }
}
}  // namespace

// ---------------------------------------------------------------------------
// Device plumbing for callers without a GPU framework of their own: buffers,
// streams and events of this library's HIP runtime.
// ---------------------------------------------------------------------------
extern "C" {

int hnswgpu_malloc(uint64_t bytes, void** out) {
  return guarded([&] {
    if (out == nullptr) throw std::invalid_argument("malloc: null out");
    check(hipMalloc(out, std::max<uint64_t>(bytes, 1)), "hipMalloc");
  });
}

int hnswgpu_mfree(void* p) {
  return guarded([&] { check(hipFree(p), "hipFree"); });
}

int hnswgpu_memcpy(void* dst, const void* src, uint64_t bytes, int kind) {
  return guarded([&] {
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    check(hipMemcpy(dst, src, bytes, k), "hipMemcpy");
  });
}

int hnswgpu_copy_kernel(void* dst, const void* src, uint64_t bytes) {
  return guarded([&] {
    if (bytes % 16 != 0) throw std::invalid_argument("copy_kernel: bytes must be a multiple of 16");
    const size_t n = bytes / 16;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 65536));
    if (n == 0) return;
    hipLaunchKernelGGL(copy_words, dim3(blocks), dim3(256), 0, nullptr, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n);
    check(hipGetLastError(), "copy kernel launch");
    check(hipDeviceSynchronize(), "copy kernel");
  });
}

int hnswgpu_stream_create(void** out) {
  return guarded([&] {
    hipStream_t s = nullptr;
    check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *out = s;
  });
}

int hnswgpu_stream_destroy(void* s) {
  return guarded([&] { check(hipStreamDestroy(static_cast<hipStream_t>(s)), "hipStreamDestroy"); });
}

int hnswgpu_stream_synchronize(void* s) {
  return guarded([&] { check(hipStreamSynchronize(static_cast<hipStream_t>(s)), "hipStreamSynchronize"); });
}

int hnswgpu_event_create(void** out) {
  return guarded([&] {
    hipEvent_t e = nullptr;
    check(hipEventCreate(&e), "hipEventCreate");
    *out = e;
  });
}

int hnswgpu_event_destroy(void* e) {
  return guarded([&] { check(hipEventDestroy(static_cast<hipEvent_t>(e)), "hipEventDestroy"); });
}

int hnswgpu_event_record(void* e, void* s) {
  return guarded([&] {
    check(hipEventRecord(static_cast<hipEvent_t>(e), static_cast<hipStream_t>(s)), "hipEventRecord");
  });
}

int hnswgpu_event_elapsed_ms(void* begin, void* end, float* ms) {
  return guarded([&] {
    check(hipEventSynchronize(static_cast<hipEvent_t>(end)), "hipEventSynchronize");
    check(hipEventElapsedTime(ms, static_cast<hipEvent_t>(begin), static_cast<hipEvent_t>(end)),
          "hipEventElapsedTime");
  });
}

int hnswgpu_device_info(char* name, int len, int64_t* info) {
  return guarded([&] {
    hipDeviceProp_t p{};
    check(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties");
    if (name != nullptr && len > 0) std::snprintf(name, static_cast<size_t>(len), "%s", p.gcnArchName);
    if (info != nullptr) {
      info[0] = p.multiProcessorCount;
      info[1] = static_cast<int64_t>(p.totalGlobalMem);
      info[2] = static_cast<int64_t>(p.sharedMemPerMultiprocessor);
      info[3] = p.clockRate;
      info[4] = p.memoryClockRate;
      info[5] = p.memoryBusWidth;
    }
  });
}

}  // extern "C"

extern "C" int hnswgpu_get_row(hnswgpu_index* h, uint32_t id, int storage, void* out, float* scale_norm) {
  return guarded([&] {
    if (h == nullptr || out == nullptr) throw std::invalid_argument("get_row: null pointer");
    std::shared_lock<std::shared_mutex> lk(h->mu);
    if (id >= h->view.count) throw std::out_of_range("get_row: id out of range");
    const size_t p = h->view.padded;
    if (storage == 0) {
      check(hipMemcpy(out, h->view.vecs + id * p, p * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy row");
      return;
    }
    const DevIndexView v = hnswamd::dev::view_for(*h, storage);
    const size_t width = storage == 1 ? 2 : 1;
    check(hipMemcpy(out, static_cast<const uint8_t*>(v.vlow) + id * p * width, p * width,
                    hipMemcpyDeviceToHost), "hipMemcpy row");
    if (scale_norm != nullptr) {
      float sn[2] = {1.0f, 0.0f};
      if (storage == 2) check(hipMemcpy(&sn[0], v.vscale + id, sizeof(float), hipMemcpyDeviceToHost), "scale");
      check(hipMemcpy(&sn[1], v.vnorm + id, sizeof(float), hipMemcpyDeviceToHost), "norm");
      scale_norm[0] = sn[0];
      scale_norm[1] = sn[1];
    }
  });
}
