// Device probe of the wave builder (host/graph_build.cpp): phase 1 of an
// insertion wave with the layer-0 beams of the wave's level-0 nodes searched
// on the GPU.
//
// A level-0 node's phase 1 is exactly a query search: greedy descent from the
// frozen entry point to layer 1, then Alg. 2 on layer 0 with ef =
// ef_construction from the node found, returning the whole beam sorted by
// (distance, id). So it runs on the search kernels unchanged, with the node's
// stored (already prepared) row as the query, a cosine scale of exactly 1,
// k = ef = ef_construction, and labels that are the node ids. The kernels
// compute the host's summation order and Alg. 2's order and rules bit for bit,
// so the beams equal the host's and the graph equals a host build with the
// same waves. Nodes of higher level (about 1 in M) start their layer-0 beam
// from several seeds; the host probes those while the GPU runs.
//
// The session holds a device copy of the growing graph: every row of the
// batch (appended before the first wave), identity labels, and the lists,
// refreshed before each wave from the ids the previous wave changed. Every
// refreshed list is checked against the frozen node count and the levels
// first, so the kernels' accesses stay in bounds.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../host/graph.hpp"
#include "../host/graph_build.hpp"
#include "build_probe.hpp"
#include "index_state.hpp"
#include "launch.hpp"

namespace hnswamd {
namespace {

using dev::check;
using dev::DevBuf;
using dev::PinnedBuf;

/// rows[r] (stride u32 words) -> l0 row ids[r].
__global__ void scatter_rows(uint32_t* __restrict__ l0, const uint32_t* __restrict__ rows,
                             const uint32_t* __restrict__ ids, uint32_t n, uint32_t stride) {
  const uint64_t total = static_cast<uint64_t>(n) * stride;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t r = i / stride;
    l0[static_cast<uint64_t>(ids[r]) * stride + (i - r * stride)] = rows[i];
  }
}

/// out[r] = the first dim floats of stored row ids[r] (the search's query layout).
__global__ void gather_rows(const float* __restrict__ vecs, uint32_t padded, uint32_t dim,
                            const uint32_t* __restrict__ ids, uint32_t n, float* __restrict__ out) {
  const uint64_t total = static_cast<uint64_t>(n) * dim;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t r = i / dim;
    out[i] = vecs[static_cast<uint64_t>(ids[r]) * padded + (i - r * dim)];
  }
}

__global__ void fill_value(float* __restrict__ p, uint32_t n, float v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

uint32_t grid_for(uint64_t items) {
  return static_cast<uint32_t>(std::min<uint64_t>(4096, std::max<uint64_t>(1, (items + 255) / 256)));
}

double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

class Session {
 public:
  Session(const Graph& g, uint32_t frozen, DeviceProbeStats* stats, Team& team) : stats_(stats) {
    if (g.padded_dim > 4096) throw std::invalid_argument("add_batch_device: padded_dim must be <= 4096");
    if (2 * g.M > gpu::kMaxDegree0) throw std::invalid_argument("add_batch_device: M must be <= 32");
    if (g.count >= 0xFFFFFFFFull || g.upper_used >= 0xFFFFFFFFull) {
      throw std::invalid_argument("add_batch_device: index too large for 32-bit device offsets");
    }
    h_ = std::make_unique<hnswgpu_index>();
    check(hipStreamCreateWithFlags(&h_->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreate(&ev0_), "hipEventCreate");
    check(hipEventCreate(&ev1_), "hipEventCreate");
    const uint64_t n = g.count;
    const hipStream_t st = h_->stream;
    // Every list of the frozen graph is checked once; lists of nodes not yet
    // linked are empty.
    const uint64_t chunk = 4096;
    team.run((n + chunk - 1) / chunk, [&](std::size_t c, std::size_t) {
      for (uint64_t id = c * chunk; id < std::min<uint64_t>(n, (c + 1) * chunk); ++id) {
        check_lists(g, static_cast<uint32_t>(id), frozen);
      }
    });
    h_->view.has_dups = dups_.load() ? 1 : 0;
    auto put = [&](DevBuf& b, const void* src, size_t bytes) {
      void* d = b.need(std::max<size_t>(bytes, 4));
      if (bytes > 0) check(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, st), "hipMemcpy upload");
      return d;
    };
    std::vector<uint64_t> ids(n);
    for (uint64_t i = 0; i < n; ++i) ids[i] = i;
    std::vector<uint32_t> off(n);
    for (uint64_t i = 0; i < n; ++i) off[i] = static_cast<uint32_t>(g.upper_offset[i]);
    gpu::DevIndexView& v = h_->view;
    v.vecs = static_cast<const float*>(put(h_->vecs, g.vecs, n * g.padded_dim * sizeof(float)));
    v.l0 = static_cast<const uint32_t*>(put(h_->l0, g.l0, n * g.l0_stride * sizeof(uint32_t)));
    v.upper = static_cast<const uint32_t*>(put(h_->upper, g.upper, g.upper_used * sizeof(uint32_t)));
    v.upper_off = static_cast<const uint32_t*>(put(h_->upper_off, off.data(), n * sizeof(uint32_t)));
    v.labels = static_cast<const uint64_t*>(put(h_->labels, ids.data(), n * sizeof(uint64_t)));
    v.count = static_cast<uint32_t>(n);
    v.padded = static_cast<uint32_t>(g.padded_dim);
    v.M = static_cast<uint32_t>(g.M);
    v.M0 = static_cast<uint32_t>(g.M0);
    v.metric = g.params.metric;
    v.storage = 0;
    h_->dim = g.dim;
    h_->upper_used = g.upper_used;
    h_->deleted_host.assign(n, 0);
    ones_ = static_cast<float*>(ones_buf_.need(kMaxWave * sizeof(float)));
    hipLaunchKernelGGL(fill_value, dim3((kMaxWave + 255) / 256), dim3(256), 0, st, ones_,
                       static_cast<uint32_t>(kMaxWave), 1.0f);
    check(hipGetLastError(), "fill kernel");
    check(hipStreamSynchronize(st), "upload");  // the host vectors above go out of scope
  }

  ~Session() {
    if (h_ && h_->stream != nullptr) (void)hipStreamSynchronize(h_->stream);
    if (ev0_ != nullptr) (void)hipEventDestroy(ev0_);
    if (ev1_ != nullptr) (void)hipEventDestroy(ev1_);
  }

  /// Lists changed by the previous wave: checked, staged and scattered into
  /// the device layer-0 array; the (small) upper section is copied whole.
  void refresh(const Graph& g, const std::vector<uint32_t>& changed, uint32_t frozen, Team& team) {
    const auto t0 = std::chrono::steady_clock::now();
    const hipStream_t st = h_->stream;
    check(hipStreamSynchronize(st), "refresh");  // staging buffers are reused
    const size_t m = changed.size();
    if (m > 0) {
      auto* ids = static_cast<uint32_t*>(stage_changed_.need(m * sizeof(uint32_t)));
      auto* rows = static_cast<uint32_t*>(stage_rows_.need(m * g.l0_stride * sizeof(uint32_t)));
      const size_t chunk = 1024;
      team.run((m + chunk - 1) / chunk, [&](std::size_t c, std::size_t) {
        for (size_t i = c * chunk; i < std::min(m, (c + 1) * chunk); ++i) {
          const uint32_t id = changed[i];
          if (id >= h_->view.count) throw std::logic_error("add_batch_device: changed id out of range");
          check_lists(g, id, frozen);
          ids[i] = id;
          std::copy(g.l0 + static_cast<size_t>(id) * g.l0_stride,
                    g.l0 + static_cast<size_t>(id + 1) * g.l0_stride, rows + i * g.l0_stride);
        }
      });
      if (dups_.load()) h_->view.has_dups = 1;
      auto* d_ids = static_cast<uint32_t*>(d_changed_.need(m * sizeof(uint32_t)));
      auto* d_rows = static_cast<uint32_t*>(d_rows_.need(m * g.l0_stride * sizeof(uint32_t)));
      check(hipMemcpyAsync(d_ids, ids, m * sizeof(uint32_t), hipMemcpyHostToDevice, st), "hipMemcpy ids");
      check(hipMemcpyAsync(d_rows, rows, m * g.l0_stride * sizeof(uint32_t), hipMemcpyHostToDevice, st),
            "hipMemcpy rows");
      hipLaunchKernelGGL(scatter_rows, dim3(grid_for(m * g.l0_stride)), dim3(256), 0, st,
                         static_cast<uint32_t*>(h_->l0.p), d_rows, d_ids, static_cast<uint32_t>(m),
                         static_cast<uint32_t>(g.l0_stride));
      check(hipGetLastError(), "scatter kernel");
    }
    if (g.upper_used > 0) {
      check(hipMemcpyAsync(h_->upper.p, g.upper, g.upper_used * sizeof(uint32_t), hipMemcpyHostToDevice, st),
            "hipMemcpy upper");
    }
    if (stats_ != nullptr) stats_->refresh_ms += ms_since(t0);
  }

  /// Phase 1 of wave [s, e) on the frozen graph `f`. Its buffers are apart
  /// from refresh()'s, whose copies may still be in flight.
  void probe(const Graph& g, const WaveFrame& f, uint32_t s, uint32_t e,
             std::vector<std::vector<std::vector<Key>>>& found, Team& team) {
    std::vector<uint32_t> on_gpu, on_host;
    for (uint32_t id = s; id < e; ++id) (g.levels[id] == 0 ? on_gpu : on_host).push_back(id);
    const hipStream_t st = h_->stream;
    const uint32_t nq = static_cast<uint32_t>(on_gpu.size());
    const uint32_t k = static_cast<uint32_t>(g.efc);
    if (nq > kMaxWave) throw std::logic_error("add_batch_device: wave larger than the probe's buffers");
    gpu::OutRow* out = nullptr;
    if (nq > 0) {
      if (f.entry >= h_->view.count || f.top < 0 || g.levels[f.entry] != f.top) {
        throw std::logic_error("add_batch_device: invalid entry point");
      }
      h_->view.entry = f.entry;
      h_->view.max_level = f.top;
      auto* ids = static_cast<uint32_t*>(stage_ids_.need(nq * sizeof(uint32_t)));
      std::copy(on_gpu.begin(), on_gpu.end(), ids);
      auto* d_ids = static_cast<uint32_t*>(d_ids_.need(nq * sizeof(uint32_t)));
      auto* d_q = static_cast<float*>(d_q_.need(static_cast<size_t>(nq) * g.dim * sizeof(float)));
      auto* d_out = static_cast<gpu::OutRow*>(d_out_.need(static_cast<size_t>(nq) * k * sizeof(gpu::OutRow)));
      out = static_cast<gpu::OutRow*>(stage_out_.need(static_cast<size_t>(nq) * k * sizeof(gpu::OutRow)));
      check(hipMemcpyAsync(d_ids, ids, nq * sizeof(uint32_t), hipMemcpyHostToDevice, st), "hipMemcpy ids");
      hipLaunchKernelGGL(gather_rows, dim3(grid_for(static_cast<uint64_t>(nq) * g.dim)), dim3(256), 0, st,
                         h_->view.vecs, h_->view.padded, static_cast<uint32_t>(g.dim), d_ids, nq, d_q);
      check(hipGetLastError(), "gather kernel");
      gpu::SearchRequest r{};
      r.queries = d_q;
      r.qscale = h_->view.metric == 2 ? ones_ : nullptr;  // stored rows are already unit length
      r.dim = static_cast<uint32_t>(g.dim);
      r.nq = nq;
      r.k = k;
      r.ef = k;
      r.out = d_out;
      dev::plan_request(*h_, r, nq);
      gpu::SearchScratch sc = dev::bind_scratch(*h_, scratch_, r, nq, h_->spill_waves, st);
      check(hipEventRecord(ev0_, st), "hipEventRecord");
      check(gpu::launch_prepare(h_->view, r, sc, st), "prepare kernel");
      check(gpu::launch_search(h_->view, r, sc, st), "search kernel");
      check(hipEventRecord(ev1_, st), "hipEventRecord");
      check(hipMemcpyAsync(out, d_out, static_cast<size_t>(nq) * k * sizeof(gpu::OutRow), hipMemcpyDeviceToHost, st),
            "hipMemcpy beams");
    }
    // The other nodes on the host meanwhile.
    team.run(on_host.size(), [&](std::size_t i, std::size_t) {
      probe_on_host(g, f, on_host[i], found[on_host[i] - s]);
    });
    if (nq == 0) return;
    check(hipStreamSynchronize(st), "probe");
    const size_t chunk = 256;
    team.run((nq + chunk - 1) / chunk, [&](std::size_t c, std::size_t) {
      for (size_t q = c * chunk; q < std::min<size_t>(nq, (c + 1) * chunk); ++q) {
        std::vector<std::vector<Key>>& by_layer = found[on_gpu[q] - s];
        by_layer.assign(1, {});
        std::vector<Key>& beam = by_layer[0];
        beam.reserve(k);
        for (uint32_t j = 0; j < k; ++j) {
          const gpu::OutRow& o = out[q * k + j];
          if (o.label == ~0ULL) break;
          beam.push_back(Key{o.distance, static_cast<uint32_t>(o.label)});
        }
      }
    });
    if (stats_ != nullptr) {
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, ev0_, ev1_), "hipEventElapsedTime");
      stats_->device_ms += ms;
      stats_->device_nodes += nq;
      stats_->host_nodes += on_host.size();
      stats_->waves += 1;
    }
  }

 private:
  /// Every list of `id` holds at most its layer's degree of ids below
  /// `frozen` that reach that layer; a node at or above `frozen` has none. A
  /// layer-0 list that repeats an id switches the kernels to their
  /// duplicate-safe visited marking (the builder writes none).
  void check_lists(const Graph& g, uint32_t id, uint32_t frozen) {
    for (int L = 0; L <= static_cast<int>(g.levels[id]); ++L) {
      const uint32_t* l = g.list(id, L);
      if (l[0] > g.degree_cap(L) || (id >= frozen && l[0] != 0)) {
        throw std::logic_error("add_batch_device: list of node " + std::to_string(id) + " is malformed");
      }
      for (uint32_t j = 0; j < l[0]; ++j) {
        const uint32_t nb = l[1 + j];
        if (nb >= frozen || static_cast<int>(g.levels[nb]) < L) {
          throw std::logic_error("add_batch_device: list of node " + std::to_string(id) + " leaves the frozen graph");
        }
      }
      if (L == 0 && l[0] > 1) {
        uint32_t ids[gpu::kMaxDegree0];
        std::copy(l + 1, l + 1 + l[0], ids);
        std::sort(ids, ids + l[0]);
        if (std::adjacent_find(ids, ids + l[0]) != ids + l[0]) dups_.store(true);
      }
    }
  }

  std::unique_ptr<hnswgpu_index> h_;
  dev::Scratch scratch_;
  DevBuf ones_buf_, d_changed_, d_rows_, d_ids_, d_q_, d_out_;
  PinnedBuf stage_changed_, stage_rows_, stage_ids_, stage_out_;
  float* ones_ = nullptr;
  std::atomic<bool> dups_{false};
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  DeviceProbeStats* stats_;
};

}  // namespace

ProbeFn make_device_probe(uint32_t min_device_wave, DeviceProbeStats* stats) {
  auto session = std::make_shared<std::unique_ptr<Session>>();
  return [session, min_device_wave, stats](const Graph& g, const WaveFrame& f, uint32_t s, uint32_t e,
                                           std::vector<std::vector<std::vector<Key>>>& found, Team& team) {
    std::unique_ptr<Session>& ses = *session;
    if (!ses && (e - s < min_device_wave || f.entry == kNoNode)) {
      team.run(e - s, [&](std::size_t i, std::size_t) {
        probe_on_host(g, f, s + static_cast<uint32_t>(i), found[i]);
      });
      if (stats != nullptr) stats->host_nodes += e - s;
      return;
    }
    if (!ses) {
      ses = std::make_unique<Session>(g, s, stats, team);
    } else {
      ses->refresh(g, *f.changed, s, team);
    }
    ses->probe(g, f, s, e, found, team);
  };
}

}  // namespace hnswamd
