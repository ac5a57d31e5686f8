// Serving loop (hnswgpu_server_*): continuous batching of search requests over
// a resident index.
//
// Requests (host buffers; any number of queries, k, ef, filter and storage)
// are queued by submit. A dispatcher thread forms launches: whenever one of the
// slots is free it takes pending requests in arrival order -- those with the
// storage of the oldest one -- until the slot's capacity (queries, result rows,
// distinct filters) is reached, splitting a request that does not fit, and
// enqueues on the slot's stream: one host-to-device copy of the slot's staging
// area, the eligibility bitmaps of the launch's filters, the mixed-task search
// (device work queue, one task per query), and one device-to-host copy of the
// results. With graphs on, that sequence is a hipGraph captured once per slot
// and launch shape and replayed with the two copy sizes updated; the number of
// queries travels in the staging area. A completion thread waits for the slots
// in launch order, copies each request's rows to its buffers and records its
// latency. Every slot has its own stream and scratch (visited tables included),
// so a launch formed while others run overlaps them on the device: its blocks
// take the CUs their tails leave idle.
//
// Device memory per slot is sized at creation from the configuration and
// checked against the budget (default: a quarter of the free HBM); the slot
// count is reduced to fit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <tuple>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../../include/hnswgpu.h"
#include "../error.hpp"
#include "index_state.hpp"
#include "launch.hpp"

namespace {

using hnswamd::dev::check;
using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Request {
  uint64_t ticket = 0;
  hnswgpu_request r{};
  int storage = 0;
  double cost = 1.0;    // estimated cost of one of its queries (max(ef, k) / filter density)
  bool heavy = false;   // cost >= heavy_cost: launched only with other heavy requests
  uint64_t placed = 0;  // queries placed in launches
  uint64_t done = 0;    // queries completed
  int status = HNSWGPU_OK;
  std::string error;
  bool finished = false;
  Clock::time_point t_submit;
  double latency_ms = 0.0;  // submit to completion, set when finished
};

/// A run of one request's queries inside a launch.
struct Piece {
  std::shared_ptr<Request> req;
  uint64_t first;    // first query of the request
  uint32_t n;        // queries
  uint32_t task;     // first task index in the launch
  uint64_t row;      // first result row in the launch
};

struct Graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t h2d = nullptr;
  hipGraphNode_t d2h = nullptr;
  std::vector<uint64_t> shape;  // launch parameters baked into the graph's kernel nodes
};

void destroy_graph(Graph& g) {
  if (g.exec != nullptr) (void)hipGraphExecDestroy(g.exec);
  if (g.graph != nullptr) (void)hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
}

/// Every value a captured launch keeps in its kernel arguments: the launch
/// shape, the scratch pointers (which bind_scratch may reallocate when an
/// option or the tuned shapes change) and the index view. A graph is replayed
/// only while these are unchanged; otherwise it is captured again.
std::vector<uint64_t> shape_words(const hnswamd::gpu::SearchRequest& r, const hnswamd::gpu::SearchScratch& s,
                                  const hnswamd::gpu::DevIndexView& ix, const uint32_t* del_bits) {
  auto P = [](const void* p) { return static_cast<uint64_t>(reinterpret_cast<uintptr_t>(p)); };
  return {P(r.tasks), P(r.nq_dev), P(r.queries), P(r.qscale), r.dim, r.nq, r.k, r.ef, P(r.elig),
          static_cast<uint64_t>(r.deleted_only), P(r.out), P(r.counts), P(r.stats), r.hash_log2, r.waves_per_block,
          r.blocks, static_cast<uint64_t>(r.force_spill), static_cast<uint64_t>(r.generic_gather),
          static_cast<uint64_t>(r.wide_hash), static_cast<uint64_t>(r.row_trip), r.min_class,
          P(s.counters), P(s.visit_tables), s.visit_log2, P(s.overflow), P(s.qprep), P(s.spill_words),
          P(s.spill_heaps), s.spill_heap_stride, s.spill_waves, s.spill_bitmap_words, P(s.cand_runs),
          P(ix.vecs), P(ix.l0), P(ix.upper), P(ix.upper_off), P(ix.labels), ix.count, ix.padded, ix.M, ix.M0,
          ix.entry, static_cast<uint64_t>(static_cast<int64_t>(ix.max_level)), static_cast<uint64_t>(ix.metric),
          ix.has_dups, static_cast<uint64_t>(ix.storage), P(ix.vlow), P(ix.vscale), P(ix.vnorm), P(del_bits)};
}

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hnswamd::dev::Scratch sc;
  hnswamd::dev::PinnedBuf stage, result;
  hnswamd::dev::DevBuf in, out, elig;
  std::map<std::tuple<int, uint32_t, int>, Graph> graphs;  // (storage, ef class, deletions-only kernel)
  std::vector<Piece> pieces;
  bool heavy = false;     // the launch holds heavy requests
  bool launched = false;  // a launch is in flight (its completion thread owns the slot)
  std::thread completer;
  uint32_t nq = 0;
  uint64_t rows = 0;
  std::string error;  // launch failure, reported to the launch's requests
};

}  // namespace

struct hnswgpu_server {
  hnswgpu_index* h = nullptr;
  hnswgpu_server_config cfg{};
  uint64_t dim = 0;
  uint32_t padded = 0;
  uint64_t filter_words_max = 0;
  // Layout of a slot's input area (device and staging alike):
  //   [0, 256)          nq (u32)
  //   descs_off         FilterDesc x (max_filters + 1)  (last one: deletions only)
  //   tasks_off         QueryTask x max_batch
  //   var_off           queries (nq x dim f32), then the filter words of the launch
  size_t descs_off = 0, tasks_off = 0, var_off = 0, in_bytes = 0;
  // Output area: [0, counts_bytes) counts (u32 x max_batch), then rows (16 B x max_rows).
  size_t counts_bytes = 0, out_bytes = 0;
  size_t words = 0;  // eligibility bitmap words per filter
  size_t slot_bytes = 0;

  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_req;
  std::deque<std::shared_ptr<Request>> pending;
  std::deque<Slot*> free_slots;
  uint32_t heavy_inflight = 0;  // launches of heavy requests in flight
  std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
  uint64_t next_ticket = 1;
  bool stopping = false;
  std::thread dispatcher;

  // statistics since the last reset
  std::vector<double> latencies_ms;
  uint64_t queries_done = 0, launches = 0, graph_launches = 0, launched_queries = 0, recaptures = 0;
  bool any_submit = false;
  Clock::time_point t_first, t_last;
  double device_busy_s = 0.0;

  ~hnswgpu_server() {
    for (auto& s : slots) {
      for (auto& kv : s->graphs) destroy_graph(kv.second);
      if (s->done != nullptr) (void)hipEventDestroy(s->done);
      if (s->stream != nullptr) (void)hipStreamDestroy(s->stream);
    }
  }
};

namespace {

using hnswamd::gpu::FilterDesc;
using hnswamd::gpu::QueryTask;

uint32_t ef_class(uint64_t max_ef) {
  if (max_ef <= 64) return 64;
  if (max_ef <= 128) return 128;
  if (max_ef <= 256) return 256;
  return hnswamd::gpu::kFastMaxEf;  // larger tasks go to the general kernel one by one
}

/// Take pending requests into `slot` (caller holds srv.mu). Returns the
/// launch's filters (host words pointer, size) in descriptor order.
std::vector<std::pair<const uint64_t*, uint64_t>> form_launch(hnswgpu_server& srv, Slot& slot, int& storage,
                                                              uint64_t& max_ef) {
  const hnswgpu_server_config& c = srv.cfg;
  slot.pieces.clear();
  slot.nq = 0;
  slot.rows = 0;
  std::vector<std::pair<const uint64_t*, uint64_t>> filters;
  size_t var_used = 0;
  // The oldest request that may start a launch now: heavy ones only while a
  // slot stays free for the others.
  const bool heavy_ok = srv.heavy_inflight + 1 < srv.cfg.slots || srv.cfg.slots == 1;
  const Request* first = nullptr;
  for (const auto& p : srv.pending) {
    if (!p->heavy || heavy_ok) {
      first = p.get();
      break;
    }
  }
  if (first == nullptr) return {};
  storage = first->storage;
  slot.heavy = first->heavy;
  max_ef = 1;
  double cost = 0.0;  // estimated cost of the launch so far
  const double max_cost = c.max_cost != 0 ? static_cast<double>(c.max_cost) : 1e300;
  for (auto it = srv.pending.begin(); it != srv.pending.end() && slot.nq < c.max_batch;) {
    Request& q = **it;
    if (q.storage != storage || q.heavy != slot.heavy) {
      ++it;
      continue;
    }
    const uint64_t k = q.r.k;
    const size_t qbytes = srv.dim * sizeof(float);
    uint64_t can = std::min<uint64_t>(q.r.nq - q.placed, c.max_batch - slot.nq);
    can = std::min<uint64_t>(can, (c.max_rows - slot.rows) / k);
    // Launches are bounded by cost, so that expensive queries (large ef,
    // sparse filters) do not make every query of the launch wait for them.
    const double room = max_cost - cost;
    const uint64_t by_cost = room > 0.0 ? static_cast<uint64_t>(room / q.cost) : 0;
    can = std::min<uint64_t>(can, slot.nq == 0 ? std::max<uint64_t>(by_cost, 1) : by_cost);
    int fidx = -1;
    size_t fbytes = 0;  // words of a filter this launch does not hold yet
    if (q.r.filter_words != nullptr) {
      for (size_t f = 0; f < filters.size(); ++f) {
        if (filters[f].first == q.r.filter_words && filters[f].second == q.r.filter_size) fidx = static_cast<int>(f);
      }
      if (fidx < 0) {
        if (filters.size() == c.max_filters) can = 0;
        fbytes = ((q.r.filter_size + 63) / 64) * sizeof(uint64_t);
      }
    }
    const size_t var_cap = srv.in_bytes - srv.var_off;
    const size_t avail = var_cap > var_used + fbytes ? var_cap - var_used - fbytes : 0;
    can = std::min<uint64_t>(can, avail / qbytes);
    if (can == 0) {
      // Does not fit this launch; later requests may (smaller k, known filter).
      ++it;
      continue;
    }
    if (q.r.filter_words != nullptr && fidx < 0) {
      fidx = static_cast<int>(filters.size());
      filters.emplace_back(q.r.filter_words, q.r.filter_size);
      var_used += fbytes;
    }
    slot.pieces.push_back(Piece{*it, q.placed, static_cast<uint32_t>(can), slot.nq, slot.rows});
    (void)fidx;
    cost += static_cast<double>(can) * q.cost;
    var_used += can * qbytes;
    slot.nq += static_cast<uint32_t>(can);
    slot.rows += can * k;
    max_ef = std::max<uint64_t>(max_ef, hnswamd::dev::effective_ef(*srv.h, q.r.k, q.r.ef));
    q.placed += can;
    if (q.placed == q.r.nq) {
      it = srv.pending.erase(it);
    } else {
      break;  // the launch is full
    }
  }
  return filters;
}

/// Fill the staging area, then enqueue the launch on the slot's stream (as a
/// graph replay when enabled). Returns the number of bytes to copy in and out.
void launch(hnswgpu_server& srv, Slot& slot, int storage, uint64_t max_ef,
            const std::vector<std::pair<const uint64_t*, uint64_t>>& filters) {
  hnswgpu_index& h = *srv.h;
  const hnswgpu_server_config& c = srv.cfg;
  auto* base = static_cast<uint8_t*>(slot.stage.p);
  auto* dbase = static_cast<uint8_t*>(slot.in.p);
  auto* descs = reinterpret_cast<FilterDesc*>(base + srv.descs_off);
  auto* tasks = reinterpret_cast<QueryTask*>(base + srv.tasks_off);
  const auto* elig = static_cast<const uint32_t*>(slot.elig.p);
  auto* dout = static_cast<uint8_t*>(slot.out.p);

  std::shared_lock<std::shared_mutex> ilk(h.mu);  // index state (deletions, views) stays put while we read it
  const hnswamd::gpu::DevIndexView ix = hnswamd::dev::view_for(h, storage);
  const bool deletions = h.num_deleted > 0;
  const uint32_t side_list = hnswamd::dev::deleted_side_list(h);
  *reinterpret_cast<uint32_t*>(base) = slot.nq;
  size_t var = srv.var_off;
  for (uint32_t f = 0; f <= c.max_filters; ++f) descs[f] = FilterDesc{nullptr, 0, 0u, 0u};
  for (size_t f = 0; f < filters.size(); ++f) {
    const size_t nw = (filters[f].second + 63) / 64;
    std::memcpy(base + var, filters[f].first, nw * sizeof(uint64_t));
    descs[f] = FilterDesc{reinterpret_cast<const uint64_t*>(dbase + var), filters[f].second, 1u, 0u};
    var += nw * sizeof(uint64_t);
  }
  descs[c.max_filters] = FilterDesc{nullptr, 0, deletions ? 1u : 0u, 0u};
  std::vector<int> fidx(slot.pieces.size(), -1);
  std::vector<double> cost(slot.pieces.size());
  for (size_t pi = 0; pi < slot.pieces.size(); ++pi) {
    const hnswgpu_request& r = slot.pieces[pi].req->r;
    for (size_t f = 0; f < filters.size(); ++f) {
      if (filters[f].first == r.filter_words && filters[f].second == r.filter_size) fidx[pi] = static_cast<int>(f);
    }
    cost[pi] = slot.pieces[pi].req->cost;
  }
  // The work queue runs the costliest queries first, so that no long query
  // starts at the end of the launch (longest processing time first). Each
  // query keeps its output slots.
  std::vector<size_t> order(slot.pieces.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
  uint32_t next_task = 0;
  std::vector<float> scales;
  for (size_t pi : order) {
    const Piece& p = slot.pieces[pi];
    const hnswgpu_request& r = p.req->r;
    const uint32_t* e = fidx[pi] >= 0 ? elig + static_cast<size_t>(fidx[pi]) * srv.words
                                      : (deletions ? elig + static_cast<size_t>(c.max_filters) * srv.words : nullptr);
    const uint64_t ef = hnswamd::dev::effective_ef(h, r.k, r.ef);
    const size_t qbytes = srv.dim * sizeof(float);
    std::memcpy(base + var, r.queries + p.first * srv.dim, p.n * qbytes);
    if (ix.metric == 2) {
      // Cosine: the query scales in double on the host, as the reference computes them.
      scales.resize(p.n);
      hnswamd::dev::host_query_scales(r.queries + p.first * srv.dim, p.n, srv.dim, scales.data());
    }
    for (uint32_t i = 0; i < p.n; ++i) {
      QueryTask q{};
      q.query = reinterpret_cast<const float*>(dbase + var + i * qbytes);
      q.out = reinterpret_cast<hnswamd::gpu::OutRow*>(dout + srv.counts_bytes) + p.row + i * r.k;
      q.count = reinterpret_cast<uint32_t*>(dout) + p.task + i;
      q.elig = e;
      q.k = static_cast<uint32_t>(r.k);
      q.ef = static_cast<uint32_t>(ef);
      q.has_scale = ix.metric == 2 ? 1u : 0u;
      q.scale = ix.metric == 2 ? scales[i] : 1.0f;
      tasks[next_task++] = q;
    }
    var += p.n * qbytes;
  }
  const size_t in_used = var;
  const size_t out_used = srv.counts_bytes + slot.rows * sizeof(hnswgpu_neighbor);

  const uint32_t cls = ef_class(max_ef);
  hnswamd::gpu::SearchRequest r{};
  r.tasks = reinterpret_cast<const QueryTask*>(dbase + srv.tasks_off);
  r.nq_dev = reinterpret_cast<const uint32_t*>(dbase);
  r.dim = static_cast<uint32_t>(srv.dim);
  r.nq = c.max_batch;  // the grid covers the capacity; the kernels read the count
  r.ef = cls;
  r.force_spill = h.force_spill;
  // Deletions and no filter in the whole launch: the side-list kernel.
  r.deleted_only = deletions && filters.empty() ? static_cast<int>(side_list) : 0;
  hnswamd::dev::plan_request(h, r, r.nq);
  hnswamd::gpu::SearchScratch s = hnswamd::dev::bind_scratch(h, slot.sc, r, r.nq, c.spill_waves, slot.stream);
  auto* d_descs = reinterpret_cast<const FilterDesc*>(dbase + srv.descs_off);
  auto* d_elig = static_cast<uint32_t*>(slot.elig.p);
  const uint32_t* del_bits = static_cast<const uint32_t*>(h.deleted.p);
  ilk.unlock();

  auto enqueue = [&](size_t in_n, size_t out_n) {
    check(hipMemcpyAsync(slot.in.p, slot.stage.p, in_n, hipMemcpyHostToDevice, slot.stream), "hipMemcpy in");
    check(hnswamd::gpu::launch_eligibility_multi(ix, del_bits, d_descs, c.max_filters + 1, d_elig, slot.stream),
          "eligibility kernel");
    check(hnswamd::gpu::launch_prepare(ix, r, s, slot.stream), "prepare kernel");
    check(hnswamd::gpu::launch_search(ix, r, s, slot.stream), "search kernel");
    check(hipMemcpyAsync(slot.result.p, slot.out.p, out_n, hipMemcpyDeviceToHost, slot.stream), "hipMemcpy out");
  };
  const auto key = std::make_tuple(storage, cls, r.deleted_only);
  std::vector<uint64_t> shape = shape_words(r, s, ix, del_bits);
  auto it = slot.graphs.find(key);
  if (it != slot.graphs.end() && it->second.shape != shape) {
    // Captured with other launch parameters or scratch that has since been
    // reallocated: capture again.
    destroy_graph(it->second);
    slot.graphs.erase(it);
    it = slot.graphs.end();
    ++srv.recaptures;
  }
  if (c.use_graphs && it != slot.graphs.end()) {
    Graph& g = it->second;
    check(hipGraphExecMemcpyNodeSetParams1D(g.exec, g.h2d, slot.in.p, slot.stage.p, in_used, hipMemcpyHostToDevice),
          "graph copy-in size");
    check(hipGraphExecMemcpyNodeSetParams1D(g.exec, g.d2h, slot.result.p, slot.out.p, out_used,
                                            hipMemcpyDeviceToHost), "graph copy-out size");
    check(hipGraphLaunch(g.exec, slot.stream), "hipGraphLaunch");
    ++srv.graph_launches;
  } else {
    // Eager; the first launch of a shape also sets the kernels up, after
    // which the shape is captured for the next launches.
    enqueue(in_used, out_used);
    if (c.use_graphs) {
      Graph g;
      check(hipStreamBeginCapture(slot.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        enqueue(srv.in_bytes, srv.out_bytes);
      } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(slot.stream, &junk);
        if (junk != nullptr) (void)hipGraphDestroy(junk);
        throw;
      }
      check(hipStreamEndCapture(slot.stream, &g.graph), "hipStreamEndCapture");
      size_t n = 0;
      check(hipGraphGetNodes(g.graph, nullptr, &n), "hipGraphGetNodes");
      std::vector<hipGraphNode_t> nodes(n);
      check(hipGraphGetNodes(g.graph, nodes.data(), &n), "hipGraphGetNodes");
      for (hipGraphNode_t node : nodes) {
        hipGraphNodeType type;
        check(hipGraphNodeGetType(node, &type), "hipGraphNodeGetType");
        if (type != hipGraphNodeTypeMemcpy) continue;
        hipMemcpy3DParms p{};
        check(hipGraphMemcpyNodeGetParams(node, &p), "hipGraphMemcpyNodeGetParams");
        if (p.kind == hipMemcpyHostToDevice) g.h2d = node;
        if (p.kind == hipMemcpyDeviceToHost) g.d2h = node;
      }
      if (g.h2d == nullptr || g.d2h == nullptr) {
        (void)hipGraphDestroy(g.graph);
        throw hnswamd::DeviceError("server: captured graph lacks its copies");
      }
      check(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      g.shape = std::move(shape);
      slot.graphs.emplace(key, std::move(g));
    }
  }
  check(hipEventRecord(slot.done, slot.stream), "hipEventRecord");
}

void dispatch_loop(hnswgpu_server* srv) {
  for (;;) {
    std::unique_lock<std::mutex> lk(srv->mu);
    // A launch can start when a slot is free and some pending request may
    // start it (a heavy one only while another slot stays free).
    auto can_start = [&] {
      if (srv->pending.empty() || srv->free_slots.empty()) return false;
      if (srv->heavy_inflight + 1 < srv->cfg.slots || srv->cfg.slots == 1) return true;
      for (const auto& p : srv->pending) {
        if (!p->heavy) return true;
      }
      return false;
    };
    srv->cv_work.wait(lk, [&] { return srv->stopping || can_start(); });
    if (srv->stopping) return;
    Slot* slot = srv->free_slots.front();
    srv->free_slots.pop_front();
    int storage = 0;
    uint64_t max_ef = 1;
    const auto filters = form_launch(*srv, *slot, storage, max_ef);
    if (slot->pieces.empty()) {
      // Nothing fits (cannot happen for requests validated at submit); fail the oldest.
      auto q = srv->pending.front();
      srv->pending.pop_front();
      q->status = HNSWGPU_EINVAL;
      q->error = "server: request does not fit a launch";
      q->finished = true;
      srv->free_slots.push_back(slot);
      srv->cv_req.notify_all();
      continue;
    }
    ++srv->launches;
    srv->launched_queries += slot->nq;
    if (slot->heavy) ++srv->heavy_inflight;
    lk.unlock();
    slot->error.clear();
    try {
      launch(*srv, *slot, storage, max_ef, filters);
    } catch (...) {
      hnswamd::status_from_current_exception();
      slot->error = hnswamd::last_error();
      (void)hipStreamSynchronize(slot->stream);  // whatever was enqueued before the failure
    }
    lk.lock();
    slot->launched = true;
    srv->cv_done.notify_all();
  }
}

/// Completion of one slot's launches (each slot has its own thread, so a
/// short launch completes as soon as it is done, whatever else is in flight):
/// copy each request's rows to its buffers, record latencies, free the slot.
void complete_loop(hnswgpu_server* srv, Slot* slot) {
  for (;;) {
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] { return slot->launched || srv->stopping; });
    if (!slot->launched) return;
    lk.unlock();
    std::string err = slot->error;
    if (err.empty()) {
      const hipError_t e = hipEventSynchronize(slot->done);
      if (e != hipSuccess) err = std::string("server launch: ") + hipGetErrorString(e);
    }
    if (err.empty()) {
      const auto* base = static_cast<const uint8_t*>(slot->result.p);
      const auto* counts = reinterpret_cast<const uint32_t*>(base);
      const auto* rows = reinterpret_cast<const hnswgpu_neighbor*>(base + srv->counts_bytes);
      for (const Piece& p : slot->pieces) {
        const hnswgpu_request& r = p.req->r;
        std::memcpy(r.out + p.first * r.k, rows + p.row, p.n * r.k * sizeof(hnswgpu_neighbor));
        if (r.counts != nullptr) std::memcpy(r.counts + p.first, counts + p.task, p.n * sizeof(uint32_t));
      }
    }
    const auto now = Clock::now();
    lk.lock();
    for (const Piece& p : slot->pieces) {
      Request& q = *p.req;
      if (!err.empty() && q.status == HNSWGPU_OK) {
        q.status = HNSWGPU_EDEVICE;
        q.error = err;
      }
      q.done += p.n;
      if (q.done == q.r.nq) {
        q.finished = true;
        q.latency_ms = std::chrono::duration<double, std::milli>(now - q.t_submit).count();
        srv->latencies_ms.push_back(q.latency_ms);
        srv->queries_done += q.r.nq;
        srv->t_last = now;
      }
    }
    slot->pieces.clear();
    if (slot->heavy) --srv->heavy_inflight;
    slot->launched = false;
    srv->free_slots.push_back(slot);
    srv->cv_work.notify_one();
    srv->cv_req.notify_all();
  }
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

extern "C" {

int hnswgpu_server_create(hnswgpu_index* h, const hnswgpu_server_config* config, hnswgpu_server** out) {
  return guarded([&] {
    if (h == nullptr || out == nullptr) throw std::invalid_argument("server_create: null pointer");
    auto srv = std::make_unique<hnswgpu_server>();
    srv->h = h;
    hnswgpu_server_config c{};
    if (config != nullptr) c = *config;
    if (c.slots == 0) c.slots = 2;
    if (c.max_batch == 0) c.max_batch = 4096;
    if (c.max_rows == 0) c.max_rows = 64ull * c.max_batch;
    if (c.max_filters == 0) c.max_filters = 8;
    if (c.spill_waves == 0) c.spill_waves = 16;
    if (c.heavy_cost == 0) c.heavy_cost = 1024;
    if (c.slots > 16 || c.max_batch > (1u << 20) || c.max_filters > 256) {
      throw std::invalid_argument("server_create: slots <= 16, max_batch <= 2^20, max_filters <= 256");
    }
    {
      std::unique_lock<std::shared_mutex> lk(h->mu);
      if (h->view.count == 0) throw std::invalid_argument("server_create: empty index");
      for (int st = 1; st <= 2; ++st) {
        if (c.storages & (1u << st)) hnswamd::dev::build_reduced(*h, st);
      }
      srv->dim = h->dim;
      srv->padded = h->view.padded;
      if (c.max_filter_labels == 0) {
        // Enough for every label of the index.
        std::vector<uint64_t> labels(h->view.count);
        check(hipMemcpy(labels.data(), h->view.labels, labels.size() * sizeof(uint64_t), hipMemcpyDeviceToHost),
              "hipMemcpy labels");
        c.max_filter_labels = *std::max_element(labels.begin(), labels.end()) + 1;
      }
      srv->words = (h->view.count + 31) / 32;
    }
    srv->filter_words_max = (c.max_filter_labels + 63) / 64;
    srv->cfg = c;
    srv->descs_off = kAlign;
    srv->tasks_off = align_up(srv->descs_off + (c.max_filters + 1) * sizeof(FilterDesc));
    srv->var_off = align_up(srv->tasks_off + static_cast<size_t>(c.max_batch) * sizeof(QueryTask));
    srv->in_bytes = align_up(srv->var_off + static_cast<size_t>(c.max_batch) * srv->dim * sizeof(float) +
                             c.max_filters * srv->filter_words_max * sizeof(uint64_t));
    srv->counts_bytes = align_up(static_cast<size_t>(c.max_batch) * sizeof(uint32_t));
    srv->out_bytes = srv->counts_bytes + c.max_rows * sizeof(hnswgpu_neighbor);

    // Device memory of one slot, and the admission budget.
    const size_t resident = 256ull * 32;  // wavefront tables for any launch shape
    const size_t visit = resident * (static_cast<size_t>(1) << h->visit_log2) * sizeof(uint32_t);
    const size_t spill = static_cast<size_t>(c.spill_waves) * ((h->view.count + 31) / 32 * 4 +
                                                               (2ull * h->view.count + 4) * 8);
    srv->slot_bytes = srv->in_bytes + srv->out_bytes + (c.max_filters + 1) * srv->words * 4 + visit + spill +
                      static_cast<size_t>(c.max_batch) * (srv->padded * 4 + 4);
    size_t free_b = 0, total_b = 0;
    check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t budget = c.hbm_budget != 0 ? c.hbm_budget : free_b / 4;
    const uint32_t fit = static_cast<uint32_t>(std::min<size_t>(c.slots, budget / std::max<size_t>(1, srv->slot_bytes)));
    if (fit == 0) {
      throw std::invalid_argument("server_create: one slot needs " + std::to_string(srv->slot_bytes) +
                                  " bytes of device memory, over the budget of " + std::to_string(budget));
    }
    srv->cfg.slots = fit;

    for (uint32_t i = 0; i < fit; ++i) {
      auto slot = std::make_unique<Slot>();
      check(hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking), "hipStreamCreate");
      check(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming), "hipEventCreate");
      slot->stage.need(srv->in_bytes);
      slot->result.need(srv->out_bytes);
      slot->in.need(srv->in_bytes);
      slot->out.need(srv->out_bytes);
      slot->elig.need((c.max_filters + 1) * srv->words * sizeof(uint32_t));
      // Scratch for the largest launch shape, allocated and cleared now (not
      // inside a captured graph).
      std::shared_lock<std::shared_mutex> lk(h->mu);
      for (uint32_t cls : {64u, 128u, 256u, hnswamd::gpu::kFastMaxEf}) {
        hnswamd::gpu::SearchRequest r{};
        r.ef = cls;
        hnswamd::dev::plan_request(*h, r, c.max_batch);
        (void)hnswamd::dev::bind_scratch(*h, slot->sc, r, c.max_batch, c.spill_waves, slot->stream);
      }
      check(hipStreamSynchronize(slot->stream), "server scratch");
      srv->free_slots.push_back(slot.get());
      srv->slots.push_back(std::move(slot));
    }
    hnswgpu_server* s = srv.release();
    h->servers.fetch_add(1);
    s->dispatcher = std::thread(dispatch_loop, s);
    for (auto& slot : s->slots) slot->completer = std::thread(complete_loop, s, slot.get());
    *out = s;
  });
}

int hnswgpu_server_submit(hnswgpu_server* s, const hnswgpu_request* req, uint64_t* ticket) {
  return guarded([&] {
    if (s == nullptr || req == nullptr || ticket == nullptr) throw std::invalid_argument("server_submit: null pointer");
    const hnswgpu_request& r = *req;
    if (r.k == 0) throw std::invalid_argument("search: k must be > 0");
    if (r.nq > 0 && (r.queries == nullptr || r.out == nullptr)) throw std::invalid_argument("search_batch: null pointer");
    if (r.k > s->cfg.max_rows) throw std::invalid_argument("server_submit: k exceeds the launch's result rows");
    if (r.filter_words != nullptr && (r.filter_size + 63) / 64 > s->filter_words_max) {
      throw std::invalid_argument("server_submit: filter covers more labels than the server was sized for");
    }
    auto q = std::make_shared<Request>();
    q->r = r;
    // Cost of one query: its ef over the fraction of labels its filter allows
    // (a filtered search expands about that many more nodes).
    double density = 1.0;
    if (r.filter_words != nullptr) {
      uint64_t ones = 0;
      const uint64_t nw = (r.filter_size + 63) / 64;
      for (uint64_t i = 0; i < nw; ++i) {
        uint64_t w = r.filter_words[i];
        if (i + 1 == nw && r.filter_size % 64 != 0) w &= (1ULL << (r.filter_size % 64)) - 1;
        ones += static_cast<uint64_t>(__builtin_popcountll(w));
      }
      density = r.filter_size > 0 ? static_cast<double>(ones) / static_cast<double>(r.filter_size) : 0.0;
    }
    {
      std::shared_lock<std::shared_mutex> ilk(s->h->mu);
      q->storage = r.storage < 0 ? s->h->view.storage : r.storage;
      (void)hnswamd::dev::view_for(*s->h, q->storage);  // throws if that copy is not built
      if (hnswamd::dev::effective_ef(*s->h, r.k, r.ef) > 0xFFFFFFFFull) {
        throw std::invalid_argument("server_submit: ef out of range");
      }
      q->cost = static_cast<double>(hnswamd::dev::effective_ef(*s->h, r.k, r.ef)) / std::max(density, 0.02);
      q->heavy = q->cost >= static_cast<double>(s->cfg.heavy_cost);
    }
    q->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) throw std::runtime_error("server_submit: server is stopping");
    q->ticket = s->next_ticket++;
    if (!s->any_submit) {
      s->any_submit = true;
      s->t_first = q->t_submit;
    }
    s->requests.emplace(q->ticket, q);
    if (r.nq == 0) {
      q->finished = true;
    } else {
      s->pending.push_back(q);
      s->cv_work.notify_one();
    }
    *ticket = q->ticket;
  });
}

int hnswgpu_server_wait_latency(hnswgpu_server* s, uint64_t ticket, double* latency_ms) {
  std::shared_ptr<Request> q;
  const int rc = guarded([&] {
    if (s == nullptr) throw std::invalid_argument("server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) throw std::invalid_argument("server_wait: unknown ticket");
    q = it->second;
    s->cv_req.wait(lk, [&] { return q->finished; });
    s->requests.erase(it);
  });
  if (rc != HNSWGPU_OK) return rc;
  if (latency_ms != nullptr) *latency_ms = q->latency_ms;
  if (q->status != HNSWGPU_OK) hnswamd::set_last_error(q->error);
  return q->status;
}

int hnswgpu_server_wait(hnswgpu_server* s, uint64_t ticket) { return hnswgpu_server_wait_latency(s, ticket, nullptr); }

int hnswgpu_server_stats(hnswgpu_server* s, double* out, uint64_t n) {
  return guarded([&] {
    if (s == nullptr || out == nullptr) throw std::invalid_argument("server_stats: null pointer");
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> l = s->latencies_ms;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
      if (l.empty()) return 0.0;
      const size_t i = std::min(l.size() - 1, static_cast<size_t>(p * static_cast<double>(l.size() - 1) + 0.5));
      return l[i];
    };
    double mean = 0.0;
    for (double v : l) mean += v;
    if (!l.empty()) mean /= static_cast<double>(l.size());
    const double span = s->any_submit && !l.empty() ? std::chrono::duration<double>(s->t_last - s->t_first).count() : 0.0;
    const double v[] = {static_cast<double>(l.size()),
                        static_cast<double>(s->queries_done),
                        static_cast<double>(s->launches),
                        s->launches ? static_cast<double>(s->launched_queries) / static_cast<double>(s->launches) : 0.0,
                        pct(0.50),
                        pct(0.99),
                        l.empty() ? 0.0 : l.back(),
                        span > 0 ? static_cast<double>(s->queries_done) / span : 0.0,
                        mean,
                        static_cast<double>(s->cfg.slots),
                        static_cast<double>(s->slot_bytes),
                        static_cast<double>(s->graph_launches),
                        static_cast<double>(s->recaptures)};
    for (uint64_t i = 0; i < n && i < sizeof(v) / sizeof(v[0]); ++i) out[i] = v[i];
  });
}

int hnswgpu_server_reset_stats(hnswgpu_server* s) {
  return guarded([&] {
    if (s == nullptr) throw std::invalid_argument("server_reset_stats: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->latencies_ms.clear();
    s->queries_done = s->launches = s->graph_launches = s->launched_queries = s->recaptures = 0;
    s->any_submit = false;
  });
}

int hnswgpu_server_free(hnswgpu_server* s) {
  return guarded([&] {
    if (s == nullptr) return;
    {
      // Finish what was submitted, then stop both threads.
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_req.wait(lk, [&] { return s->pending.empty() && s->free_slots.size() == s->slots.size(); });
      s->stopping = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    for (auto& slot : s->slots) {
      if (slot->completer.joinable()) slot->completer.join();
    }
    for (auto& slot : s->slots) (void)hipStreamSynchronize(slot->stream);
    hnswgpu_index* h = s->h;
    delete s;
    h->servers.fetch_sub(1);
  });
}

}  // extern "C"
