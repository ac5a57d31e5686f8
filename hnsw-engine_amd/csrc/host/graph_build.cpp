// Graph construction: Malkov & Yashunin's INSERT (Alg. 1) with the
// SELECT-NEIGHBORS-HEURISTIC (Alg. 4, no candidate extension, no kept pruned
// connections), organised as insertion waves.
//
// A wave is a run of consecutive new ids [s, e). Its nodes are inserted against
// the graph as it stood before the wave (the frozen graph: ids below s, its
// entry point and top level). A wave has three phases; each runs in parallel
// over independent items and none depends on timing:
//   1. probe (read-only): per node, the greedy descent from the frozen entry
//      point down to the node's level, then per layer lc <= min(level, top) an
//      ef_construction beam search (Alg. 2) seeded by the layer above's result.
//      Nothing writes the graph during this phase, so it takes no locks, and it
//      is the part a GPU can take over (ProbeFn);
//   2. choose: per node and layer, the probe result (on an upper layer joined
//      with the earlier nodes of the same wave that reach it, at their exact
//      distances), then Alg. 4 keeps up to M; the node's own list is written;
//   3. back-link: the wave's proposals "add source to target's list", grouped
//      by (layer, target) and applied per group in source order: appended while
//      the list has room, otherwise Alg. 4 over the old members plus the source.
// With one node per wave this is the reference's sequential insertion
// (src/index.cpp:205-306): a one-thread build writes the reference's bytes.
// Multi-threaded builds use waves that grow with the graph (1/32 of it, at most
// 16,384 nodes), so their graph depends on the input alone, not on the thread
// count, on scheduling or on where phase 1 ran. Nodes of one wave do not see each other on layer 0
// (at 1/32 of the graph that cost no recall: 100k x 128 mixture, recall@10 at
// ef 16..128 equal to a one-thread build within 0.001); on the sparse upper
// layers, and above the frozen top, they are joined as candidates.
//
// What fixes the bytes of a one-thread build, kept as the reference defines it:
//   * levels from splitmix64 (src/index.cpp:23-28) as floor(-ln(U) * mL), U in
//     (0, 1] from the top 53 bits, mL = 1 / ln(M), capped at 64 (:94-99), drawn
//     in input order before any linking (:298-305);
//   * the beam search's admission and stop rules and its (distance, id) order
//     (src/search_layer.cpp:58-134), the strict greedy descent (:29-56);
//   * M neighbours kept on every layer; layer 0 lists hold up to 2M, upper ones
//     up to M, and a list that shrinks keeps its stale tail entries.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <stdexcept>
#include <thread>
#include <unordered_set>

#include "distance.hpp"
#include "graph.hpp"
#include "graph_build.hpp"

namespace hnswamd {

template <typename T>
void AlignedArray<T>::grow(std::size_t n) {
  if (n <= n_) return;
  const std::size_t bytes = round_up(n * sizeof(T), kRowAlign);
  auto* q = static_cast<T*>(std::aligned_alloc(kRowAlign, bytes));
  if (q == nullptr) throw std::bad_alloc();
  std::memset(static_cast<void*>(q), 0, bytes);
  if (p_ != nullptr) std::memcpy(static_cast<void*>(q), p_, n_ * sizeof(T));
  std::free(p_);
  p_ = q;
  n_ = n;
}
template class AlignedArray<float>;
template class AlignedArray<uint32_t>;

// ---------------------------------------------------------------------------
// A fixed team of threads that runs index-space jobs: run(n, fn) calls
// fn(i, worker) for every i in [0, n) and returns when all calls have ended,
// rethrowing the first exception. The calling thread is worker 0.
// ---------------------------------------------------------------------------
Team::Team(std::size_t threads) : size_(std::max<std::size_t>(1, threads)) {
  for (std::size_t w = 1; w < size_; ++w) crew_.emplace_back([this, w] { serve(w); });
}

Team::~Team() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    closing_ = true;
  }
  wake_.notify_all();
  for (auto& t : crew_) t.join();
}

void Team::drain(std::size_t worker) {
  for (;;) {
    if (failed_.load(std::memory_order_relaxed)) return;
    const std::size_t i = next_.fetch_add(1, std::memory_order_relaxed);
    if (i >= items_) return;
    try {
      (*job_)(i, worker);
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      if (!error_) error_ = std::current_exception();
      failed_.store(true);
    }
  }
}

void Team::serve(std::size_t worker) {
  uint64_t seen = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      wake_.wait(lk, [&] { return closing_ || round_ != seen; });
      if (closing_) return;
      seen = round_;
    }
    drain(worker);
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (--busy_ == 0) done_.notify_all();
    }
  }
}

void Team::run(std::size_t n, const std::function<void(std::size_t, std::size_t)>& fn) {
  if (n == 0) return;
  if (size_ == 1 || n == 1) {
    for (std::size_t i = 0; i < n; ++i) fn(i, 0);
    return;
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    job_ = &fn;
    items_ = n;
    next_.store(0);
    failed_.store(false);
    error_ = nullptr;
    busy_ = crew_.size();
    ++round_;
  }
  wake_.notify_all();
  drain(0);
  std::exception_ptr err;
  {
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return busy_ == 0; });
    job_ = nullptr;
    err = error_;
  }
  if (err) std::rethrow_exception(err);
}

namespace {

inline uint64_t splitmix64_next(uint64_t& s) noexcept {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline bool before(const Key& a, const Key& b) noexcept { return a.d < b.d || (a.d == b.d && a.id < b.id); }
inline bool after(const Key& a, const Key& b) noexcept { return before(b, a); }

/// Per-worker search state: generation stamps over the node ids (a node is
/// seen in the current search iff its stamp equals `gen`), the frontier as a
/// binary heap with the nearest key on top, and the beam as a binary heap with
/// the farthest key on top (sorted once the search ends).
struct Walker {
  std::vector<uint32_t> stamp;
  uint32_t gen = 0;
  std::vector<Key> frontier;
  std::vector<Key> beam;
  std::vector<Key> pool;  // Alg. 4 input
  std::vector<Key> kept;  // Alg. 4 output

  void fresh(std::size_t n) {
    if (stamp.size() < n) {
      stamp.assign(n, 0u);
      gen = 0;
    }
    if (++gen == 0) {  // wrapped: every old stamp could collide
      std::fill(stamp.begin(), stamp.end(), 0u);
      gen = 1;
    }
  }
  bool first_visit(uint32_t id) {
    if (stamp[id] == gen) return false;
    stamp[id] = gen;
    return true;
  }
};

/// Greedy descent on one layer: repeated passes over the list of the node that
/// is current when the pass starts; a neighbour replaces the current node when
/// strictly closer (src/search_layer.cpp:29-56).
Key descend(const Graph& g, const float* q, Key cur, int layer) {
  for (;;) {
    const uint32_t* l = g.list(cur.id, layer);
    const uint32_t* end = l + 1 + l[0];
    Key best = cur;
    for (const uint32_t* p = l + 1; p != end; ++p) {
      const float d = g.dist(q, g.vec(*p));
      if (d < best.d) best = Key{d, *p};
    }
    if (best.id == cur.id) return cur;
    cur = best;
  }
}

/// Beam search of width w on one layer (Alg. 2, src/search_layer.cpp:58-134):
/// the frontier is expanded nearest first by (distance, id) until its nearest
/// key is farther than the beam's worst while the beam is full; a fresh
/// neighbour is admitted iff the beam has room or it is strictly closer than
/// the beam's worst, which it then displaces. Leaves the beam sorted.
void beam_search(const Graph& g, const float* q, const Key* seeds, std::size_t n_seeds, std::size_t w, int layer,
                 Walker& s) {
  s.fresh(g.capacity);
  s.frontier.clear();
  s.beam.clear();
  for (std::size_t i = 0; i < n_seeds; ++i) {
    if (!s.first_visit(seeds[i].id)) continue;
    s.frontier.push_back(seeds[i]);
    s.beam.push_back(seeds[i]);
  }
  std::make_heap(s.frontier.begin(), s.frontier.end(), after);
  if (s.beam.size() > w) {  // keep the w nearest seeds
    std::nth_element(s.beam.begin(), s.beam.begin() + static_cast<std::ptrdiff_t>(w), s.beam.end(), before);
    s.beam.resize(w);
  }
  std::make_heap(s.beam.begin(), s.beam.end(), before);

  while (!s.frontier.empty()) {
    const Key c = s.frontier.front();
    if (s.beam.size() >= w && c.d > s.beam.front().d) break;
    std::pop_heap(s.frontier.begin(), s.frontier.end(), after);
    s.frontier.pop_back();
    const uint32_t* l = g.list(c.id, layer);
    const uint32_t n = l[0];
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t id = l[1 + j];
      if (!s.first_visit(id)) continue;
      const float d = g.dist(q, g.vec(id));
      const bool room = s.beam.size() < w;
      if (!room && !(d < s.beam.front().d)) continue;
      s.frontier.push_back(Key{d, id});
      std::push_heap(s.frontier.begin(), s.frontier.end(), after);
      if (!room) {  // the farthest key leaves
        std::pop_heap(s.beam.begin(), s.beam.end(), before);
        s.beam.back() = Key{d, id};
      } else {
        s.beam.push_back(Key{d, id});
      }
      std::push_heap(s.beam.begin(), s.beam.end(), before);
    }
  }
  std::sort_heap(s.beam.begin(), s.beam.end(), before);
}

/// Alg. 4: walk `sorted` (ascending by (distance to the base node, id)) and keep
/// a key unless an already kept one is strictly closer to it than the base
/// node is; stop at m kept.
void select_diverse(const Graph& g, const std::vector<Key>& sorted, std::size_t m, std::vector<Key>& kept) {
  kept.clear();
  for (const Key& e : sorted) {
    if (kept.size() == m) break;
    const float* ev = g.vec(e.id);
    const bool shadowed = std::any_of(kept.begin(), kept.end(),
                                      [&](const Key& r) { return g.dist(ev, g.vec(r.id)) < e.d; });
    if (!shadowed) kept.push_back(e);
  }
}

void store_list(uint32_t* l, const std::vector<Key>& keys) {
  for (std::size_t i = 0; i < keys.size(); ++i) l[1 + i] = keys[i].id;
  l[0] = static_cast<uint32_t>(keys.size());
}

/// "Add `source` (at distance d) to `target`'s list at `layer`".
struct Proposal {
  int layer;
  uint32_t target;
  uint32_t source;
  float d;
};

}  // namespace

// ---------------------------------------------------------------------------
// Phase 1 on the host
// ---------------------------------------------------------------------------
void probe_on_host(const Graph& g, const WaveFrame& f, uint32_t id, std::vector<std::vector<Key>>& by_layer) {
  thread_local Walker s;
  const int level = g.levels[id];
  const int reach = std::min(level, f.top);
  by_layer.assign(static_cast<std::size_t>(reach + 1), {});
  if (f.entry == kNoNode) return;
  const float* q = g.vec(id);
  Key cur{g.dist(q, g.vec(f.entry)), f.entry};
  for (int lc = f.top; lc > level; --lc) cur = descend(g, q, cur, lc);
  std::vector<Key> seeds{cur};
  for (int lc = reach; lc >= 0; --lc) {
    beam_search(g, q, seeds.data(), seeds.size(), g.efc, lc, s);
    by_layer[static_cast<std::size_t>(lc)] = s.beam;
    seeds = s.beam;
  }
}

// ---------------------------------------------------------------------------
// Waves
// ---------------------------------------------------------------------------
std::size_t wave_length(std::size_t frozen, std::size_t threads) {
  if (threads <= 1 || frozen == 0) return 1;
  return std::clamp<std::size_t>(frozen / 32, 1, kMaxWave);
}

void insert_wave(Graph& g, uint32_t s, uint32_t e, Team& team, const ProbeFn& probe,
                 std::vector<uint32_t>& changed) {
  const WaveFrame frame{g.entry_point, g.max_level, &changed};
  const std::size_t n = e - s;
  std::vector<std::vector<std::vector<Key>>> found(n);

  // 1. probe the frozen graph
  if (probe) {
    probe(g, frame, s, e, found, team);
  } else {
    team.run(n, [&](std::size_t i, std::size_t) { probe_on_host(g, frame, s + static_cast<uint32_t>(i), found[i]); });
  }

  // 2. choose: own lists, and the back-link proposals of every node
  std::vector<std::vector<Proposal>> proposed(n);
  team.run(n, [&](std::size_t i, std::size_t) {
    thread_local Walker s_;
    const uint32_t id = s + static_cast<uint32_t>(i);
    const int level = g.levels[id];
    const float* q = g.vec(id);
    if (frame.entry == kNoNode) return;  // the first node of the graph links to nothing
    for (int lc = level; lc >= 0; --lc) {
      std::vector<Key>& pool = s_.pool;
      pool.clear();
      if (static_cast<std::size_t>(lc) < found[i].size()) pool = found[i][static_cast<std::size_t>(lc)];
      if (lc > 0) {  // earlier nodes of this wave that reach this (sparse) layer
        for (uint32_t j = s; j < id; ++j) {
          if (g.levels[j] >= lc) pool.push_back(Key{g.dist(q, g.vec(j)), j});
        }
      }
      if (pool.empty()) continue;
      std::sort(pool.begin(), pool.end(), before);
      select_diverse(g, pool, g.M, s_.kept);
      store_list(g.list(id, lc), s_.kept);
      for (const Key& k : s_.kept) proposed[i].push_back(Proposal{lc, k.id, id, k.d});
    }
  });

  // 3. back-links, one group per (layer, target), sources in id order
  std::vector<Proposal> all;
  for (auto& v : proposed) all.insert(all.end(), v.begin(), v.end());
  std::sort(all.begin(), all.end(), [](const Proposal& a, const Proposal& b) {
    if (a.layer != b.layer) return a.layer < b.layer;
    if (a.target != b.target) return a.target < b.target;
    return a.source < b.source;
  });
  std::vector<std::size_t> starts;
  for (std::size_t i = 0; i < all.size(); ++i) {
    if (i == 0 || all[i].layer != all[i - 1].layer || all[i].target != all[i - 1].target) starts.push_back(i);
  }
  starts.push_back(all.size());
  team.run(starts.size() - 1, [&](std::size_t gi, std::size_t) {
    thread_local Walker s_;
    for (std::size_t p = starts[gi]; p < starts[gi + 1]; ++p) {
      const Proposal& pr = all[p];
      uint32_t* l = g.list(pr.target, pr.layer);
      const uint32_t have = l[0];
      if (std::find(l + 1, l + 1 + have, pr.source) != l + 1 + have) continue;
      if (have < g.degree_cap(pr.layer)) {
        l[1 + have] = pr.source;
        l[0] = have + 1;
        continue;
      }
      const float* tv = g.vec(pr.target);
      std::vector<Key>& pool = s_.pool;
      pool.assign(1, Key{pr.d, pr.source});
      for (uint32_t j = 0; j < have; ++j) pool.push_back(Key{g.dist(tv, g.vec(l[1 + j])), l[1 + j]});
      std::sort(pool.begin(), pool.end(), before);
      select_diverse(g, pool, g.degree_cap(pr.layer), s_.kept);
      store_list(l, s_.kept);
    }
  });

  // what the next wave's probe has to refresh: the new nodes and the targets
  std::vector<uint32_t> next_changed;
  next_changed.reserve(n + starts.size());
  for (uint32_t id = s; id < e; ++id) next_changed.push_back(id);
  for (std::size_t gi = 0; gi + 1 < starts.size(); ++gi) {
    const uint32_t t = all[starts[gi]].target;
    if (t < s || t >= e) next_changed.push_back(t);
  }
  std::sort(next_changed.begin(), next_changed.end());
  next_changed.erase(std::unique(next_changed.begin(), next_changed.end()), next_changed.end());
  changed.swap(next_changed);

  // the highest new level becomes the entry point (the first such node)
  for (uint32_t id = s; id < e; ++id) {
    if (static_cast<int>(g.levels[id]) > g.max_level) {
      g.entry_point = id;
      g.max_level = g.levels[id];
    }
  }
}

// ---------------------------------------------------------------------------
// Graph
// ---------------------------------------------------------------------------
Graph::Graph(const BuildParams& p) : params(p) {
  if (p.dim == 0) throw std::invalid_argument("Index: dim must be > 0");
  if (p.M < 2) throw std::invalid_argument("Index: M must be >= 2");
  if (p.M > 2048) throw std::invalid_argument("Index: M must be <= 2048");
  if (p.ef_construction == 0) throw std::invalid_argument("Index: ef_construction must be > 0");
  if (p.max_elements > std::numeric_limits<uint32_t>::max() - 1) {
    throw std::invalid_argument("Index: max_elements exceeds the 32-bit internal id space");
  }
  if (p.metric < 0 || p.metric > 2) throw std::invalid_argument("Index: unknown metric");
  dim = p.dim;
  padded_dim = round_up(dim, kPadFloats);
  M = p.M;
  M0 = 2 * p.M;
  efc = std::max(p.ef_construction, p.M);
  mL = 1.0 / std::log(static_cast<double>(M));
  l0_stride = 1 + M0;
  upper_stride = 1 + M;
  rng_state = p.seed;
  dist_fn = pick_distance(p.metric);
  if (p.max_elements > 0) reserve(p.max_elements);
}

// The builder's distances: the engine's fixed summation order (distance.hpp),
// compiled a second time for AVX-512 hosts. std::fma and the fixed order make
// both versions bit-identical; the wider one only runs the 16 lanes of each
// accumulator in one register.
namespace {
float l2_base(const float* a, const float* b, std::size_t n) noexcept { return l2_sq(a, b, n); }
float ip_base(const float* a, const float* b, std::size_t n) noexcept { return 1.0f - dot(a, b, n); }
__attribute__((target("avx512f,avx2,fma"))) float l2_wide(const float* a, const float* b, std::size_t n) noexcept {
  return ordered_sum<true>(a, b, n);
}
__attribute__((target("avx512f,avx2,fma"))) float ip_wide(const float* a, const float* b, std::size_t n) noexcept {
  return 1.0f - ordered_sum<false>(a, b, n);
}
}  // namespace

Graph::DistFn Graph::pick_distance(int metric) {
  const bool wide = __builtin_cpu_supports("avx512f");
  if (metric == kL2) return wide ? l2_wide : l2_base;
  return wide ? ip_wide : ip_base;
}

void Graph::rebind() {
  vecs = vec_store.data();
  l0 = l0_store.data();
  upper = upper_store.data();
  labels = label_store.data();
  levels = level_store.data();
}

void Graph::reserve(std::size_t n) {
  if (n <= capacity) return;
  if (read_only) throw std::runtime_error("index is read-only (memory-mapped)");
  if (n > std::numeric_limits<uint32_t>::max() - 1) {
    throw std::runtime_error("Index: capacity exceeds the 32-bit internal id space");
  }
  vec_store.grow(n * padded_dim);
  l0_store.grow(n * l0_stride);
  label_store.resize(n, 0);
  level_store.resize(n, 0);
  upper_offset.resize(n, 0);
  deleted.resize(n, 0);
  capacity = n;
  rebind();
}

uint32_t Graph::append(const float* v, uint64_t label) {
  const auto id = static_cast<uint32_t>(count);
  prepare_vector(params.metric, v, vec_store.data() + static_cast<std::size_t>(id) * padded_dim, dim);
  label_store[id] = label;
  const double u = 1.0 - static_cast<double>(splitmix64_next(rng_state) >> 11) * 0x1.0p-53;
  const int level = static_cast<int>(std::min(std::floor(-std::log(u) * mL), static_cast<double>(kMaxLevel)));
  level_store[id] = static_cast<uint8_t>(level);
  if (level > 0) {
    upper_offset[id] = upper_used;
    upper_used += static_cast<std::size_t>(level) * upper_stride;
    if (upper_store.size() < upper_used) {
      upper_store.resize(std::max(upper_used, upper_store.size() * 2), 0u);
      upper = upper_store.data();
    }
  }
  l0_store.data()[static_cast<std::size_t>(id) * l0_stride] = 0;
  label_to_id.emplace(label, id);
  ++count;
  return id;
}

void Graph::add_batch(const float* vin, const uint64_t* lin, std::size_t n, std::size_t threads,
                      const ProbeFn& probe) {
  if (n == 0) return;
  if (vin == nullptr || lin == nullptr) throw std::invalid_argument("add_batch: null input");
  if (read_only) throw std::runtime_error("index is read-only (memory-mapped); load without mmap to modify");
  if (n > 1) {
    std::unordered_set<uint64_t> seen;
    seen.reserve(n);
    for (std::size_t i = 0; i < n; ++i) {
      if (!seen.insert(lin[i]).second) {
        throw std::invalid_argument("add_batch: duplicate label " + std::to_string(lin[i]) + " in batch");
      }
    }
  }
  for (std::size_t i = 0; i < n; ++i) {
    if (label_to_id.count(lin[i]) != 0) {
      throw std::invalid_argument("add: label " + std::to_string(lin[i]) + " already exists");
    }
  }
  const std::size_t need = count + n;
  if (need > capacity) {
    if (params.max_elements > 0) {
      throw std::runtime_error("Index is full (max_elements = " + std::to_string(params.max_elements) + ")");
    }
    reserve(std::max({need, capacity * 2, std::size_t{1024}}));
  }
  label_to_id.reserve(count + n);

  // Rows, labels and levels in input order, then the waves.
  const auto first = static_cast<uint32_t>(count);
  for (std::size_t i = 0; i < n; ++i) append(vin + i * dim, lin[i]);
  const std::size_t nt = threads == 0 ? std::max(1u, std::thread::hardware_concurrency()) : threads;
  Team team(std::min(nt, n));
  std::vector<uint32_t> changed;
  for (uint32_t s = first, end = first + static_cast<uint32_t>(n); s < end;) {
    const uint32_t e = std::min<uint32_t>(end, s + static_cast<uint32_t>(wave_length(s, nt)));
    insert_wave(*this, s, e, team, probe, changed);
    s = e;
  }
}

std::size_t Graph::memory_usage() const {
  return count * (padded_dim * sizeof(float) + l0_stride * sizeof(uint32_t) + sizeof(uint64_t) * 2 + 2) +
         upper_used * sizeof(uint32_t);
}

}  // namespace hnswamd
