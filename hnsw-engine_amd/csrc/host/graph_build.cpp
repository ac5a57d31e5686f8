// CPU graph construction in the reference's algorithm (Malkov & Yashunin,
// Alg. 1 INSERT with Alg. 4 SELECT-NEIGHBORS-HEURISTIC), so that indexes for
// the GPU can be created on any host and, built with one thread, are
// byte-identical to the reference's files. The steps that fix the bytes of the
// output are kept as the reference has them (src/index.cpp:17-306):
//   * levels from splitmix64 (src/index.cpp:23-28) as floor(-ln(U) * mL), U in
//     (0, 1] from the top 53 bits, mL = 1 / ln(M), capped at 64 (:94-99);
//   * insertion: greedy descent above the node's level, then per layer an
//     ef_construction search, the heuristic keeps up to M neighbours (also on
//     layer 0) and reverse edges are added, a full list being re-pruned with the
//     heuristic over its old members plus the new node (:145-241);
//   * ids are assigned and levels drawn in input order before any linking
//     (:298-305), so the thread count does not change the level sequence.
// Ordering everywhere is (distance, id), and a list that shrinks keeps its stale
// tail entries, exactly as the file format records them.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <queue>
#include <stdexcept>
#include <thread>
#include <unordered_set>

#include "distance.hpp"
#include "graph.hpp"

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

namespace {

inline uint64_t splitmix64_next(uint64_t& s) noexcept {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Item {
  float d;
  uint32_t id;
};
inline bool item_less(const Item& a, const Item& b) noexcept {
  return a.d < b.d || (a.d == b.d && a.id < b.id);
}
struct NearFirst {
  bool operator()(const Item& a, const Item& b) const noexcept { return item_less(b, a); }
};
struct FarFirst {
  bool operator()(const Item& a, const Item& b) const noexcept { return item_less(a, b); }
};

/// Per-worker scratch: epoch-stamped visited marks and heap storage.
struct Workspace {
  std::vector<uint32_t> stamp;
  uint32_t epoch = 0;
  std::vector<uint32_t> nbuf;
  void begin(std::size_t n) {
    if (stamp.size() < n) {
      stamp.assign(n, 0u);
      epoch = 0;
    }
    if (++epoch == 0) {
      std::fill(stamp.begin(), stamp.end(), 0u);
      epoch = 1;
    }
  }
};

/// Copy a node's list at `level` under its lock.
void snapshot(const Graph& g, uint32_t id, int level, std::vector<uint32_t>& out) {
  std::lock_guard<std::mutex> lk(g.lock_of(id));
  const uint32_t* l = g.list(id, level);
  out.assign(l + 1, l + 1 + l[0]);
}

Item descend(const Graph& g, const float* q, Item cur, int level, Workspace& ws) {
  for (bool moved = true; moved;) {
    moved = false;
    snapshot(g, cur.id, level, ws.nbuf);
    for (uint32_t nb : ws.nbuf) {
      const float d = g.dist(q, g.vec(nb));
      if (d < cur.d) {
        cur = Item{d, nb};
        moved = true;
      }
    }
  }
  return cur;
}

/// Alg. 2 on one layer from several entry points; returns the result set
/// sorted by (distance, id).
std::vector<Item> layer_search(const Graph& g, const float* q, const std::vector<Item>& eps, std::size_t ef,
                               int level, Workspace& ws) {
  ws.begin(g.capacity);
  std::priority_queue<Item, std::vector<Item>, NearFirst> cand;
  std::priority_queue<Item, std::vector<Item>, FarFirst> res;
  for (const Item& e : eps) {
    if (ws.stamp[e.id] == ws.epoch) continue;
    ws.stamp[e.id] = ws.epoch;
    cand.push(e);
    res.push(e);
  }
  while (res.size() > ef) res.pop();
  while (!cand.empty()) {
    const Item c = cand.top();
    if (res.size() >= ef && c.d > res.top().d) break;
    cand.pop();
    snapshot(g, c.id, level, ws.nbuf);
    for (uint32_t id : ws.nbuf) {
      if (ws.stamp[id] == ws.epoch) continue;
      ws.stamp[id] = ws.epoch;
      const float d = g.dist(q, g.vec(id));
      if (res.size() < ef || d < res.top().d) {
        cand.push(Item{d, id});
        res.push(Item{d, id});
        if (res.size() > ef) res.pop();
      }
    }
  }
  std::vector<Item> out(res.size());
  for (std::size_t i = out.size(); i-- > 0;) {
    out[i] = res.top();
    res.pop();
  }
  return out;
}

/// Alg. 4 without candidate extension or pruned-connection keeping: scan in
/// increasing (distance, id) and keep e unless some kept r is strictly closer
/// to e than the base node is.
void heuristic(const Graph& g, const std::vector<Item>& sorted, std::size_t m, std::vector<Item>& kept) {
  kept.clear();
  for (const Item& e : sorted) {
    if (kept.size() >= m) break;
    const float* ev = g.vec(e.id);
    bool keep = true;
    for (const Item& r : kept) {
      if (g.dist(ev, g.vec(r.id)) < e.d) {
        keep = false;
        break;
      }
    }
    if (keep) kept.push_back(e);
  }
}

void write_list(uint32_t* l, const std::vector<Item>& items) {
  l[0] = static_cast<uint32_t>(items.size());
  for (std::size_t i = 0; i < items.size(); ++i) l[1 + i] = items[i].id;
}

void connect(Graph& g, uint32_t id, int level, const std::vector<Item>& chosen) {
  const std::size_t cap = g.degree_cap(level);
  std::vector<Item> pool, kept;
  {
    std::lock_guard<std::mutex> lk(g.lock_of(id));
    uint32_t* l = g.list(id, level);
    if (l[0] == 0) {
      write_list(l, chosen);
    } else {
      // Only in parallel builds: another inserter linked to this node on this
      // layer first. Merge its edges with ours and re-prune on overflow.
      const float* q = g.vec(id);
      pool = chosen;
      for (uint32_t j = 0; j < l[0]; ++j) {
        const uint32_t e = l[1 + j];
        bool dup = false;
        for (const Item& c : pool) dup = dup || c.id == e;
        if (!dup) pool.push_back(Item{g.dist(q, g.vec(e)), e});
      }
      std::sort(pool.begin(), pool.end(), item_less);
      if (pool.size() > cap) {
        heuristic(g, pool, cap, kept);
        pool.swap(kept);
      }
      write_list(l, pool);
    }
  }
  for (const Item& nb : chosen) {
    std::lock_guard<std::mutex> lk(g.lock_of(nb.id));
    uint32_t* l = g.list(nb.id, level);
    const uint32_t n = l[0];
    bool present = false;
    for (uint32_t j = 0; j < n; ++j) present = present || l[1 + j] == id;
    if (present) continue;
    if (n < cap) {
      l[1 + n] = id;
      l[0] = n + 1;
      continue;
    }
    const float* nv = g.vec(nb.id);
    pool.clear();
    pool.push_back(Item{nb.d, id});
    for (uint32_t j = 0; j < n; ++j) pool.push_back(Item{g.dist(nv, g.vec(l[1 + j])), l[1 + j]});
    std::sort(pool.begin(), pool.end(), item_less);
    heuristic(g, pool, cap, kept);
    write_list(l, kept);
  }
}

void insert(Graph& g, uint32_t id, Workspace& ws) {
  const int level = g.levels[id];
  const float* q = g.vec(id);
  std::unique_lock<std::mutex> top(g.top_mutex);
  const uint32_t ep = g.entry_point;
  const int top_level = g.max_level;
  if (ep == kNoNode) {
    g.entry_point = id;
    g.max_level = level;
    return;
  }
  if (level <= top_level) top.unlock();  // hold it only when this node becomes the top

  Item cur{g.dist(q, g.vec(ep)), ep};
  for (int lc = top_level; lc > level; --lc) cur = descend(g, q, cur, lc, ws);
  std::vector<Item> eps{cur}, chosen;
  for (int lc = std::min(level, top_level); lc >= 0; --lc) {
    std::vector<Item> w = layer_search(g, q, eps, g.efc, lc, ws);
    w.erase(std::remove_if(w.begin(), w.end(), [id](const Item& c) { return c.id == id; }), w.end());
    heuristic(g, w, g.M, chosen);
    connect(g, id, lc, chosen);
    eps.swap(w);
  }
  if (level > top_level) {
    g.entry_point = id;
    g.max_level = level;
  }
}

}  // namespace

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
  if (p.max_elements > 0) reserve(p.max_elements);
}

float Graph::dist(const float* a, const float* b) const noexcept {
  return distance(params.metric, a, b, padded_dim);
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

void Graph::add_batch(const float* vin, const uint64_t* lin, std::size_t n, std::size_t threads) {
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

  // Sequential part: rows, labels and levels in input order.
  const auto first = static_cast<uint32_t>(count);
  for (std::size_t i = 0; i < n; ++i) {
    const auto id = static_cast<uint32_t>(count);
    prepare_vector(params.metric, vin + i * dim, vec_store.data() + static_cast<std::size_t>(id) * padded_dim,
                   dim);
    label_store[id] = lin[i];
    const double u = 1.0 - static_cast<double>(splitmix64_next(rng_state) >> 11) * 0x1.0p-53;
    const double lv = std::floor(-std::log(u) * mL);
    const int level = static_cast<int>(std::min(lv, static_cast<double>(kMaxLevel)));
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
    label_to_id.emplace(lin[i], id);
    ++count;
  }

  // Parallel part: link with an atomic ticket over the new ids.
  std::size_t nt = threads == 0 ? std::max(1u, std::thread::hardware_concurrency()) : threads;
  nt = std::min(nt, n);
  std::atomic<std::size_t> next{0};
  std::atomic<bool> failed{false};
  std::exception_ptr err;
  std::mutex err_mu;
  auto worker = [&] {
    Workspace ws;
    try {
      while (!failed.load(std::memory_order_relaxed)) {
        const std::size_t i = next.fetch_add(1, std::memory_order_relaxed);
        if (i >= n) break;
        insert(*this, first + static_cast<uint32_t>(i), ws);
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(err_mu);
      if (!err) err = std::current_exception();
      failed.store(true);
    }
  };
  std::vector<std::thread> pool;
  for (std::size_t t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
}

std::size_t Graph::memory_usage() const {
  return count * (padded_dim * sizeof(float) + l0_stride * sizeof(uint32_t) + sizeof(uint64_t) * 2 + 2) +
         upper_used * sizeof(uint32_t);
}

}  // namespace hnswamd
