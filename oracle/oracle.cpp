// TEST INFRASTRUCTURE ONLY -- the oracle. Nothing in the product links, loads or
// calls this file; tests/, __graft_entry__.smoke() and the CPU-baseline leg of
// bench.py use it as the checker.
//
// A plain CPU restatement of the reference's query path, written from its
// algorithm (not its code), with three things the reference does not have:
//   * a selectable fp32 summation order, so a distance can be reproduced bit for
//     bit in the order of the reference's scalar kernel, of its AVX2 kernel or of
//     its AVX-512 kernel (the GPU computes in the AVX-512 order);
//   * per-query work counters (distances, layer-0 and upper-layer list reads),
//     which give the algorithmic HBM bytes per query of the roofline figure;
//   * an eligibility view that takes the deleted bytes and a label bitset.
//
// Reference files followed (file:line):
//   distances   src/kernels_scalar.cpp:9-22, src/kernels_avx2.cpp:11-69,
//               src/kernels_avx512.cpp:10-58, src/space.hpp:14-27
//   query prep  src/space.hpp:31-40
//   greedy      src/search_layer.cpp:29-56
//   layer 0     src/search_layer.cpp:58-134 (Alg. 2), eligibility src/index_impl.hpp:259-261
//   query       src/search_layer.cpp:146-171, batch padding :180-194
//   order       src/index_impl.hpp:58-71 (Cand, CandLess, CandGreater)
//   brute force src/brute_force.cpp:15-17,34-69
//   file        docs/FORMAT.md (header offsets, sections, upper offsets)
//
// Pinning (tests/test_oracle.py): every distance order is compared bit for bit
// with the compiled reference's own kernels (oracle/_ref) over dims 1..1024, and
// exactly on small integers (closed form); greedy + layer-0 search and brute
// force are compared with reference-generated golden vectors (tests/golden),
// with the reference itself where it builds, with FP64 brute force, and with
// the 1-D line known answer.
//
// Built with -ffp-contract=off so that no multiply-add is fused except where a
// reference kernel uses an FMA (written here as std::fma).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <limits>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

enum Order { kScalar = 0, kAvx2 = 1, kAvx512 = 2 };
enum MetricId { kL2 = 0, kIP = 1, kCos = 2 };

// ---------------------------------------------------------------------------
// Distances in the three summation orders.
// ---------------------------------------------------------------------------

// Sequential: sum += (a-b)^2 with a separate multiply and add
// (src/kernels_scalar.cpp:9-22; that TU is compiled without FMA).
float l2_scalar(const float* a, const float* b, std::size_t n) {
  float s = 0.0f;
  for (std::size_t i = 0; i < n; ++i) {
    const float d = a[i] - b[i];
    const float p = d * d;
    s = s + p;
  }
  return s;
}
float dot_scalar(const float* a, const float* b, std::size_t n) {
  float s = 0.0f;
  for (std::size_t i = 0; i < n; ++i) {
    const float p = a[i] * b[i];
    s = s + p;
  }
  return s;
}

// 64 partial sums: lane p of accumulator q holds positions 64t + 16q + p
// (src/kernels_avx512.cpp:12-27); whole 16-chunks past the last 64-chunk go to
// accumulator 0 (:28-31); a final partial chunk goes to accumulator 1 (:32-36).
// Then (acc0+acc1)+(acc2+acc3) per lane, then the 16-lane horizontal sum as a
// halving tree: p+8, p+4, p+2, p+1 (GCC's _mm512_reduce_add_ps).
template <bool kL2Metric>
float avx512_order(const float* a, const float* b, std::size_t n) {
  float acc[4][16] = {};
  auto term = [](float x, float y, float acc_in) {
    if (kL2Metric) {
      const float d = x - y;
      return std::fma(d, d, acc_in);
    }
    return std::fma(x, y, acc_in);
  };
  std::size_t i = 0;
  for (; i + 64 <= n; i += 64)
    for (int q = 0; q < 4; ++q)
      for (int p = 0; p < 16; ++p) acc[q][p] = term(a[i + 16 * q + p], b[i + 16 * q + p], acc[q][p]);
  for (; i + 16 <= n; i += 16)
    for (int p = 0; p < 16; ++p) acc[0][p] = term(a[i + p], b[i + p], acc[0][p]);
  if (i < n)
    for (std::size_t p = 0; p < n - i; ++p) acc[1][p] = term(a[i + p], b[i + p], acc[1][p]);
  float v[16];
  for (int p = 0; p < 16; ++p) v[p] = (acc[0][p] + acc[1][p]) + (acc[2][p] + acc[3][p]);
  for (int w = 8; w >= 1; w >>= 1)
    for (int p = 0; p < w; ++p) v[p] = v[p] + v[p + w];
  return v[0];
}

// 32 partial sums over 32-chunks (src/kernels_avx2.cpp:22-38), 8-chunks into
// accumulator 0 (:39-42), lanes combined as (acc0+acc1)+(acc2+acc3), then the
// 8-lane sum of hsum256 (:11-18): p+4, then p+2, then p+1. The element tail
// (:44-47) is compiled FMA-contracted; it never occurs for padded rows, so this
// order is offered only for n % 8 == 0.
template <bool kL2Metric>
float avx2_order(const float* a, const float* b, std::size_t n) {
  if (n % 8 != 0) throw std::invalid_argument("avx2 order: dim must be a multiple of 8");
  float acc[4][8] = {};
  auto term = [](float x, float y, float acc_in) {
    if (kL2Metric) {
      const float d = x - y;
      return std::fma(d, d, acc_in);
    }
    return std::fma(x, y, acc_in);
  };
  std::size_t i = 0;
  for (; i + 32 <= n; i += 32)
    for (int q = 0; q < 4; ++q)
      for (int p = 0; p < 8; ++p) acc[q][p] = term(a[i + 8 * q + p], b[i + 8 * q + p], acc[q][p]);
  for (; i + 8 <= n; i += 8)
    for (int p = 0; p < 8; ++p) acc[0][p] = term(a[i + p], b[i + p], acc[0][p]);
  float v[8];
  for (int p = 0; p < 8; ++p) v[p] = (acc[0][p] + acc[1][p]) + (acc[2][p] + acc[3][p]);
  for (int w = 4; w >= 1; w >>= 1)
    for (int p = 0; p < w; ++p) v[p] = v[p] + v[p + w];
  return v[0];
}

float raw_l2(const float* a, const float* b, std::size_t n, int order) {
  switch (order) {
    case kScalar: return l2_scalar(a, b, n);
    case kAvx2: return avx2_order<true>(a, b, n);
    case kAvx512: return avx512_order<true>(a, b, n);
  }
  throw std::invalid_argument("unknown summation order");
}
float raw_dot(const float* a, const float* b, std::size_t n, int order) {
  switch (order) {
    case kScalar: return dot_scalar(a, b, n);
    case kAvx2: return avx2_order<false>(a, b, n);
    case kAvx512: return avx512_order<false>(a, b, n);
  }
  throw std::invalid_argument("unknown summation order");
}

// L2 -> squared distance; inner product and cosine -> 1 - dot (src/space.hpp:23-26).
float metric_dist(int metric, const float* a, const float* b, std::size_t n, int order) {
  if (metric == kL2) return raw_l2(a, b, n, order);
  return 1.0f - raw_dot(a, b, n, order);
}

// Cosine: scale by float(1/sqrt(sum of squares in double)); zero stays zero
// (src/space.hpp:31-40).
void prepare(int metric, const float* src, float* dst, std::size_t n) {
  if (metric != kCos) {
    std::copy(src, src + n, dst);
    return;
  }
  double ss = 0.0;
  for (std::size_t i = 0; i < n; ++i) ss += static_cast<double>(src[i]) * static_cast<double>(src[i]);
  const float inv = ss > 0.0 ? static_cast<float>(1.0 / std::sqrt(ss)) : 1.0f;
  for (std::size_t i = 0; i < n; ++i) dst[i] = src[i] * inv;
}

// ---------------------------------------------------------------------------
// Index file (docs/FORMAT.md): the whole file is read into memory.
// ---------------------------------------------------------------------------
template <typename T>
T rd(const std::vector<unsigned char>& b, std::size_t at) {
  T v;
  std::memcpy(&v, b.data() + at, sizeof(T));
  return v;
}

struct OIndex {
  std::vector<unsigned char> file;
  uint64_t dim = 0, padded = 0, M = 0, count = 0, ef_search = 0, upper_used = 0;
  int metric = 0;
  int32_t max_level = -1;
  uint32_t entry = 0xFFFFFFFFu;
  const float* vecs = nullptr;
  const uint32_t* l0 = nullptr;
  const uint64_t* labels = nullptr;
  const uint8_t* levels = nullptr;
  const uint32_t* upper = nullptr;
  std::vector<uint64_t> upper_off;
  std::vector<uint8_t> deleted;
  std::size_t num_deleted = 0;

  const float* vec(uint32_t id) const { return vecs + static_cast<std::size_t>(id) * padded; }
  const uint32_t* list(uint32_t id, int level) const {
    if (level == 0) return l0 + static_cast<std::size_t>(id) * (1 + 2 * M);
    return upper + upper_off[id] + static_cast<std::size_t>(level - 1) * (1 + M);
  }
};

OIndex* load_index(const char* path) {
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  auto ix = new OIndex();
  const auto size = static_cast<std::size_t>(in.tellg());
  ix->file.resize(size);
  in.seekg(0);
  in.read(reinterpret_cast<char*>(ix->file.data()), static_cast<std::streamsize>(size));
  const auto& f = ix->file;
  if (size < 256 || std::memcmp(f.data(), "HNSWENG", 8) != 0 || rd<uint32_t>(f, 8) != 1) {
    delete ix;
    throw std::runtime_error("not an index file (version 1)");
  }
  ix->dim = rd<uint64_t>(f, 16);
  ix->metric = static_cast<int>(rd<uint32_t>(f, 24));
  ix->M = rd<uint64_t>(f, 32);
  ix->ef_search = rd<uint64_t>(f, 48);
  ix->count = rd<uint64_t>(f, 72);
  ix->padded = rd<uint64_t>(f, 80);
  ix->max_level = rd<int32_t>(f, 88);
  ix->entry = rd<uint32_t>(f, 92);
  ix->upper_used = rd<uint64_t>(f, 96);
  if (rd<uint64_t>(f, 168) != size) {
    delete ix;
    throw std::runtime_error("file size mismatch");
  }
  const unsigned char* base = f.data();
  ix->vecs = reinterpret_cast<const float*>(base + rd<uint64_t>(f, 120));
  ix->l0 = reinterpret_cast<const uint32_t*>(base + rd<uint64_t>(f, 128));
  ix->labels = reinterpret_cast<const uint64_t*>(base + rd<uint64_t>(f, 136));
  ix->levels = base + rd<uint64_t>(f, 144);
  const uint8_t* del = base + rd<uint64_t>(f, 152);
  ix->upper = reinterpret_cast<const uint32_t*>(base + rd<uint64_t>(f, 160));
  ix->upper_off.resize(ix->count);
  uint64_t pos = 0;
  for (uint64_t i = 0; i < ix->count; ++i) {
    ix->upper_off[i] = pos;
    pos += static_cast<uint64_t>(ix->levels[i]) * (1 + ix->M);
  }
  ix->deleted.assign(del, del + ix->count);
  for (uint8_t d : ix->deleted) ix->num_deleted += d;
  return ix;
}

// ---------------------------------------------------------------------------
// Query path.
// ---------------------------------------------------------------------------
struct Node {
  float d;
  uint32_t id;
};
// Total order by (distance, id) (src/index_impl.hpp:58-71).
inline bool before(const Node& x, const Node& y) { return x.d < y.d || (x.d == y.d && x.id < y.id); }
struct FarthestOnTop {
  bool operator()(const Node& x, const Node& y) const { return before(x, y); }
};
struct NearestOnTop {
  bool operator()(const Node& x, const Node& y) const { return before(y, x); }
};

struct Counters {
  uint64_t n_dist = 0, n_exp0 = 0, n_expu = 0;
};

struct OutRow {
  uint64_t label;
  float distance;
  uint32_t pad;
};

struct LabelFilter {
  const uint64_t* words = nullptr;
  uint64_t size = 0;
  bool allowed(uint64_t label) const { return label < size && ((words[label >> 6] >> (label & 63)) & 1ULL); }
};

// One query: descent, Alg. 2 on layer 0, sort, label map.
std::vector<Node> search_one(const OIndex& ix, const float* q_in, std::size_t k, std::size_t ef_req,
                             int order, const LabelFilter* f, Counters& cnt, std::vector<uint32_t>& mark,
                             uint32_t& epoch) {
  std::vector<Node> out;
  if (ix.count == 0) return out;
  const std::size_t ef = std::max(ef_req == 0 ? static_cast<std::size_t>(ix.ef_search) : ef_req, k);
  std::vector<float> q(ix.padded, 0.0f);
  prepare(ix.metric, q_in, q.data(), ix.dim);
  auto dist = [&](uint32_t id) {
    ++cnt.n_dist;
    return metric_dist(ix.metric, q.data(), ix.vec(id), ix.padded, order);
  };

  Node cur{dist(ix.entry), ix.entry};
  for (int level = ix.max_level; level > 0; --level) {
    for (bool moved = true; moved;) {
      moved = false;
      const uint32_t* l = ix.list(cur.id, level);
      ++cnt.n_expu;
      Node best = cur;
      for (uint32_t j = 0; j < l[0]; ++j) {
        const float d = dist(l[1 + j]);
        if (d < best.d) best = Node{d, l[1 + j]};
      }
      if (best.id != cur.id || best.d != cur.d) {
        cur = best;
        moved = true;
      }
    }
  }

  const bool filtered = f != nullptr || ix.num_deleted > 0;
  auto eligible = [&](uint32_t id) {
    return ix.deleted[id] == 0 && (f == nullptr || f->allowed(ix.labels[id]));
  };
  if (++epoch == 0) {
    std::fill(mark.begin(), mark.end(), 0u);
    epoch = 1;
  }
  std::priority_queue<Node, std::vector<Node>, NearestOnTop> cand;
  std::priority_queue<Node, std::vector<Node>, FarthestOnTop> res;
  mark[cur.id] = epoch;
  cand.push(cur);
  if (!filtered || eligible(cur.id)) res.push(cur);
  while (!cand.empty()) {
    const Node c = cand.top();
    if (res.size() >= ef && c.d > res.top().d) break;
    cand.pop();
    const uint32_t* l = ix.list(c.id, 0);
    ++cnt.n_exp0;
    for (uint32_t j = 0; j < l[0]; ++j) {
      const uint32_t id = l[1 + j];
      if (mark[id] == epoch) continue;
      mark[id] = epoch;
      const float d = dist(id);
      if (res.size() < ef || d < res.top().d) {
        cand.push(Node{d, id});
        if (!filtered || eligible(id)) {
          res.push(Node{d, id});
          if (res.size() > ef) res.pop();
        }
      }
    }
  }
  out.reserve(res.size());
  while (!res.empty()) {
    out.push_back(res.top());
    res.pop();
  }
  std::reverse(out.begin(), out.end());  // ascending (distance, id)
  if (out.size() > k) out.resize(k);
  return out;
}

template <typename F>
void for_each_parallel(std::size_t n, std::size_t threads, F&& fn) {
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  threads = std::min(threads, n);
  if (threads <= 1) {
    for (std::size_t i = 0; i < n; ++i) fn(i, 0);
    return;
  }
  std::atomic<std::size_t> next{0};
  std::exception_ptr err;
  std::atomic<bool> failed{false};
  auto run = [&](std::size_t w) {
    try {
      while (!failed.load()) {
        const std::size_t i = next.fetch_add(1);
        if (i >= n) break;
        fn(i, w);
      }
    } catch (...) {
      if (!failed.exchange(true)) err = std::current_exception();
    }
  };
  std::vector<std::thread> ts;
  for (std::size_t t = 1; t < threads; ++t) ts.emplace_back(run, t);
  run(0);
  for (auto& t : ts) t.join();
  if (err) std::rethrow_exception(err);
}

template <typename F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "unknown exception";
  }
  return -1;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

float orc_l2(const float* a, const float* b, uint64_t n, int order) { return raw_l2(a, b, n, order); }
float orc_dot(const float* a, const float* b, uint64_t n, int order) { return raw_dot(a, b, n, order); }
float orc_dist(int metric, const float* a, const float* b, uint64_t n, int order) {
  return metric_dist(metric, a, b, n, order);
}
void orc_prepare(int metric, const float* src, float* dst, uint64_t n) { prepare(metric, src, dst, n); }

int orc_load(const char* path, void** out) {
  return guard([&] { *out = load_index(path); });
}
void orc_free(void* h) { delete static_cast<OIndex*>(h); }

/// info[0..8] = dim, padded_dim, M, count, max_level, entry_point, metric, ef_search, num_deleted
int orc_info(void* h, int64_t* info) {
  const auto* ix = static_cast<OIndex*>(h);
  info[0] = static_cast<int64_t>(ix->dim);
  info[1] = static_cast<int64_t>(ix->padded);
  info[2] = static_cast<int64_t>(ix->M);
  info[3] = static_cast<int64_t>(ix->count);
  info[4] = ix->max_level;
  info[5] = ix->entry;
  info[6] = ix->metric;
  info[7] = static_cast<int64_t>(ix->ef_search);
  info[8] = static_cast<int64_t>(ix->num_deleted);
  return 0;
}

/// Soft delete by label (src/index.cpp:308-318). Unknown label -> -1.
int orc_mark_deleted(void* h, uint64_t label) {
  return guard([&] {
    auto* ix = static_cast<OIndex*>(h);
    for (uint64_t i = 0; i < ix->count; ++i) {
      if (ix->labels[i] == label) {
        if (ix->deleted[i] == 0) {
          ix->deleted[i] = 1;
          ++ix->num_deleted;
        }
        return;
      }
    }
    throw std::invalid_argument("mark_deleted: unknown label " + std::to_string(label));
  });
}

/// search_batch: out = nq*k rows of {label, distance, pad}, padded with
/// {~0, +inf}; counters (optional) = nq*3 of {n_dist, n_exp0, n_expu}.
int orc_search_batch(void* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef, int order,
                     const uint64_t* filter_words, uint64_t filter_size, void* out, uint64_t* counters,
                     uint64_t threads) {
  return guard([&] {
    if (k == 0) throw std::invalid_argument("search_batch: k must be > 0");
    const auto& ix = *static_cast<OIndex*>(h);
    LabelFilter lf{filter_words, filter_size};
    const LabelFilter* f = filter_words != nullptr ? &lf : nullptr;
    std::size_t nt = threads == 0 ? std::max(1u, std::thread::hardware_concurrency()) : threads;
    nt = std::max<std::size_t>(1, std::min<std::size_t>(nt, nq));
    std::vector<std::vector<uint32_t>> marks(nt, std::vector<uint32_t>(ix.count, 0u));
    std::vector<uint32_t> epochs(nt, 0u);
    auto* rows = static_cast<OutRow*>(out);
    for_each_parallel(nq, nt, [&](std::size_t i, std::size_t w) {
      Counters c;
      const auto r = search_one(ix, qs + i * ix.dim, k, ef, order, f, c, marks[w], epochs[w]);
      for (std::size_t j = 0; j < k; ++j) {
        rows[i * k + j] = j < r.size() ? OutRow{ix.labels[r[j].id], r[j].d, 0}
                                       : OutRow{~0ULL, std::numeric_limits<float>::infinity(), 0};
      }
      if (counters != nullptr) {
        counters[3 * i] = c.n_dist;
        counters[3 * i + 1] = c.n_exp0;
        counters[3 * i + 2] = c.n_expu;
      }
    });
  });
}

/// Distance of query i (prepared as the search prepares it) to the node of
/// each of its k labels, in the given summation order; +inf for an invalid or
/// unknown label. Checks that a reduced-precision traversal returns the exact
/// fp32 distance of every id it returns.
int orc_label_dists(void* h, const float* qs, uint64_t nq, const uint64_t* labels, uint64_t k, int order,
                    float* out, uint64_t threads) {
  return guard([&] {
    const auto& ix = *static_cast<OIndex*>(h);
    std::vector<std::pair<uint64_t, uint32_t>> by_label(ix.count);
    for (uint64_t i = 0; i < ix.count; ++i) by_label[i] = {ix.labels[i], static_cast<uint32_t>(i)};
    std::sort(by_label.begin(), by_label.end());
    for_each_parallel(nq, threads, [&](std::size_t qi, std::size_t) {
      std::vector<float> q(ix.padded, 0.0f);
      prepare(ix.metric, qs + qi * ix.dim, q.data(), ix.dim);
      for (uint64_t j = 0; j < k; ++j) {
        const uint64_t lab = labels[qi * k + j];
        auto it = std::lower_bound(by_label.begin(), by_label.end(), std::make_pair(lab, 0u));
        out[qi * k + j] = (it != by_label.end() && it->first == lab)
                              ? metric_dist(ix.metric, q.data(), ix.vec(it->second), ix.padded, order)
                              : std::numeric_limits<float>::infinity();
      }
    });
  });
}

/// Exact k-NN over n unpadded rows, best k by (distance, label), filter first
/// (src/brute_force.cpp:34-69). labels == null -> 0..n-1.
int orc_brute_force(uint64_t dim, int metric, const float* data, const uint64_t* labels, uint64_t n,
                    const float* qs, uint64_t nq, uint64_t k, int order, const uint64_t* filter_words,
                    uint64_t filter_size, void* out, uint64_t threads) {
  return guard([&] {
    if (k == 0) throw std::invalid_argument("search: k must be > 0");
    std::vector<float> base(n * dim);
    for (uint64_t i = 0; i < n; ++i) prepare(metric, data + i * dim, base.data() + i * dim, dim);
    LabelFilter lf{filter_words, filter_size};
    auto* rows = static_cast<OutRow*>(out);
    for_each_parallel(nq, threads, [&](std::size_t qi, std::size_t) {
      std::vector<float> q(dim);
      prepare(metric, qs + qi * dim, q.data(), dim);
      std::vector<OutRow> all;
      for (uint64_t i = 0; i < n; ++i) {
        const uint64_t lab = labels != nullptr ? labels[i] : i;
        if (filter_words != nullptr && !lf.allowed(lab)) continue;
        all.push_back(OutRow{lab, metric_dist(metric, q.data(), base.data() + i * dim, dim, order), 0});
      }
      auto lt = [](const OutRow& x, const OutRow& y) {
        return x.distance < y.distance || (x.distance == y.distance && x.label < y.label);
      };
      const std::size_t m = std::min<std::size_t>(k, all.size());
      std::partial_sort(all.begin(), all.begin() + static_cast<std::ptrdiff_t>(m), all.end(), lt);
      for (std::size_t j = 0; j < k; ++j) {
        rows[qi * k + j] = j < m ? all[j] : OutRow{~0ULL, std::numeric_limits<float>::infinity(), 0};
      }
    });
  });
}
}
