// The C++ API (include/hnsw/*.hpp) and the hnsw_index_* / hnsw_brute_force_*
// part of the C ABI (include/hnswgpu.h).
//
// hnsw::Index keeps the reference's contract (include/hnsw/index.hpp:31-93 and
// src/search_layer.cpp:146-194 of the reference):
//   * search: k == 0 -> invalid_argument("search: k must be > 0"), null query ->
//     invalid_argument("search: null query"); returns min(k, #eligible) rows;
//   * search_batch: k == 0 -> invalid_argument("search_batch: k must be > 0");
//     nq == 0 returns before the null check; null -> "search_batch: null pointer";
//     rows padded with {kInvalidLabel, +inf};
//   * a filtered search runs iff a filter is given or some node is deleted;
//   * searches share a lock, mutations take it exclusively.
// The search itself runs on the GPU through the device layer.

#include <algorithm>
#include <bit>
#include <cstring>
#include <limits>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/hnsw/index.hpp"
#include "../../../include/hnswgpu.h"
#include "../error.hpp"
#include "../gpu/brute.hpp"
#include "../gpu/build_probe.hpp"
#include "distance.hpp"
#include "graph.hpp"

namespace hnswamd {
void device_search_host(hnswgpu_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                        const uint64_t* filter_words, uint64_t filter_size, int by_id, hnswgpu_neighbor* out,
                        uint32_t* counts);
}

namespace hnsw {

static_assert(sizeof(Neighbor) == sizeof(hnswgpu_neighbor), "Neighbor and hnswgpu_neighbor must match");

// ---------------------------------------------------------------------------
// metric / filter
// ---------------------------------------------------------------------------
Metric parse_metric(std::string_view name) {
  if (name == "l2" || name == "L2" || name == "euclidean") return Metric::L2;
  if (name == "ip" || name == "inner_product" || name == "dot") return Metric::InnerProduct;
  if (name == "cosine" || name == "angular") return Metric::Cosine;
  throw std::invalid_argument("unknown metric '" + std::string(name) + "' (expected l2, ip or cosine)");
}

std::string_view metric_name(Metric m) noexcept {
  switch (m) {
    case Metric::L2: return "l2";
    case Metric::InnerProduct: return "ip";
    case Metric::Cosine: return "cosine";
  }
  return "unknown";
}

float l2_sq(const float* a, const float* b, std::size_t dim) noexcept { return hnswamd::l2_sq(a, b, dim); }
float dot(const float* a, const float* b, std::size_t dim) noexcept { return hnswamd::dot(a, b, dim); }

FunctionFilter::FunctionFilter(std::function<bool(uint64_t)> fn) : fn_(std::move(fn)) {
  if (!fn_) throw std::invalid_argument("FunctionFilter: empty function");
}

BitsetFilter::BitsetFilter(uint64_t num_labels)
    : words_(static_cast<std::size_t>((num_labels + 63) / 64), 0), size_(num_labels) {}

BitsetFilter BitsetFilter::from_labels(const uint64_t* labels, std::size_t n) {
  uint64_t top = 0;
  for (std::size_t i = 0; i < n; ++i) top = std::max(top, labels[i]);
  BitsetFilter f(n == 0 ? 0 : top + 1);
  for (std::size_t i = 0; i < n; ++i) f.set(labels[i]);
  return f;
}

void BitsetFilter::set(uint64_t label, bool value) {
  if (label >= size_) throw std::out_of_range("BitsetFilter::set: label out of range");
  const uint64_t bit = 1ULL << (label & 63);
  if (value) {
    words_[label >> 6] |= bit;
  } else {
    words_[label >> 6] &= ~bit;
  }
}

std::size_t BitsetFilter::count() const noexcept {
  std::size_t c = 0;
  for (uint64_t w : words_) c += static_cast<std::size_t>(std::popcount(w));
  return c;
}

// ---------------------------------------------------------------------------
// Index
// ---------------------------------------------------------------------------
namespace {
/// Waves smaller than this before the first device wave are probed on the
/// host (too few queries to fill the GPU).
constexpr uint32_t kMinDeviceWave = 512;

hnswamd::BuildParams to_build(const Params& p) {
  hnswamd::BuildParams b;
  b.dim = p.dim;
  b.metric = static_cast<int>(p.metric);
  b.M = p.M;
  b.ef_construction = p.ef_construction;
  b.max_elements = p.max_elements;
  b.seed = p.seed;
  return b;
}
}  // namespace

struct Index::Impl {
  std::unique_ptr<hnswamd::Graph> g;
  Params params;
  mutable std::shared_mutex api;
  mutable std::mutex dev_mu;
  mutable hnswgpu_index* dev = nullptr;
  mutable bool dev_stale = true;

  ~Impl() {
    if (dev != nullptr) hnswgpu_free(dev);
  }

  void sync_params() {
    params.dim = g->dim;
    params.metric = static_cast<Metric>(g->params.metric);
    params.M = g->params.M;
    params.ef_construction = g->params.ef_construction;
    params.max_elements = g->params.max_elements;
    params.seed = g->params.seed;
  }

  hnswgpu_index_view view() const {
    const hnswamd::Graph& gg = *g;
    hnswgpu_index_view v{};
    v.dim = gg.dim;
    v.padded_dim = gg.padded_dim;
    v.M = gg.M;
    v.count = gg.count;
    v.ef_search = gg.ef_search;
    v.upper_used = gg.upper_used;
    v.metric = gg.params.metric;
    v.max_level = gg.max_level;
    v.entry_point = gg.entry_point;
    v.vectors = gg.vecs;
    v.level0 = gg.l0;
    v.upper = gg.upper;
    v.labels = gg.labels;
    v.levels = gg.levels;
    v.deleted = gg.deleted.data();
    return v;
  }

  /// Upload on first use; after an add the device copy is refreshed in place,
  /// so a handle given out by device_handle() stays valid (with its options).
  hnswgpu_index* device() const {
    std::lock_guard<std::mutex> lk(dev_mu);
    if (dev != nullptr && !dev_stale) return dev;
    const hnswgpu_index_view v = view();
    const int rc = dev == nullptr ? hnswgpu_upload(&v, &dev) : hnswgpu_reupload(dev, &v);
    if (rc != HNSWGPU_OK) hnswamd::throw_status(rc, hnswgpu_last_error());
    dev_stale = false;
    return dev;
  }

  /// An add changes the arrays a server's slots were sized for.
  void refuse_if_served(const char* what) const {
    std::lock_guard<std::mutex> lk(dev_mu);
    if (dev == nullptr) return;
    int64_t servers = 0;
    const int rc = hnswgpu_get_option(dev, "servers", &servers);
    if (rc != HNSWGPU_OK) hnswamd::throw_status(rc, hnswgpu_last_error());
    if (servers != 0) throw std::invalid_argument(std::string(what) + ": a server is attached to this index");
  }

  void search_rows(const float* qs, std::size_t nq, std::size_t k, std::size_t ef, Neighbor* out,
                   uint32_t* counts, const Filter* f) const {
    const uint64_t* words = nullptr;
    uint64_t size = 0;
    int by_id = 0;
    std::vector<uint64_t> idbits;
    if (f != nullptr) {
      if (const auto* bs = dynamic_cast<const BitsetFilter*>(f)) {
        words = bs->words();
        size = bs->size();
        if (size == 0) {
          idbits.assign(1, 0);
          words = idbits.data();
        }
      } else {
        // Any other predicate: evaluate it once per node, on the host.
        idbits.assign((g->count + 63) / 64 + 1, 0);
        for (std::size_t i = 0; i < g->count; ++i) {
          if (f->allowed(g->labels[i])) idbits[i >> 6] |= 1ULL << (i & 63);
        }
        words = idbits.data();
        size = g->count;
        by_id = 1;
      }
    }
    hnswgpu_index* d = g->count > 0 ? device() : nullptr;
    if (d == nullptr) {
      for (std::size_t i = 0; i < nq * k; ++i) {
        out[i] = Neighbor{kInvalidLabel, std::numeric_limits<float>::infinity()};
      }
      if (counts != nullptr) std::fill(counts, counts + nq, 0u);
      return;
    }
    try {
      hnswamd::device_search_host(d, qs, nq, k, ef, words, size, by_id, reinterpret_cast<hnswgpu_neighbor*>(out),
                                  counts);
    } catch (const hnswamd::DeviceError&) {
      throw;
    }
  }
};

Index::Index(const Params& params) : impl_(std::make_unique<Impl>()) {
  impl_->g = std::make_unique<hnswamd::Graph>(to_build(params));
  impl_->params = params;
}
Index::Index(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}
Index::~Index() = default;
Index::Index(Index&&) noexcept = default;
Index& Index::operator=(Index&&) noexcept = default;

void Index::add(const float* vec, uint64_t label) { add_batch(vec, &label, 1, 1); }

void Index::add_batch(const float* vecs, const uint64_t* labels, std::size_t n, std::size_t num_threads) {
  if (n == 0) return;
  if (vecs == nullptr || labels == nullptr) throw std::invalid_argument("add_batch: null input");
  std::unique_lock<std::shared_mutex> lk(impl_->api);
  impl_->refuse_if_served("add_batch");
  impl_->g->add_batch(vecs, labels, n, num_threads);
  std::lock_guard<std::mutex> dk(impl_->dev_mu);
  impl_->dev_stale = true;
}

void Index::add_batch_device(const float* vecs, const uint64_t* labels, std::size_t n, std::size_t num_threads,
                             DeviceBuildStats* stats) {
  if (n == 0) return;
  if (vecs == nullptr || labels == nullptr) throw std::invalid_argument("add_batch: null input");
  std::unique_lock<std::shared_mutex> lk(impl_->api);
  impl_->refuse_if_served("add_batch");
  hnswamd::DeviceProbeStats st;
  {
    // The probe's device session ends (and frees its memory) with the call.
    impl_->g->add_batch(vecs, labels, n, num_threads, hnswamd::make_device_probe(kMinDeviceWave, &st));
  }
  if (stats != nullptr) *stats = DeviceBuildStats{st.waves, st.device_nodes, st.host_nodes, st.device_ms, st.refresh_ms};
  std::lock_guard<std::mutex> dk(impl_->dev_mu);
  impl_->dev_stale = true;
}

std::vector<Neighbor> Index::search(const float* q, std::size_t k, std::size_t ef, const Filter* f) const {
  if (k == 0) throw std::invalid_argument("search: k must be > 0");
  if (q == nullptr) throw std::invalid_argument("search: null query");
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  std::vector<Neighbor> rows(k);
  uint32_t n = 0;
  impl_->search_rows(q, 1, k, ef, rows.data(), &n, f);
  rows.resize(n);
  return rows;
}

void Index::search_batch(const float* qs, std::size_t nq, std::size_t k, std::size_t ef, Neighbor* out,
                         std::size_t /*num_threads*/, const Filter* f) const {
  if (k == 0) throw std::invalid_argument("search_batch: k must be > 0");
  if (nq == 0) return;
  if (qs == nullptr || out == nullptr) throw std::invalid_argument("search_batch: null pointer");
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  impl_->search_rows(qs, nq, k, ef, out, nullptr, f);
}

void Index::mark_deleted(uint64_t label) {
  std::unique_lock<std::shared_mutex> lk(impl_->api);
  hnswamd::Graph& g = *impl_->g;
  auto it = g.label_to_id.find(label);
  if (it == g.label_to_id.end()) throw std::invalid_argument("mark_deleted: unknown label " + std::to_string(label));
  if (g.deleted[it->second] != 0) return;
  g.deleted[it->second] = 1;
  ++g.num_deleted;
  std::lock_guard<std::mutex> dk(impl_->dev_mu);
  if (impl_->dev != nullptr && !impl_->dev_stale) {
    const int rc = hnswgpu_set_deleted(impl_->dev, it->second, 1);
    if (rc != HNSWGPU_OK) hnswamd::throw_status(rc, hnswgpu_last_error());
  }
}

bool Index::is_deleted(uint64_t label) const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  auto it = impl_->g->label_to_id.find(label);
  return it != impl_->g->label_to_id.end() && impl_->g->deleted[it->second] != 0;
}

bool Index::contains(uint64_t label) const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  return impl_->g->label_to_id.count(label) != 0;
}

std::size_t Index::size() const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  return impl_->g->count;
}
std::size_t Index::dim() const noexcept { return impl_->g->dim; }
Metric Index::metric() const noexcept { return impl_->params.metric; }
const Params& Index::params() const noexcept { return impl_->params; }
std::size_t Index::capacity() const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  return impl_->g->capacity;
}
std::size_t Index::ef() const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  return impl_->g->ef_search;
}
void Index::set_ef(std::size_t ef) {
  if (ef == 0) throw std::invalid_argument("set_ef: ef must be > 0");
  std::unique_lock<std::shared_mutex> lk(impl_->api);
  impl_->g->ef_search = ef;
  std::lock_guard<std::mutex> dk(impl_->dev_mu);
  if (impl_->dev != nullptr) {
    const int rc = hnswgpu_set_ef(impl_->dev, ef);
    if (rc != HNSWGPU_OK) hnswamd::throw_status(rc, hnswgpu_last_error());
  }
}
int Index::max_level() const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  return impl_->g->max_level;
}
bool Index::read_only() const noexcept { return impl_->g->read_only; }
std::size_t Index::memory_usage() const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  return impl_->g->memory_usage();
}
std::vector<float> Index::get_vector(uint64_t label) const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  auto it = impl_->g->label_to_id.find(label);
  if (it == impl_->g->label_to_id.end()) throw std::invalid_argument("get_vector: unknown label");
  const float* v = impl_->g->vec(it->second);
  return std::vector<float>(v, v + impl_->g->dim);
}

void Index::save(const std::string& path) const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  impl_->g->save(path);
}

Index Index::load(const std::string& path, bool mmap) {
  auto impl = std::make_unique<Impl>();
  impl->g = hnswamd::Graph::load(path, mmap);
  impl->sync_params();
  return Index(std::move(impl));
}

void* Index::device_handle() const {
  std::shared_lock<std::shared_mutex> lk(impl_->api);
  return impl_->device();
}

// ---------------------------------------------------------------------------
// BruteForceIndex
// ---------------------------------------------------------------------------
struct BruteForceIndex::Device {
  hnswamd::gpu::BruteSet* set = nullptr;
  ~Device() { hnswamd::gpu::brute_free(set); }
};

BruteForceIndex::BruteForceIndex(std::size_t dim, Metric metric) : dim_(dim), metric_(metric) {
  if (dim == 0) throw std::invalid_argument("BruteForceIndex: dim must be > 0");
}
BruteForceIndex::~BruteForceIndex() = default;
BruteForceIndex::BruteForceIndex(BruteForceIndex&&) noexcept = default;
BruteForceIndex& BruteForceIndex::operator=(BruteForceIndex&&) noexcept = default;

void BruteForceIndex::add(const float* vecs, const uint64_t* labels, std::size_t n) {
  const std::size_t old = labels_.size();
  data_.resize((old + n) * dim_);
  labels_.resize(old + n);
  for (std::size_t i = 0; i < n; ++i) {
    hnswamd::prepare_vector(static_cast<int>(metric_), vecs + i * dim_, data_.data() + (old + i) * dim_, dim_);
    labels_[old + i] = labels != nullptr ? labels[i] : static_cast<uint64_t>(old + i);
  }
  dev_.reset();
}

std::vector<Neighbor> BruteForceIndex::search(const float* q, std::size_t k, const Filter* f) const {
  if (k == 0) throw std::invalid_argument("search: k must be > 0");
  std::vector<Neighbor> rows(k);
  uint32_t n = 0;
  run(q, 1, k, rows.data(), &n, f);
  rows.resize(n);
  return rows;
}

void BruteForceIndex::run(const float* qs, std::size_t nq, std::size_t k, Neighbor* out, uint32_t* counts,
                          const Filter* f) const {
  if (labels_.empty()) {
    for (std::size_t i = 0; i < nq * k; ++i) out[i] = Neighbor{kInvalidLabel, std::numeric_limits<float>::infinity()};
    if (counts != nullptr) std::fill(counts, counts + nq, 0u);
    return;
  }
  if (!dev_) {
    dev_ = std::make_unique<Device>();
    dev_->set = hnswamd::gpu::brute_create(dim_, static_cast<int>(metric_), data_.data(), labels_.data(),
                                           labels_.size());
    for (const auto& [name, value] : options_) hnswamd::gpu::brute_set_option(dev_->set, name, value);
  }
  std::vector<uint64_t> allow;
  const uint64_t* allow_ptr = nullptr;
  if (f != nullptr) {
    allow.assign((labels_.size() + 63) / 64 + 1, 0);
    for (std::size_t i = 0; i < labels_.size(); ++i) {
      if (f->allowed(labels_[i])) allow[i >> 6] |= 1ULL << (i & 63);
    }
    allow_ptr = allow.data();
  }
  hnswamd::gpu::brute_search(dev_->set, qs, nq, k, allow_ptr, reinterpret_cast<hnswgpu_neighbor*>(out), counts);
}

void BruteForceIndex::set_device_option(const std::string& name, int64_t value) {
  if (name.rfind("last_", 0) == 0) throw std::invalid_argument("read-only brute-force option: " + name);
  if (dev_) hnswamd::gpu::brute_set_option(dev_->set, name, value);
  bool found = false;
  for (auto& [n, v] : options_) {
    if (n == name) {
      v = value;
      found = true;
    }
  }
  if (!found) {
    // Validate the name before keeping it for a device set created later.
    if (name != "force_scan" && name != "kernel_timing" && name != "chunk_rows" && name != "mfma_variant") {
      throw std::invalid_argument("unknown brute-force option: " + name);
    }
    options_.emplace_back(name, value);
  }
}

int64_t BruteForceIndex::device_option(const std::string& name) const {
  if (dev_) return hnswamd::gpu::brute_get_option(dev_->set, name);
  for (const auto& [n, v] : options_) {
    if (n == name) return v;
  }
  if (name == "mfma_variant") return 1;
  if (name == "force_scan" || name == "kernel_timing" || name == "chunk_rows" || name.rfind("last_", 0) == 0) return 0;
  throw std::invalid_argument("unknown brute-force option: " + name);
}

void BruteForceIndex::search_batch(const float* qs, std::size_t nq, std::size_t k, Neighbor* out,
                                   std::size_t /*num_threads*/, const Filter* f) const {
  if (k == 0) throw std::invalid_argument("search: k must be > 0");
  if (nq == 0) return;
  if (qs == nullptr || out == nullptr) throw std::invalid_argument("search_batch: null pointer");
  run(qs, nq, k, out, nullptr, f);
}

}  // namespace hnsw

// ---------------------------------------------------------------------------
// C ABI: hnsw_index_* and hnsw_brute_force_*
// ---------------------------------------------------------------------------
struct hnsw_index {
  hnsw::Index idx;
  explicit hnsw_index(hnsw::Index&& i) : idx(std::move(i)) {}
};
struct hnsw_brute_force {
  hnsw::BruteForceIndex bf;
  hnsw_brute_force(std::size_t d, hnsw::Metric m) : bf(d, m) {}
};

namespace {
template <typename F>
int guarded(F&& f) {
  try {
    f();
    return HNSWGPU_OK;
  } catch (...) {
    return hnswamd::status_from_current_exception();
  }
}

hnsw::Metric metric_of(int32_t m) {
  if (m < 0 || m > 2) throw std::invalid_argument("unknown metric id " + std::to_string(m));
  return static_cast<hnsw::Metric>(m);
}

struct LabelBits final : hnsw::Filter {
  const uint64_t* words;
  uint64_t size;
  LabelBits(const uint64_t* w, uint64_t s) : words(w), size(s) {}
  bool allowed(uint64_t l) const override { return l < size && ((words[l >> 6] >> (l & 63)) & 1ULL) != 0; }
};

/// Wrap raw words as a BitsetFilter (copy) so the GPU path uploads them directly.
hnsw::BitsetFilter as_bitset(const uint64_t* words, uint64_t size) {
  hnsw::BitsetFilter f(size);
  for (uint64_t l = 0; l < size; ++l) {
    if ((words[l >> 6] >> (l & 63)) & 1ULL) f.set(l);
  }
  return f;
}
}  // namespace

extern "C" {

int hnsw_index_create(uint64_t dim, int32_t metric, uint64_t M, uint64_t ef_construction, uint64_t max_elements,
                      uint64_t seed, hnsw_index** out) {
  return guarded([&] {
    if (out == nullptr) throw std::invalid_argument("create: null out");
    hnsw::Params p;
    p.dim = dim;
    p.metric = metric_of(metric);
    p.M = M;
    p.ef_construction = ef_construction;
    p.max_elements = max_elements;
    p.seed = seed;
    *out = new hnsw_index(hnsw::Index(p));
  });
}

int hnsw_index_load(const char* path, int mmap, hnsw_index** out) {
  return guarded([&] {
    if (path == nullptr || out == nullptr) throw std::invalid_argument("load: null pointer");
    *out = new hnsw_index(hnsw::Index::load(path, mmap != 0));
  });
}

void hnsw_index_free(hnsw_index* h) { delete h; }

int hnsw_index_save(hnsw_index* h, const char* path) {
  return guarded([&] { h->idx.save(path); });
}

int hnsw_index_add_batch(hnsw_index* h, const float* vecs, const uint64_t* labels, uint64_t n, uint64_t threads) {
  return guarded([&] { h->idx.add_batch(vecs, labels, n, threads); });
}

int hnsw_index_add_batch_device(hnsw_index* h, const float* vecs, const uint64_t* labels, uint64_t n,
                                uint64_t threads, hnswgpu_build_stats* stats) {
  static_assert(sizeof(hnswgpu_build_stats) == sizeof(hnsw::Index::DeviceBuildStats), "build stats layout");
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("add_batch: null handle");
    hnsw::Index::DeviceBuildStats st;
    h->idx.add_batch_device(vecs, labels, n, threads, &st);
    if (stats != nullptr) {
      *stats = hnswgpu_build_stats{st.waves, st.device_nodes, st.host_nodes, st.device_ms, st.refresh_ms};
    }
  });
}

int hnsw_index_search_batch(hnsw_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                            hnswgpu_neighbor* out, uint64_t threads, const uint64_t* fw, uint64_t fsize) {
  return guarded([&] {
    if (fw != nullptr) {
      const hnsw::BitsetFilter f = as_bitset(fw, fsize);
      h->idx.search_batch(qs, nq, k, ef, reinterpret_cast<hnsw::Neighbor*>(out), threads, &f);
    } else {
      h->idx.search_batch(qs, nq, k, ef, reinterpret_cast<hnsw::Neighbor*>(out), threads, nullptr);
    }
  });
}

int hnsw_index_search_batch_pred(hnsw_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                                 hnswgpu_neighbor* out, uint64_t threads, hnsw_label_predicate allowed, void* ctx) {
  return guarded([&] {
    if (allowed == nullptr) {
      h->idx.search_batch(qs, nq, k, ef, reinterpret_cast<hnsw::Neighbor*>(out), threads, nullptr);
      return;
    }
    const hnsw::FunctionFilter f([allowed, ctx](uint64_t label) { return allowed(label, ctx) != 0; });
    h->idx.search_batch(qs, nq, k, ef, reinterpret_cast<hnsw::Neighbor*>(out), threads, &f);
  });
}

int hnsw_index_search(hnsw_index* h, const float* q, uint64_t k, uint64_t ef, const uint64_t* fw, uint64_t fsize,
                      hnswgpu_neighbor* out, uint64_t* n_out) {
  return guarded([&] {
    std::vector<hnsw::Neighbor> r;
    if (fw != nullptr) {
      const hnsw::BitsetFilter f = as_bitset(fw, fsize);
      r = h->idx.search(q, k, ef, &f);
    } else {
      r = h->idx.search(q, k, ef, nullptr);
    }
    for (std::size_t i = 0; i < r.size(); ++i) out[i] = hnswgpu_neighbor{r[i].label, r[i].distance, 0u};
    if (n_out != nullptr) *n_out = r.size();
  });
}

int hnsw_index_mark_deleted(hnsw_index* h, uint64_t label) {
  return guarded([&] { h->idx.mark_deleted(label); });
}
int hnsw_index_is_deleted(hnsw_index* h, uint64_t label, int* out) {
  return guarded([&] { *out = h->idx.is_deleted(label) ? 1 : 0; });
}
int hnsw_index_contains(hnsw_index* h, uint64_t label, int* out) {
  return guarded([&] { *out = h->idx.contains(label) ? 1 : 0; });
}
int hnsw_index_set_ef(hnsw_index* h, uint64_t ef) {
  return guarded([&] { h->idx.set_ef(ef); });
}
int hnsw_index_get_vector(hnsw_index* h, uint64_t label, float* out) {
  return guarded([&] {
    const auto v = h->idx.get_vector(label);
    std::memcpy(out, v.data(), v.size() * sizeof(float));
  });
}
int hnsw_index_info(hnsw_index* h, int64_t* info) {
  return guarded([&] {
    info[0] = static_cast<int64_t>(h->idx.size());
    info[1] = static_cast<int64_t>(h->idx.dim());
    info[2] = static_cast<int64_t>(h->idx.metric());
    info[3] = static_cast<int64_t>(h->idx.params().M);
    info[4] = static_cast<int64_t>(h->idx.params().ef_construction);
    info[5] = static_cast<int64_t>(h->idx.params().max_elements);
    info[6] = static_cast<int64_t>(h->idx.ef());
    info[7] = h->idx.max_level();
    info[8] = h->idx.read_only() ? 1 : 0;
    info[9] = static_cast<int64_t>(h->idx.memory_usage());
  });
}
int hnsw_index_device(hnsw_index* h, hnswgpu_index** out) {
  return guarded([&] { *out = static_cast<hnswgpu_index*>(h->idx.device_handle()); });
}

int hnsw_brute_force_create(uint64_t dim, int32_t metric, hnsw_brute_force** out) {
  return guarded([&] { *out = new hnsw_brute_force(dim, metric_of(metric)); });
}
void hnsw_brute_force_free(hnsw_brute_force* h) { delete h; }
int hnsw_brute_force_add(hnsw_brute_force* h, const float* vecs, const uint64_t* labels, uint64_t n) {
  return guarded([&] { h->bf.add(vecs, labels, n); });
}
int hnsw_brute_force_set_option(hnsw_brute_force* h, const char* name, int64_t value) {
  return guarded([&] {
    if (name == nullptr) throw std::invalid_argument("null option name");
    h->bf.set_device_option(name, value);
  });
}
int hnsw_brute_force_get_option(hnsw_brute_force* h, const char* name, int64_t* value) {
  return guarded([&] {
    if (name == nullptr || value == nullptr) throw std::invalid_argument("null pointer");
    *value = h->bf.device_option(name);
  });
}
int hnsw_brute_force_search_batch(hnsw_brute_force* h, const float* qs, uint64_t nq, uint64_t k,
                                  hnswgpu_neighbor* out, uint64_t threads, const uint64_t* fw, uint64_t fsize) {
  return guarded([&] {
    if (fw != nullptr) {
      const LabelBits f(fw, fsize);
      h->bf.search_batch(qs, nq, k, reinterpret_cast<hnsw::Neighbor*>(out), threads, &f);
    } else {
      h->bf.search_batch(qs, nq, k, reinterpret_cast<hnsw::Neighbor*>(out), threads, nullptr);
    }
  });
}

}  // extern "C"
