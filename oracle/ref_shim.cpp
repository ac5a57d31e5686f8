// TEST INFRASTRUCTURE ONLY. Not part of the product.
//
// A C driver over the reference's public C++ API (include/hnsw/index.hpp,
// include/hnsw/filter.hpp, include/hnsw/metric.hpp). It is compiled together
// with the reference's own sources (src/*.cpp, left where they lie) into
// oracle/_ref/libhnswref.so by oracle/Makefile. Tests and the CPU-baseline leg
// of bench.py call it through ctypes to (1) build and save reference-format
// index files, (2) run the reference's search_batch / BruteForceIndex on them.
//
// Every entry point returns 0 on success and -1 on a C++ exception, whose
// message is kept for ref_last_error().

#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "hnsw/filter.hpp"
#include "hnsw/index.hpp"
#include "hnsw/metric.hpp"

namespace {
thread_local std::string g_err;

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

struct RefOut {
  uint64_t label;
  float distance;
  uint32_t pad;
};
static_assert(sizeof(hnsw::Neighbor) == 16, "Neighbor layout");
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

/// Force the reference's distance kernels ("scalar", "avx2", "avx512").
int ref_set_isa(const char* name) {
  return guard([&] { hnsw::simd::set_active(hnsw::simd::parse_isa(name)); });
}

int ref_active_isa(char* buf, int len) {
  const auto n = hnsw::simd::isa_name(hnsw::simd::active().isa);
  std::snprintf(buf, static_cast<std::size_t>(len), "%.*s", static_cast<int>(n.size()), n.data());
  return 0;
}

/// Build an index with Index::add_batch and save it to `path`.
int ref_build_save(const char* path, uint64_t dim, int metric, uint64_t M, uint64_t efc, uint64_t seed,
                   const float* data, const uint64_t* labels, uint64_t n, uint64_t threads) {
  return guard([&] {
    hnsw::Params p;
    p.dim = dim;
    p.metric = static_cast<hnsw::Metric>(metric);
    p.M = M;
    p.ef_construction = efc;
    p.seed = seed;
    hnsw::Index idx(p);
    idx.add_batch(data, labels, n, threads);
    idx.save(path);
  });
}

int ref_load(const char* path, int mmap, void** out) {
  return guard([&] { *out = new hnsw::Index(hnsw::Index::load(path, mmap != 0)); });
}

void ref_free(void* h) { delete static_cast<hnsw::Index*>(h); }

int ref_mark_deleted(void* h, uint64_t label) {
  return guard([&] { static_cast<hnsw::Index*>(h)->mark_deleted(label); });
}

int ref_set_ef(void* h, uint64_t ef) {
  return guard([&] { static_cast<hnsw::Index*>(h)->set_ef(ef); });
}

int ref_save(void* h, const char* path) {
  return guard([&] { static_cast<hnsw::Index*>(h)->save(path); });
}

/// search_batch with an optional BitsetFilter given as label-indexed words.
int ref_search_batch(void* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef, void* out,
                     uint64_t threads, const uint64_t* filter_words, uint64_t filter_size) {
  return guard([&] {
    auto* idx = static_cast<hnsw::Index*>(h);
    std::unique_ptr<hnsw::BitsetFilter> f;
    if (filter_words != nullptr) {
      f = std::make_unique<hnsw::BitsetFilter>(filter_size);
      for (uint64_t l = 0; l < filter_size; ++l) {
        if ((filter_words[l >> 6] >> (l & 63)) & 1ULL) f->set(l);
      }
    }
    idx->search_batch(qs, nq, k, ef, static_cast<hnsw::Neighbor*>(out), threads, f.get());
  });
}

/// BruteForceIndex over n vectors (labels may be null = 0..n-1).
int ref_brute_force(uint64_t dim, int metric, const float* data, const uint64_t* labels, uint64_t n,
                    const float* qs, uint64_t nq, uint64_t k, void* out, uint64_t threads,
                    const uint64_t* filter_words, uint64_t filter_size) {
  return guard([&] {
    hnsw::BruteForceIndex bf(dim, static_cast<hnsw::Metric>(metric));
    bf.add(data, labels, n);
    std::unique_ptr<hnsw::BitsetFilter> f;
    if (filter_words != nullptr) {
      f = std::make_unique<hnsw::BitsetFilter>(filter_size);
      for (uint64_t l = 0; l < filter_size; ++l) {
        if ((filter_words[l >> 6] >> (l & 63)) & 1ULL) f->set(l);
      }
    }
    bf.search_batch(qs, nq, k, static_cast<hnsw::Neighbor*>(out), threads, f.get());
  });
}

/// The reference's parse_metric: the metric id, or -1 when it throws.
int ref_parse_metric(const char* name) {
  int id = -1;
  if (guard([&] { id = static_cast<int>(hnsw::parse_metric(name)); }) != 0) return -1;
  return id;
}

/// The reference's active l2_sq / dot kernels on raw rows.
float ref_l2_sq(const float* a, const float* b, uint64_t dim) { return hnsw::l2_sq(a, b, dim); }
float ref_dot(const float* a, const float* b, uint64_t dim) { return hnsw::dot(a, b, dim); }
}
