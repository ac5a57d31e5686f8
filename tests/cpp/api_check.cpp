// The engine's C++ API (include/hnsw/index.hpp, include/hnsw/filter.hpp) used
// the way the reference's C++ callers use it, searching on the GPU. Checks the
// behaviours the reference's C++ tests check (tests/cpp/test_index.cpp:78-154,
// test_filter.cpp, test_persist.cpp, test_concurrency.cpp, test_brute_force.cpp
// of the reference), and writes one index, its queries and the rows of one
// search_batch under the output directory for the oracle to check
// (tests/test_gpu_cpp_api.py).
//
//   api_check <output directory>      exit status 0 iff every check held

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "hnsw/filter.hpp"
#include "hnsw/index.hpp"

namespace {

int g_checks = 0, g_failed = 0;

void check(bool ok, const char* what, int line) {
  ++g_checks;
  if (!ok) {
    ++g_failed;
    std::fprintf(stderr, "FAILED line %d: %s\n", line, what);
  }
}
#define CHECK(cond) check((cond), #cond, __LINE__)

template <typename E>
bool throws(const std::function<void()>& f) {
  try {
    f();
  } catch (const E&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

std::vector<float> clustered(std::size_t n, std::size_t dim, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  std::vector<float> centres(10 * dim);
  for (auto& c : centres) c = 4.0f * nd(rng);
  std::vector<float> out(n * dim);
  for (std::size_t i = 0; i < n; ++i) {
    const std::size_t c = rng() % 10;
    for (std::size_t j = 0; j < dim; ++j) out[i * dim + j] = centres[c * dim + j] + nd(rng);
  }
  return out;
}

std::vector<uint64_t> iota(std::size_t n, uint64_t from = 0) {
  std::vector<uint64_t> v(n);
  for (std::size_t i = 0; i < n; ++i) v[i] = from + i;
  return v;
}

bool same_rows(const std::vector<hnsw::Neighbor>& a, const std::vector<hnsw::Neighbor>& b) {
  if (a.size() != b.size()) return false;
  for (std::size_t i = 0; i < a.size(); ++i) {
    if (a[i].label != b[i].label || std::memcmp(&a[i].distance, &b[i].distance, sizeof(float)) != 0) return false;
  }
  return true;
}

void write_file(const std::string& path, const void* p, std::size_t bytes) {
  std::ofstream f(path, std::ios::binary);
  f.write(static_cast<const char*>(p), static_cast<std::streamsize>(bytes));
  if (!f) throw std::runtime_error("cannot write " + path);
}

}  // namespace

int main(int argc, char** argv) {
  using hnsw::Index;
  using hnsw::Neighbor;
  using hnsw::Params;
  if (argc != 2) {
    std::fprintf(stderr, "usage: api_check <output directory>\n");
    return 2;
  }
  const std::string out_dir = argv[1];

  // Parameters are validated as the reference validates them.
  CHECK(throws<std::invalid_argument>([] { Index bad(Params{0}); }));
  CHECK(throws<std::invalid_argument>([] {
    Params p{4};
    p.M = 1;
    Index bad(p);
  }));
  CHECK(throws<std::invalid_argument>([] {
    Params p{4};
    p.ef_construction = 0;
    Index bad(p);
  }));
  {
    Index ix(Params{4});
    CHECK(throws<std::invalid_argument>([&] { ix.set_ef(0); }));
  }

  // An empty index: no results, padded rows.
  {
    Index ix(Params{8});
    std::vector<float> q(8, 0.0f);
    CHECK(ix.search(q.data(), 5).empty());
    std::vector<Neighbor> out(3);
    ix.search_batch(q.data(), 1, 3, 10, out.data());
    for (const auto& nb : out) CHECK(nb.label == hnsw::kInvalidLabel && std::isinf(nb.distance));
    ix.search_batch(nullptr, 0, 3, 10, nullptr);  // nq = 0 returns before any pointer is read
  }

  // A line of 200 points: the known answer 42, 43, 41.
  {
    Index ix(Params{1});
    for (int i = 0; i < 200; ++i) {
      const float v = static_cast<float>(i);
      ix.add(&v, static_cast<uint64_t>(i));
    }
    const float q = 42.2f;
    const auto r = ix.search(&q, 3, 50);
    CHECK(r.size() == 3 && r[0].label == 42 && r[1].label == 43 && r[2].label == 41);
  }

  // k above the size; ef below k is raised to k; k = 0 is an error.
  {
    const auto data = clustered(5, 4, 1);
    Index ix(Params{4});
    ix.add_batch(data.data(), iota(5).data(), 5);
    const auto r = ix.search(data.data(), 50, 1);
    CHECK(r.size() == 5 && r[0].label == 0);
    for (std::size_t i = 1; i < r.size(); ++i) CHECK(r[i - 1].distance <= r[i].distance);
    CHECK(throws<std::invalid_argument>([&] { ix.search(data.data(), 0); }));
  }

  // Duplicate labels: a failed add inserts nothing.
  {
    const auto data = clustered(4, 4, 2);
    Index ix(Params{4});
    ix.add(data.data(), 7);
    CHECK(throws<std::invalid_argument>([&] { ix.add(data.data() + 4, 7); }));
    const uint64_t dup[2] = {9, 9}, clash[2] = {10, 7};
    CHECK(throws<std::invalid_argument>([&] { ix.add_batch(data.data(), dup, 2); }));
    CHECK(throws<std::invalid_argument>([&] { ix.add_batch(data.data(), clash, 2); }));
    CHECK(ix.size() == 1 && !ix.contains(10));
  }

  // Identical vectors: k distinct labels at distance 0.
  {
    const std::size_t n = 300, dim = 12;
    std::vector<float> data(n * dim, 0.5f);
    Index ix(Params{dim});
    ix.add_batch(data.data(), iota(n).data(), n, 1);
    const auto r = ix.search(data.data(), 10, 64);
    std::set<uint64_t> uniq;
    for (const auto& nb : r) {
      CHECK(nb.distance == 0.0f);
      uniq.insert(nb.label);
    }
    CHECK(r.size() == 10 && uniq.size() == 10);
  }

  // The main index: 3,000 x 32, labels 1000.., searched in one batch and one
  // query at a time; its file, queries and rows go to the oracle.
  const std::size_t n = 3000, dim = 32, nq = 200, k = 10, ef = 64;
  const auto base = clustered(n, dim, 3);
  const auto queries = clustered(nq, dim, 4);
  Index ix(Params{dim});
  ix.add_batch(base.data(), iota(n, 1000).data(), n);
  ix.save(out_dir + "/api.idx");
  std::vector<Neighbor> rows(nq * k);
  ix.search_batch(queries.data(), nq, k, ef, rows.data());
  for (std::size_t i = 0; i < nq; i += 17) {
    const auto one = ix.search(queries.data() + i * dim, k, ef);
    CHECK(same_rows(one, std::vector<Neighbor>(rows.begin() + i * k, rows.begin() + (i + 1) * k)));
  }
  write_file(out_dir + "/queries.f32", queries.data(), queries.size() * sizeof(float));
  write_file(out_dir + "/rows.bin", rows.data(), rows.size() * sizeof(Neighbor));

  // Filters: a BitsetFilter and a FunctionFilter allowing the same labels
  // return the same rows, and only allowed labels.
  {
    hnsw::BitsetFilter bits(1000 + n);
    for (uint64_t l = 1000; l < 1000 + n; l += 5) bits.set(l);
    const hnsw::FunctionFilter fn([](uint64_t l) { return l >= 1000 && (l - 1000) % 5 == 0; });
    std::vector<Neighbor> a(nq * k), b(nq * k);
    ix.search_batch(queries.data(), nq, k, 128, a.data(), 0, &bits);
    ix.search_batch(queries.data(), nq, k, 128, b.data(), 0, &fn);
    CHECK(same_rows(a, b));
    bool only_allowed = true;
    for (const auto& nb : a) only_allowed = only_allowed && (nb.label == hnsw::kInvalidLabel || bits.allowed(nb.label));
    CHECK(only_allowed);
  }

  // Concurrent searches on one index equal the sequential rows.
  {
    std::vector<std::vector<Neighbor>> got(4, std::vector<Neighbor>(nq / 4 * k));
    std::vector<std::thread> ts;
    for (int t = 0; t < 4; ++t) {
      ts.emplace_back([&, t] {
        for (int rep = 0; rep < 3; ++rep) {
          ix.search_batch(queries.data() + t * (nq / 4) * dim, nq / 4, k, ef, got[t].data());
        }
      });
    }
    for (auto& t : ts) t.join();
    for (int t = 0; t < 4; ++t) {
      CHECK(same_rows(got[t], std::vector<Neighbor>(rows.begin() + t * (nq / 4) * k,
                                                    rows.begin() + (t + 1) * (nq / 4) * k)));
    }
  }

  // Soft deletes are never returned; a saved and memory-mapped copy answers
  // the same and refuses changes.
  {
    for (uint64_t l = 1000; l < 1000 + n; l += 3) ix.mark_deleted(l);
    CHECK(ix.is_deleted(1000) && !ix.is_deleted(1001));
    std::vector<Neighbor> a(nq * k);
    ix.search_batch(queries.data(), nq, k, ef, a.data());
    bool none_deleted = true;
    for (const auto& nb : a) none_deleted = none_deleted && (nb.label - 1000) % 3 != 0;
    CHECK(none_deleted);
    ix.save(out_dir + "/api_deleted.idx");
    Index ro = Index::load(out_dir + "/api_deleted.idx", true);
    CHECK(ro.read_only() && ro.size() == n);
    std::vector<Neighbor> b(nq * k);
    ro.search_batch(queries.data(), nq, k, ef, b.data());
    CHECK(same_rows(a, b));
    CHECK(throws<std::runtime_error>([&] { ro.add(base.data(), 999999); }));
    CHECK(throws<std::runtime_error>([&] { Index::load(out_dir + "/missing.idx"); }));
  }

  // Brute force equals a double-precision scan on continuous data.
  {
    const std::size_t bn = 500, bq = 20;
    const auto bd = clustered(bn, dim, 5);
    hnsw::BruteForceIndex bf(dim, hnsw::Metric::L2);
    bf.add(bd.data(), iota(bn, 7).data(), bn);
    std::vector<Neighbor> got(bq * k);
    bf.search_batch(queries.data(), bq, k, got.data());
    for (std::size_t i = 0; i < bq; ++i) {
      std::vector<std::pair<double, uint64_t>> all(bn);
      for (std::size_t r = 0; r < bn; ++r) {
        double s = 0.0;
        for (std::size_t j = 0; j < dim; ++j) {
          const double d = static_cast<double>(queries[i * dim + j]) - bd[r * dim + j];
          s += d * d;
        }
        all[r] = {s, 7 + r};
      }
      std::sort(all.begin(), all.end());
      for (std::size_t j = 0; j < k; ++j) CHECK(got[i * k + j].label == all[j].second);
    }
  }

  std::printf("api_check: %d checks, %d failed\n", g_checks, g_failed);
  return g_failed == 0 ? 0 : 1;
}
