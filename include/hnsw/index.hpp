#pragma once

// The C++ API of the engine: the reference's hnsw::Index and BruteForceIndex
// (include/hnsw/index.hpp:14-118 of the reference) with the same signatures and
// semantics, searching on the GPU. Indexes are built on the host (the
// reference's algorithm) and stored in the reference's file format, so a file
// written by either library loads in the other. The device copy is uploaded on
// first search and kept resident; add() marks it stale, mark_deleted() and
// set_ef() update it in place.

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "hnsw/filter.hpp"
#include "hnsw/metric.hpp"

namespace hnsw {

struct Params {
  std::size_t dim = 0;
  Metric metric = Metric::L2;
  std::size_t M = 16;
  std::size_t ef_construction = 200;
  std::size_t max_elements = 0;  ///< 0 = grow geometrically
  uint64_t seed = 100;
};

struct Neighbor {
  uint64_t label;
  float distance;
};

/// Label that pads search_batch rows when fewer than k results exist.
inline constexpr uint64_t kInvalidLabel = ~uint64_t{0};

/// Thread safety: search / search_batch / accessors may run concurrently;
/// add / add_batch / mark_deleted / set_ef are exclusive.
class Index {
 public:
  explicit Index(const Params& params);
  ~Index();
  Index(Index&&) noexcept;
  Index& operator=(Index&&) noexcept;
  Index(const Index&) = delete;
  Index& operator=(const Index&) = delete;

  void add(const float* vec, uint64_t label);
  void add_batch(const float* vecs, const uint64_t* labels, std::size_t n, std::size_t num_threads = 0);
  /// What add_batch_device did: waves probed on the GPU, nodes whose layer-0
  /// insertion search ran there or on the host, GPU time of those searches
  /// and host time spent refreshing the device copy of the graph.
  struct DeviceBuildStats {
    uint64_t waves = 0, device_nodes = 0, host_nodes = 0;
    double device_ms = 0.0, refresh_ms = 0.0;
  };
  /// add_batch with the insertion searches of level-0 nodes on the GPU (no
  /// counterpart in the reference). The graph equals add_batch's with the
  /// same num_threads (>= 2: insertion waves; 1: one node per wave, the
  /// reference's sequential order).
  void add_batch_device(const float* vecs, const uint64_t* labels, std::size_t n, std::size_t num_threads = 0,
                        DeviceBuildStats* stats = nullptr);

  /// k nearest neighbours sorted by increasing distance; ef = 0 uses the index
  /// default, ef is clamped to >= k; returns min(k, #eligible) results.
  std::vector<Neighbor> search(const float* q, std::size_t k, std::size_t ef = 0,
                               const Filter* f = nullptr) const;
  /// nq queries into nq*k rows, padded with {kInvalidLabel, +inf}. num_threads
  /// is accepted for signature compatibility; the work runs on the GPU.
  void search_batch(const float* qs, std::size_t nq, std::size_t k, std::size_t ef, Neighbor* out,
                    std::size_t num_threads = 0, const Filter* f = nullptr) const;

  void mark_deleted(uint64_t label);
  bool is_deleted(uint64_t label) const;
  bool contains(uint64_t label) const;

  std::size_t size() const;
  std::size_t dim() const noexcept;
  Metric metric() const noexcept;
  const Params& params() const noexcept;
  std::size_t capacity() const;
  std::size_t ef() const;
  void set_ef(std::size_t ef);
  int max_level() const;
  bool read_only() const noexcept;
  std::size_t memory_usage() const;
  std::vector<float> get_vector(uint64_t label) const;

  void save(const std::string& path) const;
  static Index load(const std::string& path, bool mmap = false);

  struct Impl;
  /// The device-layer handle (struct hnswgpu_index from hnswgpu.h), uploading
  /// or refreshing it first if needed.
  void* device_handle() const;

 private:
  explicit Index(std::unique_ptr<Impl> impl);
  std::unique_ptr<Impl> impl_;
};

/// Exact search over all vectors (GPU); the recall oracle.
class BruteForceIndex {
 public:
  BruteForceIndex(std::size_t dim, Metric metric);
  ~BruteForceIndex();
  BruteForceIndex(BruteForceIndex&&) noexcept;
  BruteForceIndex& operator=(BruteForceIndex&&) noexcept;
  void add(const float* vecs, const uint64_t* labels, std::size_t n);
  std::vector<Neighbor> search(const float* q, std::size_t k, const Filter* f = nullptr) const;
  void search_batch(const float* qs, std::size_t nq, std::size_t k, Neighbor* out, std::size_t num_threads = 0,
                    const Filter* f = nullptr) const;
  std::size_t size() const noexcept { return labels_.size(); }
  std::size_t dim() const noexcept { return dim_; }

  /// GPU options (no counterpart in the reference): force_scan, kernel_timing,
  /// chunk_rows; read-only last_mfma, last_redo, last_chunks, last_gemm_ns,
  /// last_total_ns describe the last search.
  void set_device_option(const std::string& name, int64_t value);
  int64_t device_option(const std::string& name) const;

  struct Device;

 private:
  void run(const float* qs, std::size_t nq, std::size_t k, Neighbor* out, uint32_t* counts,
           const Filter* f) const;
  std::size_t dim_;
  Metric metric_;
  std::vector<float> data_;
  std::vector<uint64_t> labels_;
  std::vector<std::pair<std::string, int64_t>> options_;
  mutable std::unique_ptr<Device> dev_;
};

}  // namespace hnsw
