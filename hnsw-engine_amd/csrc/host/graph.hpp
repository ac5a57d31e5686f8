#pragma once

// Host-side index state in the reference's on-disk layout (docs/FORMAT.md):
// padded fp32 rows, a fixed-stride layer-0 adjacency array, per-node upper-layer
// blocks addressed by a prefix sum of levels, u64 labels, u8 levels and u8
// tombstones. The same arrays are what the device upload copies into HBM.
//
// The state either owns its arrays (built here, or loaded by reading the file)
// or points into a read-only file mapping (mmap load).

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <memory>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

namespace hnswamd {

inline constexpr std::size_t kRowAlign = 64;      // bytes; file sections and rows
inline constexpr std::size_t kPadFloats = 16;     // rows padded to 16 floats
inline constexpr uint32_t kNoNode = 0xFFFFFFFFu;  // empty entry point
inline constexpr int kMaxLevel = 64;

inline std::size_t round_up(std::size_t v, std::size_t m) { return (v + m - 1) / m * m; }

/// 64-byte aligned, zero-initialised growable buffer of trivially copyable T.
template <typename T>
class AlignedArray {
 public:
  AlignedArray() = default;
  AlignedArray(const AlignedArray&) = delete;
  AlignedArray& operator=(const AlignedArray&) = delete;
  ~AlignedArray() { std::free(p_); }
  T* data() const noexcept { return p_; }
  std::size_t size() const noexcept { return n_; }
  /// Grow to n elements; new elements are zero, old ones are kept.
  void grow(std::size_t n);

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

/// Read-only POSIX mapping of a whole file.
class FileMap {
 public:
  FileMap() = default;
  explicit FileMap(const std::string& path);
  FileMap(const FileMap&) = delete;
  FileMap& operator=(const FileMap&) = delete;
  FileMap(FileMap&& o) noexcept;
  FileMap& operator=(FileMap&& o) noexcept;
  ~FileMap();
  const unsigned char* data() const noexcept { return p_; }
  std::size_t size() const noexcept { return n_; }

 private:
  const unsigned char* p_ = nullptr;
  std::size_t n_ = 0;
};

struct BuildParams {
  std::size_t dim = 0;
  int metric = 0;
  std::size_t M = 16;
  std::size_t ef_construction = 200;
  std::size_t max_elements = 0;
  uint64_t seed = 100;
};

struct Graph;
class Team;

/// A node at a distance; ordered by (distance, id) everywhere.
struct Key {
  float d;
  uint32_t id;
};

/// The graph as it stood before an insertion wave: its entry point and top
/// level, and the ids whose lists the previous wave of the same add_batch
/// changed (on any layer; empty before the first wave).
struct WaveFrame {
  uint32_t entry;
  int top;
  const std::vector<uint32_t>* changed;
};

/// Phase 1 of an insertion wave [s, e): for every node i of the wave,
/// found[i - s][lc] = the ef_construction beam on layer lc (sorted by
/// (distance, id)) for lc = min(level, frame.top) .. 0, searched on the frozen
/// graph. A replacement (the device probe) must return exactly what the host's
/// probe returns.
using ProbeFn = std::function<void(const Graph& g, const WaveFrame& frame, uint32_t s, uint32_t e,
                                   std::vector<std::vector<std::vector<Key>>>& found, Team& team)>;

struct Graph {
  // parameters
  BuildParams params;
  std::size_t dim = 0, padded_dim = 0, M = 0, M0 = 0, efc = 0, ef_search = 10;
  std::size_t l0_stride = 0, upper_stride = 0;
  double mL = 0.0;

  // owned storage (empty under mmap)
  AlignedArray<float> vec_store;
  AlignedArray<uint32_t> l0_store;
  std::vector<uint32_t> upper_store;
  std::vector<uint64_t> label_store;
  std::vector<uint8_t> level_store;

  // the arrays every algorithm reads
  const float* vecs = nullptr;
  uint32_t* l0 = nullptr;
  uint32_t* upper = nullptr;
  const uint64_t* labels = nullptr;
  const uint8_t* levels = nullptr;

  std::vector<uint64_t> upper_offset;  // per node, u32 units
  std::vector<uint8_t> deleted;        // per node tombstone
  std::size_t num_deleted = 0;
  std::unordered_map<uint64_t, uint32_t> label_to_id;

  std::size_t count = 0, capacity = 0, upper_used = 0;
  uint32_t entry_point = kNoNode;
  int max_level = -1;
  uint64_t rng_state = 0;

  FileMap mapping;
  bool read_only = false;

  explicit Graph(const BuildParams& p);

  const float* vec(uint32_t id) const noexcept { return vecs + static_cast<std::size_t>(id) * padded_dim; }
  uint32_t* list(uint32_t id, int level) const noexcept {
    return level == 0 ? l0 + static_cast<std::size_t>(id) * l0_stride
                      : upper + upper_offset[id] + static_cast<std::size_t>(level - 1) * upper_stride;
  }
  std::size_t degree_cap(int level) const noexcept { return level == 0 ? M0 : M; }
  using DistFn = float (*)(const float*, const float*, std::size_t) noexcept;
  DistFn dist_fn = nullptr;  // the metric's distance for this host's ISA
  static DistFn pick_distance(int metric);
  float dist(const float* a, const float* b) const noexcept { return dist_fn(a, b, padded_dim); }

  void reserve(std::size_t n);
  void rebind();
  /// add_batch: validates every label first, then appends (levels drawn in
  /// input order) and links the new nodes in insertion waves on `threads`
  /// workers (0 = all cores). `probe`, if set, replaces the host's phase 1
  /// and must return exactly what the host's probe would.
  void add_batch(const float* vecs_in, const uint64_t* labels_in, std::size_t n, std::size_t threads,
                 const ProbeFn& probe = {});
  /// Store one row, label and level (the sequential part of an insertion).
  uint32_t append(const float* v, uint64_t label);

  void save(const std::string& path) const;
  static std::unique_ptr<Graph> load(const std::string& path, bool mmap);

  /// Bytes of vectors + graph, as the reference reports it.
  std::size_t memory_usage() const;
};

}  // namespace hnswamd
