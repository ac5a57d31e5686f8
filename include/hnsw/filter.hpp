#pragma once

// Result filters of the C++ API, as in the reference's include/hnsw/filter.hpp:12-48:
// a predicate over external labels; filtered-out nodes are still traversed but
// never returned. On the GPU a BitsetFilter is uploaded as its words; any other
// filter is evaluated on the host once per node and uploaded as a bitmap.

#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

namespace hnsw {

class Filter {
 public:
  virtual ~Filter() = default;
  virtual bool allowed(uint64_t label) const = 0;
  bool operator()(uint64_t label) const { return allowed(label); }
};

/// Wraps any callable. It must be safe to call from several threads.
class FunctionFilter final : public Filter {
 public:
  explicit FunctionFilter(std::function<bool(uint64_t)> fn);
  bool allowed(uint64_t label) const override { return fn_(label); }

 private:
  std::function<bool(uint64_t)> fn_;
};

/// Allow-list as a bitset indexed by label; labels >= size() are rejected.
class BitsetFilter final : public Filter {
 public:
  BitsetFilter() = default;
  explicit BitsetFilter(uint64_t num_labels);
  static BitsetFilter from_labels(const uint64_t* labels, std::size_t n);

  void set(uint64_t label, bool value = true);
  bool allowed(uint64_t label) const override {
    return label < size_ && ((words_[label >> 6] >> (label & 63)) & 1ULL) != 0;
  }
  uint64_t size() const noexcept { return size_; }
  std::size_t count() const noexcept;
  /// The bit words (bit l of word l / 64 = label l), for upload to the device.
  const uint64_t* words() const noexcept { return words_.data(); }

 private:
  std::vector<uint64_t> words_;
  uint64_t size_ = 0;
};

}  // namespace hnsw
