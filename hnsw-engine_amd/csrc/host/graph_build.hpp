#pragma once

// Pieces of the wave builder (graph_build.cpp) that a device-side probe needs:
// the worker team, the host probe and the wave schedule (Key, WaveFrame and
// ProbeFn are declared with the graph).

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "graph.hpp"

namespace hnswamd {

/// A fixed team of threads running index-space jobs; the caller is worker 0.
class Team {
 public:
  explicit Team(std::size_t threads);
  Team(const Team&) = delete;
  Team& operator=(const Team&) = delete;
  ~Team();
  std::size_t size() const noexcept { return size_; }
  /// fn(i, worker) for every i in [0, n); returns when all have ended and
  /// rethrows the first exception.
  void run(std::size_t n, const std::function<void(std::size_t, std::size_t)>& fn);

 private:
  void serve(std::size_t worker);
  void drain(std::size_t worker);

  std::size_t size_;
  std::vector<std::thread> crew_;
  std::mutex mu_;
  std::condition_variable wake_, done_;
  const std::function<void(std::size_t, std::size_t)>* job_ = nullptr;
  std::size_t items_ = 0, busy_ = 0;
  uint64_t round_ = 0;
  bool closing_ = false;
  std::atomic<std::size_t> next_{0};
  std::atomic<bool> failed_{false};
  std::exception_ptr error_;
};

/// Phase 1 for one node on the host.
void probe_on_host(const Graph& g, const WaveFrame& frame, uint32_t id, std::vector<std::vector<Key>>& by_layer);

/// Largest insertion wave.
inline constexpr std::size_t kMaxWave = 16384;

/// Nodes per wave when `frozen` nodes are linked: 1 for one-thread builds
/// (the reference's sequential order), else 1/32 of the graph, 1..kMaxWave.
std::size_t wave_length(std::size_t frozen, std::size_t threads);

/// Insert ids [s, e), already appended, as one wave. `changed` holds the ids
/// whose lists the previous wave changed and is replaced by this wave's.
void insert_wave(Graph& g, uint32_t s, uint32_t e, Team& team, const ProbeFn& probe,
                 std::vector<uint32_t>& changed);

}  // namespace hnswamd
