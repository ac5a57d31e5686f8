#pragma once

// Error plumbing of the C ABI: C++ exceptions inside, int status + thread-local
// text at the boundary (include/hnswgpu.h).

#include <stdexcept>
#include <string>

namespace hnswamd {

/// A HIP runtime failure.
class DeviceError : public std::runtime_error {
 public:
  explicit DeviceError(const std::string& what) : std::runtime_error(what) {}
};

/// Record the in-flight exception as the thread's last error and map its type
/// to a status code. Call only from a catch block.
int status_from_current_exception() noexcept;

/// Rethrow a status + message recorded by a C-ABI call as the matching
/// exception type (used by the C++ facade over the device layer).
[[noreturn]] void throw_status(int status, const std::string& message);

void set_last_error(const std::string& message) noexcept;
const std::string& last_error() noexcept;

}  // namespace hnswamd
