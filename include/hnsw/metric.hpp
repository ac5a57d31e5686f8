#pragma once

// Distance metrics of the C++ API, as in the reference's include/hnsw/metric.hpp:9-30.
// (The reference's simd:: kernel table has no counterpart: the GPU always
// computes in one fixed order, that of the reference's AVX-512 kernels.)

#include <cstddef>
#include <string>
#include <string_view>

namespace hnsw {

enum class Metric { L2, InnerProduct, Cosine };

/// "l2" / "L2" / "euclidean", "ip" / "inner_product" / "dot", "cosine" / "angular".
/// Throws std::invalid_argument otherwise.
Metric parse_metric(std::string_view name);
std::string_view metric_name(Metric m) noexcept;

/// Squared L2 and dot product in the engine's summation order (host side).
float l2_sq(const float* a, const float* b, std::size_t dim) noexcept;
float dot(const float* a, const float* b, std::size_t dim) noexcept;

}  // namespace hnsw
