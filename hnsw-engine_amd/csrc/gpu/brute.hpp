#pragma once

// Exact k-NN on the GPU behind BruteForceIndex (reference src/brute_force.cpp).

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../../include/hnswgpu.h"

namespace hnswamd {
namespace gpu {

struct BruteSet;

/// Upload n prepared (cosine-normalised) rows of `dim` floats with their labels.
BruteSet* brute_create(std::size_t dim, int metric, const float* rows, const uint64_t* labels, std::size_t n);
void brute_free(BruteSet* s);
/// nq host queries -> nq x k host rows, best k by (distance, label), padded
/// with {~0, +inf}. allow_by_row (nullable) is a bitmap over rows. counts
/// (nullable) receives min(k, #allowed) per query.
void brute_search(BruteSet* s, const float* qs, std::size_t nq, std::size_t k, const uint64_t* allow_by_row,
                  hnswgpu_neighbor* out, uint32_t* counts);
/// Options: force_scan (0/1), kernel_timing (0/1), chunk_rows (rows per MFMA
/// chunk, 0 = automatic). Read-only: last_mfma, last_redo (queries re-run by
/// the scan), last_chunks, last_gemm_ns, last_total_ns (with kernel_timing).
void brute_set_option(BruteSet* s, const std::string& name, int64_t value);
int64_t brute_get_option(const BruteSet* s, const std::string& name);

}  // namespace gpu
}  // namespace hnswamd
