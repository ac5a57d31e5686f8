#pragma once

// Internal interface between the C-ABI layer and the HIP kernels.

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hnswamd {
namespace gpu {

/// The resident index in HBM (reference layout; rows padded to 16 floats).
struct DevIndexView {
  const float* vecs;         // count x padded fp32
  const uint32_t* l0;        // count x (1 + M0)
  const uint32_t* upper;     // upper_used
  const uint32_t* upper_off; // count, in u32 units
  const uint64_t* labels;    // count
  uint32_t count;
  uint32_t padded;
  uint32_t M;
  uint32_t M0;
  uint32_t entry;
  int32_t max_level;
  int32_t metric;            // 0 = L2, 1 = inner product, 2 = cosine
  uint32_t has_dups;         // some adjacency list repeats an id
  int32_t storage;           // traversal copy: 0 = the fp32 rows, 1 = bf16, 2 = fp8 e4m3 + scale
  const void* vlow;          // count x padded bf16 or fp8 (storage != 0)
  const float* vscale;       // count fp8 scales (storage == 2)
  const float* vnorm;        // count |x^|^2 of the stored rows (storage != 0)
};

/// Per-wavefront LDS words of the search kernels for an LDS hash of 2^log2h slots.
__host__ __device__ inline uint32_t lds_words_per_wave(uint32_t log2h, uint32_t padded) {
  return (1u << log2h) + 192u + padded + padded / 2u + 4u;
}

struct OutRow {
  uint64_t label;
  float distance;
  uint32_t pad;
};

/// Scratch the kernels share: [0] query ticket, [1] queries queued for the
/// general kernel, [2] general-kernel ticket, [3] queries whose visited set
/// went past the LDS hash into HBM, [4] queries queued because C dropped a
/// live entry, [5] queries queued because the HBM visited table filled.
struct SearchScratch {
  uint32_t* counters;        // 8 words
  uint32_t* visit_tables;    // one HBM visited table per resident wavefront
  uint32_t visit_log2;       // slots per table = 1 << visit_log2
  uint32_t* overflow;        // capacity >= nq
  float* qprep;              // nq x padded
  uint32_t* spill_words;     // spill_waves x visited-bitmap words
  uint64_t* spill_heaps;     // spill_waves x (count + 1 + ef + 1) keys
  uint64_t spill_heap_stride;
  uint32_t spill_waves;
  uint32_t spill_bitmap_words;
  uint64_t* cand_runs;       // per resident wavefront: kRunSlots x 64 keys (candidate overflow), or null
};

/// Sorted runs of 64 candidates a wavefront may spill to HBM (filtered search).
constexpr uint32_t kRunSlots = 128;

/// One query of a mixed launch (device work queue entry): its own k, ef,
/// eligibility bitmap and output rows, all device pointers.
struct QueryTask {
  const float* query;        // dim floats
  OutRow* out;               // k rows
  uint32_t* count;           // result count, or null
  const uint32_t* elig;      // id bitmap (deletions and filter combined), or null = unfiltered
  uint32_t k;
  uint32_t ef;               // clamped: max(ef or ef_search, k)
  float scale;               // cosine: the query's 1/|q|, computed by the host (has_scale = 1)
  uint32_t has_scale;        // 0: the device computes the scale from the query
};

struct SearchRequest {
  const QueryTask* tasks;    // when set: per-query parameters; k / ef / elig / out / counts / queries
                             // below are unused and ef holds the largest task ef
  const uint32_t* nq_dev;    // when set: the number of queries is read here (nq is then a capacity)
  const float* queries;      // nq x dim (device)
  const float* qscale;       // cosine: nq scales 1/|q| computed by the host, or null (computed on the device)
  uint32_t dim;
  uint32_t nq;
  uint32_t k;
  uint32_t ef;               // already clamped: max(ef or ef_search, k)
  const uint32_t* elig;      // id bitmap, 1 = may be returned; null = unfiltered
  int deleted_only;          // 1: every elig (tasks' included) excludes only a few deleted nodes, and is
                             // set: the unfiltered kernel keeps the deleted nodes it admits in a side list
  OutRow* out;               // nq x k (device)
  uint32_t* counts;          // nq result counts (device) or null
  uint32_t* stats;           // nq x 3 work counts {distances, layer-0 lists, upper lists} or null
  uint32_t hash_log2;        // visited-hash slots per wave = 1 << hash_log2
  uint32_t waves_per_block;
  uint32_t blocks;
  int force_spill;           // route every query through the general kernel
  int generic_gather;        // 1: the row loop for any padded dim even where a fixed-width one exists
  int wide_hash;             // 1: 32-bit LDS visited entries even where 16-bit ones hold the ids
  int row_trip;              // 16-row steps per memory round trip in the layer-0 expansion (1 or 2)
  uint32_t min_class;        // smallest capacity class to run (0..3: ef <= 64 / 128 / 256 / 512)
};

/// Build the reduced copy of the rows for `storage` (1 = bf16, 2 = fp8) into
/// vlow / vscale / vnorm (device buffers sized by the caller).
hipError_t launch_build_storage(const DevIndexView& ix, int storage, void* vlow, float* vscale, float* vnorm,
                                hipStream_t st);
/// Prepare queries (copy into padded rows; unit-normalise for cosine).
hipError_t launch_prepare(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st);
/// Reset counters, run the fast kernel, then the general kernel on overflow.
/// When both events are given they bracket the fast kernel on `st`.
hipError_t launch_search(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st,
                         hipEvent_t fast_begin = nullptr, hipEvent_t fast_end = nullptr);
/// Eligibility bitmap: 1 iff not deleted and (no filter, or the filter allows
/// the node: filter bits indexed by label, or by internal id when by_id).
hipError_t launch_eligibility(const DevIndexView& ix, const uint32_t* deleted_bits, const uint64_t* filter_words,
                              uint64_t filter_size, int by_id, uint32_t* elig, hipStream_t st);

/// One eligibility bitmap of a launch with several filters: filter words by
/// label (null = deletions only); inactive entries are skipped.
struct FilterDesc {
  const uint64_t* words;
  uint64_t size;
  uint32_t active;
  uint32_t pad;
};
/// Bitmap f of `elig` (words_per = ceil(count / 32) words each) from descs[f],
/// f < nf; the grid covers nf, so a launch of fixed shape serves any number of
/// active filters (graph replays).
hipError_t launch_eligibility_multi(const DevIndexView& ix, const uint32_t* deleted_bits, const FilterDesc* descs,
                                    uint32_t nf, uint32_t* elig, hipStream_t st);

/// Largest ef the fast kernel holds in registers.
constexpr uint32_t kFastMaxEf = 512;
/// Largest M0 (layer-0 degree) the kernels load in one wave.
constexpr uint32_t kMaxDegree0 = 64;

}  // namespace gpu
}  // namespace hnswamd
