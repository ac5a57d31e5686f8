// One slice of the fast search kernel's instantiations: the storage
// HNSWGPU_ST (0 = fp32 rows, 1 = bf16, 2 = fp8 e4m3 + scale) and the row width
// HNSWGPU_PC (8 = padded dim 128 with the fixed-width row loop, 0 = any width),
// with every filter mode and capacity class. The Makefile compiles this file
// once per (storage, row width), so the slices build in parallel.

#include "search_device.hpp"

#if !defined(HNSWGPU_ST) || !defined(HNSWGPU_PC)
#error "compile with -DHNSWGPU_ST=<0|1|2> -DHNSWGPU_PC=<0|8>"
#endif

namespace hnswamd {
namespace gpu {

template <int ST, int PC>
hipError_t launch_fast(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st) {
  return pick_filter<PC, ST>(ix, r, s, st);
}

template hipError_t launch_fast<HNSWGPU_ST, HNSWGPU_PC>(const DevIndexView& ix, const SearchRequest& r,
                                                        SearchScratch& s, hipStream_t st);

}  // namespace gpu
}  // namespace hnswamd
