// Graph search launchers, the general search kernel and the small kernels
// (query preparation, eligibility bitmaps, reduced-precision copies). The
// device code of the search itself, and the semantics it follows, are in
// search_device.hpp; the fast kernel's instantiations are compiled by
// fast_variant.hip.

#include "search_device.hpp"

namespace hnswamd {
namespace gpu {

namespace {

// ---------------------------------------------------------------------------
// General kernel: binary heaps and a visited bitmap in HBM; lane 0 does the
// heap work while the wave computes distances.
// ---------------------------------------------------------------------------
__device__ void heap_push(uint64_t* h, uint32_t& n, uint64_t x, bool max_heap) {
  uint32_t i = n++;
  while (i > 0) {
    const uint32_t p = (i - 1) >> 1;
    const bool up = max_heap ? (h[p] < x) : (x < h[p]);
    if (!up) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = x;
}

__device__ uint64_t heap_pop(uint64_t* h, uint32_t& n, bool max_heap) {
  const uint64_t top = h[0];
  const uint64_t x = h[--n];
  uint32_t i = 0;
  for (;;) {
    const uint32_t a = 2 * i + 1;
    if (a >= n) break;
    uint32_t c = a;
    if (a + 1 < n) {
      const bool right = max_heap ? (h[a] < h[a + 1]) : (h[a + 1] < h[a]);
      if (right) c = a + 1;
    }
    const bool down = max_heap ? (x < h[c]) : (h[c] < x);
    if (!down) break;
    h[i] = h[c];
    i = c;
  }
  if (n > 0) h[i] = x;
  return top;
}

__global__ void __launch_bounds__(kWavesPerBlock * 64) general_search(DevIndexView ix, SearchRequest r,
                                                                      SearchScratch s) {
  extern __shared__ __align__(16) uint32_t lds[];
  const int lane = lane_id();
  const uint32_t wave_slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave_slot >= s.spill_waves) return;
  const uint32_t per_wave = lds_words_per_wave(0, ix.padded) - 1u;  // no hash
  uint32_t* base = lds + static_cast<size_t>(threadIdx.x >> 6) * per_wave;
  const WaveLds w{nullptr,
                  base,
                  base + 64,
                  base + 128,
                  reinterpret_cast<float*>(base + 192),
                  reinterpret_cast<uint16_t*>(base + 192 + ix.padded),
                  reinterpret_cast<float*>(base + 192 + ix.padded + ix.padded / 2)};
  uint32_t* bits = s.spill_words + static_cast<size_t>(wave_slot) * s.spill_bitmap_words;
  uint64_t* Cheap = s.spill_heaps + static_cast<size_t>(wave_slot) * s.spill_heap_stride;
  uint64_t* Rheap = Cheap + (ix.count + 1u);
  const uint32_t n_over = uniform(*reinterpret_cast<volatile uint32_t*>(&s.counters[1]));

  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(&s.counters[2], 1u);
    t = uniform(t);
    if (t >= n_over) break;
    const uint32_t qi = s.overflow[t];
    const QueryArgs a = query_args(r, qi);
    const bool filtered = a.elig != nullptr;

    for (uint32_t i = static_cast<uint32_t>(lane); i < s.spill_bitmap_words; i += 64) bits[i] = 0u;
    __threadfence();
    Work work;
    const uint64_t ep = start_query<0>(ix, s, w, qi, work);
    uint32_t nC = 0, nR = 0;
    if (lane == 0) {
      atomicOr(&bits[key_id(ep) >> 5], 1u << (key_id(ep) & 31));
      heap_push(Cheap, nC, ep, false);
      if (!filtered || elig_bit(a.elig, key_id(ep))) heap_push(Rheap, nR, ep, true);
    }
    for (;;) {
      uint64_t c = kKeyInf;
      if (lane == 0) {
        if (nC > 0) {
          c = Cheap[0];
          if (nR >= a.ef && key_ord(c) > key_ord(Rheap[0])) {
            c = kKeyInf;
          } else {
            heap_pop(Cheap, nC, false);
          }
        }
      }
      c = uniform64(c);
      if (c == kKeyInf) break;
      const uint32_t U = fresh_neighbours(ix, w, key_id(c), a.elig, [&](uint32_t id) {
        const uint32_t old = atomicOr(&bits[id >> 5], 1u << (id & 31));
        return (old & (1u << (id & 31))) == 0u;
      });
      work.dist += U;
      ++work.exp0;
      row_dists<0>(ix, w, U);
      wave_sync();
      if (lane == 0) {
        for (uint32_t u = 0; u < U; ++u) {
          const uint32_t od = w.nd[u];
          if (nR < a.ef || od < key_ord(Rheap[0])) {
            const uint64_t key = key_of(od, w.nid[u]);
            heap_push(Cheap, nC, key, false);
            if (!filtered || w.nel[u] != 0u) {
              heap_push(Rheap, nR, key, true);
              if (nR > a.ef) heap_pop(Rheap, nR, true);
            }
          }
        }
      }
      wave_sync();
    }
    // Sorted results: pop the max-heap from the back into the C area.
    if (lane == 0) {
      const uint32_t total = nR;
      for (uint32_t i = total; i-- > 0;) Cheap[i] = heap_pop(Rheap, nR, true);
      nR = total;
    }
    __threadfence();
    nR = uniform(nR);
    const uint32_t kout = nR < a.k ? nR : a.k;
    OutRow* row = a.out;
    for (uint32_t j = static_cast<uint32_t>(lane); j < kout; j += 64) {
      const uint64_t key = Cheap[j];
      row[j] = OutRow{ix.labels[key_id(key)], float_of(key_ord(key)), 0u};
    }
    write_pad(row, kout, a.k);
    if (a.count != nullptr && lane == 0) *a.count = kout;
    if (a.stats != nullptr && lane == 0) {
      a.stats[0] = work.dist;
      a.stats[1] = work.exp0;
      a.stats[2] = work.expu;
    }
    __threadfence();
  }
}

// ---------------------------------------------------------------------------
// Small kernels.
// ---------------------------------------------------------------------------
__global__ void reset_counters(uint32_t* c) {
  if (threadIdx.x < 8) c[threadIdx.x] = 0u;
}

/// One wavefront per query: copy into a zero-padded row, normalising for
/// cosine (src/space.hpp:31-40). The scale 1/|q| comes from the host when it
/// has the queries (computed there in double, as the reference does); for
/// device queries lane 0 reproduces that double arithmetic bit for bit in
/// integers (exact_scale.hpp), so no step uses FP64.
__global__ void __launch_bounds__(256) prepare_queries(DevIndexView ix, SearchRequest r, SearchScratch s) {
  const int lane = lane_id();
  const uint32_t qi = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (qi >= launch_nq(r)) return;
  const QueryArgs a = query_args(r, qi);
  const float* src = a.query;
  float* dst = s.qprep + static_cast<size_t>(qi) * ix.padded;
  float inv = 1.0f;
  if (ix.metric == 2) {
    if (a.has_scale) {
      inv = a.scale;
    } else {
      uint32_t bits = 0;
      if (lane == 0) bits = exact::inv_norm_bits(src, r.dim);
      inv = __uint_as_float(rdlane(bits, 0));
    }
  }
  for (uint32_t i = static_cast<uint32_t>(lane); i < ix.padded; i += 64) {
    dst[i] = i < r.dim ? (ix.metric == 2 ? src[i] * inv : src[i]) : 0.0f;
  }
}

__global__ void build_eligibility(uint32_t count, const uint64_t* labels, const uint32_t* deleted_bits,
                                  const uint64_t* fw, uint64_t fsize, int by_id, uint32_t* elig) {
  const uint32_t word = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t words = (count + 31u) >> 5;
  if (word >= words) return;
  uint32_t out = 0u;
  const uint32_t del = deleted_bits != nullptr ? deleted_bits[word] : 0u;
  for (uint32_t b = 0; b < 32; ++b) {
    const uint32_t id = word * 32u + b;
    if (id >= count) break;
    bool ok = ((del >> b) & 1u) == 0u;
    if (ok && fw != nullptr && by_id) {
      ok = ((fw[id >> 6] >> (id & 63)) & 1ULL) != 0ULL;
    } else if (ok && fw != nullptr) {
      const uint64_t lab = labels[id];
      ok = lab < fsize && ((fw[lab >> 6] >> (lab & 63)) & 1ULL) != 0ULL;
    }
    if (ok) out |= 1u << b;
  }
  elig[word] = out;
}

__global__ void build_eligibility_multi(uint32_t count, const uint64_t* labels, const uint32_t* deleted_bits,
                                        const FilterDesc* descs, uint32_t* elig) {
  const FilterDesc d = descs[blockIdx.y];
  if (d.active == 0u) return;
  const uint32_t words = (count + 31u) >> 5;
  const uint32_t word = blockIdx.x * blockDim.x + threadIdx.x;
  if (word >= words) return;
  uint32_t out = 0u;
  const uint32_t del = deleted_bits != nullptr ? deleted_bits[word] : 0u;
  for (uint32_t b = 0; b < 32; ++b) {
    const uint32_t id = word * 32u + b;
    if (id >= count) break;
    bool ok = ((del >> b) & 1u) == 0u;
    if (ok && d.words != nullptr) {
      const uint64_t lab = labels[id];
      ok = lab < d.size && ((d.words[lab >> 6] >> (lab & 63)) & 1ULL) != 0ULL;
    }
    if (ok) out |= 1u << b;
  }
  elig[static_cast<size_t>(blockIdx.y) * words + word] = out;
}


template <int PC>
hipError_t pick_variant(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st) {
  if (ix.storage == kStoreBF16) return launch_fast<kStoreBF16, PC>(ix, r, s, st);
  if (ix.storage == kStoreFP8) return launch_fast<kStoreFP8, PC>(ix, r, s, st);
  return launch_fast<kStoreF32, PC>(ix, r, s, st);
}

/// One wavefront per row: the reduced copy of an fp32 row and |x^|^2 of it.
__global__ void __launch_bounds__(256) build_storage(DevIndexView ix, int storage, void* vlow, float* vscale,
                                                     float* vnorm) {
  const int lane = lane_id();
  const uint32_t id = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (id >= ix.count) return;
  const float* x = ix.vecs + static_cast<size_t>(id) * ix.padded;
  float sq = 0.0f;
  if (storage == kStoreBF16) {
    uint16_t* out = static_cast<uint16_t*>(vlow) + static_cast<size_t>(id) * ix.padded;
    for (uint32_t i = static_cast<uint32_t>(lane); i < ix.padded; i += 64) {
      const uint32_t b = bf16_bits_rne(x[i]);
      out[i] = static_cast<uint16_t>(b);
      const float v = bf16_value(b);
      sq = __builtin_fmaf(v, v, sq);
    }
  } else {
    float m = 0.0f;
    for (uint32_t i = static_cast<uint32_t>(lane); i < ix.padded; i += 64) m = fmaxf(m, fabsf(x[i]));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    const float scale = m > 0.0f ? m / 448.0f : 1.0f;
    const float inv = m > 0.0f ? 448.0f / m : 1.0f;
    uint8_t* out = static_cast<uint8_t*>(vlow) + static_cast<size_t>(id) * ix.padded;
    // Lane pairs of elements: lane l encodes elements 2l, 2l+1 of each 128-wide slice.
    for (uint32_t i = 2u * static_cast<uint32_t>(lane); i < ix.padded; i += 128) {
      const float a = fminf(fmaxf(x[i] * inv, -448.0f), 448.0f);
      const float b = fminf(fmaxf(x[i + 1] * inv, -448.0f), 448.0f);
      const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
      out[i] = static_cast<uint8_t>(packed & 0xFF);
      out[i + 1] = static_cast<uint8_t>((packed >> 8) & 0xFF);
      const f32x2_t back = __builtin_amdgcn_cvt_pk_f32_fp8(packed, false);
      const float va = back.x * scale, vb = back.y * scale;
      sq = __builtin_fmaf(va, va, sq);
      sq = __builtin_fmaf(vb, vb, sq);
    }
    if (lane == 0) vscale[id] = scale;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o);
  if (lane == 0) vnorm[id] = sq;
}

}  // namespace

hipError_t launch_build_storage(const DevIndexView& ix, int storage, void* vlow, float* vscale, float* vnorm,
                                hipStream_t st) {
  if (ix.count == 0) return hipSuccess;
  hipLaunchKernelGGL(build_storage, dim3((ix.count + 3) / 4), dim3(256), 0, st, ix, storage, vlow, vscale, vnorm);
  return hipGetLastError();
}

hipError_t launch_prepare(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st) {
  if (r.nq == 0) return hipSuccess;  // with nq_dev, r.nq is the capacity the grid covers
  hipLaunchKernelGGL(prepare_queries, dim3((r.nq + 3) / 4), dim3(256), 0, st, ix, r, s);
  return hipGetLastError();
}

hipError_t launch_eligibility(const DevIndexView& ix, const uint32_t* deleted_bits, const uint64_t* fw,
                              uint64_t fsize, int by_id, uint32_t* elig, hipStream_t st) {
  const uint32_t words = (ix.count + 31u) >> 5;
  if (words == 0) return hipSuccess;
  hipLaunchKernelGGL(build_eligibility, dim3((words + 255) / 256), dim3(256), 0, st, ix.count, ix.labels,
                     deleted_bits, fw, fsize, by_id, elig);
  return hipGetLastError();
}

hipError_t launch_eligibility_multi(const DevIndexView& ix, const uint32_t* deleted_bits, const FilterDesc* descs,
                                    uint32_t nf, uint32_t* elig, hipStream_t st) {
  const uint32_t words = (ix.count + 31u) >> 5;
  if (words == 0 || nf == 0) return hipSuccess;
  hipLaunchKernelGGL(build_eligibility_multi, dim3((words + 255) / 256, nf), dim3(256), 0, st, ix.count, ix.labels,
                     deleted_bits, descs, elig);
  return hipGetLastError();
}

hipError_t launch_search(const DevIndexView& ix, const SearchRequest& r, SearchScratch& s, hipStream_t st,
                         hipEvent_t fast_begin, hipEvent_t fast_end) {
  hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, st, s.counters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // A mixed launch sizes the kernel for min(largest ef, kFastMaxEf); larger
  // tasks are handed to the general kernel one by one.
  const bool fast = !r.force_spill && (r.tasks != nullptr || r.ef <= kFastMaxEf);
  SearchRequest rr = r;
  rr.force_spill = fast ? 0 : 1;
  if (rr.tasks != nullptr && rr.ef > kFastMaxEf) rr.ef = kFastMaxEf;
  if (fast_begin != nullptr && fast_end != nullptr) {
    e = hipEventRecord(fast_begin, st);
    if (e != hipSuccess) return e;
  }
  if (ix.padded == 128 && !r.generic_gather) {
    e = pick_variant<8>(ix, rr, s, st);
  } else {
    e = pick_variant<0>(ix, rr, s, st);
  }
  if (e != hipSuccess) return e;
  if (fast_begin != nullptr && fast_end != nullptr) {
    e = hipEventRecord(fast_end, st);
    if (e != hipSuccess) return e;
  }
  const size_t lds = static_cast<size_t>(kWavesPerBlock) * (lds_words_per_wave(0, ix.padded) - 1u) * 4u;
  const uint32_t blocks = (s.spill_waves + kWavesPerBlock - 1) / kWavesPerBlock;
  if (blocks == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(general_search, dim3(blocks), dim3(kWavesPerBlock * 64), lds, st, ix, rr, s);
  return hipGetLastError();
}

}  // namespace gpu
}  // namespace hnswamd
