/*
 * hnswgpu.h -- the C ABI of the MI355X (gfx950) HNSW search engine.
 *
 * Plain pointers, sizes and opaque handles; no C++ or torch types cross it.
 * Every function returns HNSWGPU_OK (0) or a negative status whose text is in
 * hnswgpu_last_error() (thread-local). Status classes mirror the exception
 * types of the reference C++ API: HNSWGPU_EINVAL <-> std::invalid_argument,
 * HNSWGPU_ERUNTIME <-> std::runtime_error, HNSWGPU_ERANGE <-> std::out_of_range.
 *
 * Two layers:
 *   hnsw_index_*  the reference's hnsw::Index surface (include/hnsw/index.hpp of
 *                 the reference): build, load/save in the reference file format,
 *                 soft deletes, and search/search_batch on the GPU.
 *   hnswgpu_*     the device layer underneath: upload a loaded index to HBM once,
 *                 then search it with host or device buffers.
 * plus hnsw_brute_force_* for BruteForceIndex.
 */
#ifndef HNSWGPU_H_
#define HNSWGPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HNSWGPU_OK 0
#define HNSWGPU_EINVAL (-1)
#define HNSWGPU_ERUNTIME (-2)
#define HNSWGPU_EDEVICE (-3)
#define HNSWGPU_ERANGE (-4)

/* Metric ids as stored in the file header (docs/FORMAT.md): 0 L2, 1 inner product, 2 cosine. */
#define HNSWGPU_METRIC_L2 0
#define HNSWGPU_METRIC_IP 1
#define HNSWGPU_METRIC_COSINE 2

/* Same layout as the reference's hnsw::Neighbor {u64 label; f32 distance;}
 * (sizeof 16, include/hnsw/index.hpp:23-26). Missing results are
 * {~0 (kInvalidLabel), +inf}. */
typedef struct hnswgpu_neighbor {
  uint64_t label;
  float distance;
  uint32_t reserved;
} hnswgpu_neighbor;

/* Thread-local text of the last failure on this thread. */
const char* hnswgpu_last_error(void);

/* ------------------------------------------------------------------------ */
/* Device layer                                                              */
/* ------------------------------------------------------------------------ */

typedef struct hnswgpu_index hnswgpu_index;

/* A loaded index in the reference's in-memory layout (src/index_impl.hpp:176-237):
 * host pointers, borrowed for the duration of hnswgpu_upload only. */
typedef struct hnswgpu_index_view {
  uint64_t dim;
  uint64_t padded_dim;      /* dim rounded up to 16 */
  uint64_t M;               /* upper-layer degree; layer 0 has 2M */
  uint64_t count;
  uint64_t ef_search;       /* default ef when a search passes ef = 0 */
  uint64_t upper_used;      /* u32 slots in `upper` */
  int32_t metric;
  int32_t max_level;
  uint32_t entry_point;
  uint32_t reserved;
  const float* vectors;     /* count x padded_dim */
  const uint32_t* level0;   /* count x (1 + 2M) */
  const uint32_t* upper;    /* per node of level L >= 1: L blocks of (1 + M) */
  const uint64_t* labels;   /* count */
  const uint8_t* levels;    /* count */
  const uint8_t* deleted;   /* count tombstone bytes, or NULL for none */
} hnswgpu_index_view;

/* Replaces: the arrays Index::load leaves in memory (src/persist.cpp:477-541),
 * which the device layer uploads once to HBM and keeps resident. */
int hnswgpu_upload(const hnswgpu_index_view* view, hnswgpu_index** out);

/* Replaces: the device mirror's refresh after Index::add_batch
 * (src/index.cpp:273-306 grows the arrays the mirror copied). The arrays are
 * replaced in place: the handle, its options and its tuned launch shapes stay
 * valid, and a reduced traversal copy in use is rebuilt. EINVAL while a server
 * is attached to the handle (its slots are sized for the old index). */
int hnswgpu_reupload(hnswgpu_index* h, const hnswgpu_index_view* view);

/* Releases the device copy. EINVAL while a server is attached. */
int hnswgpu_free(hnswgpu_index* h);

/* Replaces: the tombstone write of Index::mark_deleted (src/index.cpp:308-318),
 * addressed by internal id. */
int hnswgpu_set_deleted(hnswgpu_index* h, uint32_t id, uint8_t deleted);

/* Replaces: Index::set_ef (src/index.cpp:352-356): the ef used when a search
 * passes ef = 0. */
int hnswgpu_set_ef(hnswgpu_index* h, uint64_t ef);

/* Replaces: Index::search_batch (include/hnsw/index.hpp:59-60,
 * src/search_layer.cpp:180-194). Host buffers, synchronous. qs is nq x dim
 * (unpadded), out is nq x k, padded with {~0, +inf}. The optional filter is a
 * BitsetFilter's words over labels (bit l set = label l allowed, labels >=
 * filter_size rejected; include/hnsw/filter.hpp:31-48). counts (nullable)
 * receives min(k, #results) per query, the length Index::search returns.
 * num_threads is accepted for signature parity and ignored. */
int hnswgpu_search_batch(hnswgpu_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                         const uint64_t* filter_words, uint64_t filter_size, hnswgpu_neighbor* out,
                         uint32_t* counts, uint64_t num_threads);

/* Concurrency: searches on one handle may run from several host threads at
 * once, as the reference's const search / search_batch may (shared lock,
 * include/hnsw/index.hpp:31-36). Each call in flight takes a context of its
 * own (scratch, staging, stream) from the handle's pool (option
 * "max_contexts", default 4; further callers wait for one), so their launches
 * overlap on the device. An asynchronous device call (hnswgpu_search_batch_device)
 * on a non-null stream takes the context last used on that stream, else one
 * whose work has finished, else a new one: calls on different streams overlap
 * (alternating two streams hides each launch's tail), calls on one stream
 * reuse one context. Uploads, deletions, options and tuning are exclusive. */

/* Replaces: Index::search_batch with a Filter that is not a BitsetFilter
 * (include/hnsw/filter.hpp:12-30: FunctionFilter or any subclass), which the
 * reference evaluates as eligible(id) = !deleted[id] && f->allowed(labels[id])
 * (src/index_impl.hpp:259-261). The caller evaluates its predicate once per
 * node and passes the result as id_words: ceil(count / 64) words over internal
 * ids (bit i set = node i may be returned; NULL = no filter). Soft deletions
 * still apply. Otherwise as hnswgpu_search_batch. */
int hnswgpu_search_batch_ids(hnswgpu_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                             const uint64_t* id_words, hnswgpu_neighbor* out, uint32_t* counts,
                             uint64_t num_threads);

/* Same search on device buffers, asynchronous on `stream` (a hipStream_t, NULL
 * = the default stream). Buffers must stay valid until the stream reaches the
 * end of the call's work. */
int hnswgpu_search_batch_device(hnswgpu_index* h, const float* d_qs, uint64_t nq, uint64_t k, uint64_t ef,
                                const uint64_t* d_filter_words, uint64_t filter_size,
                                hnswgpu_neighbor* d_out, uint32_t* d_counts, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hnswgpu_search_batch_device with the same arguments and
 * results, whose scratch is a fixed 96 MiB device array per GPU (allocated
 * with the library's module, held until the process ends) instead of the
 * handle's context pool. The call allocates nothing (no hipMalloc / hipFree,
 * no stream or event creation) and enqueues only kernels on `stream`, so it
 * can be captured into a graph. Limits: the per-call scratch (256 B + nq x 4
 * + nq x padded_dim x 4 + count / 8 bytes when filtered or with deletions,
 * plus 16 x count bytes for the general kernel) must fit in the 80 MiB after
 * the visited tables, else EINVAL with the sizes; the grid has at most 2048
 * wavefronts; filtered queries whose candidates outgrow their registers finish
 * in the general kernel (same rows, slower). Every call on one GPU shares the
 * array: calls that overlap on a GPU race, and the caller must order them. The
 * first call on each device resolves the array's address: make it outside
 * graph capture. */
int hnswgpu_search_batch_device_noalloc(hnswgpu_index* h, const float* d_qs, uint64_t nq, uint64_t k, uint64_t ef,
                                        const uint64_t* d_filter_words, uint64_t filter_size,
                                        hnswgpu_neighbor* d_out, uint32_t* d_counts, void* stream);

/* hnswgpu_search_batch_ids on device buffers (id_words on the device too). */
int hnswgpu_search_batch_ids_device(hnswgpu_index* h, const float* d_qs, uint64_t nq, uint64_t k, uint64_t ef,
                                    const uint64_t* d_id_words, hnswgpu_neighbor* d_out, uint32_t* d_counts,
                                    void* stream);

/* The cosine query scale the device computes for device-buffer searches:
 * float(1 / sqrt(sum of double(x[i])^2)) as the reference's prepare_vector
 * (src/space.hpp:31-40), 1 for a zero vector, reproduced with integer
 * arithmetic only (no FP64 on the GPU). This runs the same code on the host,
 * for checking; *out receives the float. Host-buffer searches compute the scale
 * on the host in double. No reference counterpart. */
int hnswgpu_query_scale(const float* x, uint64_t dim, float* out);

/* A mixed batch (no reference counterpart; it generalises search_batch, whose
 * queries share k, ef and filter): every task is one query with its own k, ef
 * and filter, and all of them share one launch through a device work queue.
 * Results are those of search_batch on the same query, k, ef and filter. */
typedef struct hnswgpu_task {
  const float* query;        /* device: dim floats */
  hnswgpu_neighbor* out;     /* device: k rows, padded with {~0, +inf} */
  uint32_t* count;           /* device: min(k, #results), or NULL */
  uint32_t k;
  uint32_t ef;               /* 0 = the index's ef_search; the search uses max(ef, k) */
  int32_t filter;            /* index into `filters`, or -1 for none */
  int32_t reserved;
} hnswgpu_task;

typedef struct hnswgpu_filter {
  const uint64_t* words;     /* device: BitsetFilter words over labels */
  uint64_t size;             /* labels covered (BitsetFilter::size) */
} hnswgpu_filter;

/* Tasks and filters are host arrays of device pointers; asynchronous on
 * `stream` like hnswgpu_search_batch_device. storage: -1 = the handle's
 * current "storage" option, else 0 (fp32), 1 (bf16) or 2 (fp8), which must
 * have been built by setting that option once. */
int hnswgpu_search_tasks_device(hnswgpu_index* h, const hnswgpu_task* tasks, uint64_t ntasks,
                                const hnswgpu_filter* filters, uint64_t nfilters, int32_t storage, void* stream);

/* ------------------------------------------------------------------------ */
/* Serving loop (no reference counterpart; the reference serves search_batch   */
/* calls from caller threads): continuous batching over a resident index.      */
/* ------------------------------------------------------------------------ */
typedef struct hnswgpu_server hnswgpu_server;

typedef struct hnswgpu_server_config {
  uint32_t slots;              /* launches in flight, each with its own stream and scratch (0 = 2) */
  uint32_t max_batch;          /* queries per launch (0 = 4096) */
  uint64_t max_rows;           /* result rows (sum of k) per launch (0 = 64 * max_batch) */
  uint32_t max_filters;        /* distinct filters per launch (0 = 8) */
  uint32_t use_graphs;         /* 1 = replay a hipGraph captured per slot and launch shape */
  uint64_t max_filter_labels;  /* labels a filter may cover (0 = every label of the index) */
  uint64_t hbm_budget;         /* device bytes all slots may use (0 = a quarter of the free memory);
                                  the slot count is reduced to fit */
  uint32_t spill_waves;        /* general-kernel wavefronts per slot (0 = 16) */
  uint32_t storages;           /* bit s set: build the reduced copy s (1 bf16, 2 fp8) now */
  uint64_t max_cost;           /* estimated cost per launch, in units of one unfiltered query at
                                  ef 1: a query costs max(ef, k) / (fraction of labels its filter
                                  allows, at least 1/50); 0 = no bound; a launch always takes at
                                  least one query */
  uint64_t heavy_cost;         /* requests whose queries cost at least this much (0 = 1024, e.g. a
                                  10 %-density filter at ef 128) form launches of their own, which
                                  hold at most all but one slot, so that cheap requests never wait
                                  for them */
} hnswgpu_server_config;

/* One request: search_batch's arguments on host buffers, which must stay valid
 * until hnswgpu_server_wait returns for it. storage: -1 = the index's current
 * "storage" option, else 0 / 1 / 2 (that copy must be built). */
typedef struct hnswgpu_request {
  const float* queries;        /* nq x dim */
  uint64_t nq;
  uint64_t k;
  uint64_t ef;                 /* 0 = ef_search; the search uses max(ef, k) */
  const uint64_t* filter_words;/* BitsetFilter words over labels, or NULL */
  uint64_t filter_size;
  int32_t storage;
  int32_t reserved;
  hnswgpu_neighbor* out;       /* nq x k */
  uint32_t* counts;            /* nq, or NULL */
} hnswgpu_request;

/* The index must stay uploaded and unchanged (no add, no reload) while a
 * server uses it: hnswgpu_reupload, hnswgpu_free and the facade's add_batch
 * refuse while one is attached. Soft deletions, options and tuning are
 * allowed; a slot whose captured graph no longer matches the launch shape or
 * scratch captures it again. */
int hnswgpu_server_create(hnswgpu_index* h, const hnswgpu_server_config* config, hnswgpu_server** out);
/* Queue a request (never blocks on the device); *ticket identifies it. */
int hnswgpu_server_submit(hnswgpu_server* s, const hnswgpu_request* request, uint64_t* ticket);
/* Block until the request has completed; returns its status. Each ticket is
 * waited for once. Results equal hnswgpu_search_batch on the same arguments. */
int hnswgpu_server_wait(hnswgpu_server* s, uint64_t ticket);
/* hnswgpu_server_wait that also returns the request's latency (ms, submit to
 * completion, as the statistics count it). */
int hnswgpu_server_wait_latency(hnswgpu_server* s, uint64_t ticket, double* latency_ms);
/* Statistics since the last reset, as doubles: [0] requests completed,
 * [1] queries completed, [2] launches, [3] mean queries per launch,
 * [4] p50 and [5] p99 request latency (ms, submit to completion), [6] max
 * latency, [7] queries/s from the first submit to the last completion,
 * [8] mean latency, [9] slots, [10] device bytes per slot, [11] graph replays,
 * [12] graphs captured again because their launch shape or scratch changed
 * (an option, the autotuner). */
int hnswgpu_server_stats(hnswgpu_server* s, double* out, uint64_t n);
int hnswgpu_server_reset_stats(hnswgpu_server* s);
/* Completes every submitted request, then stops the server. */
int hnswgpu_server_free(hnswgpu_server* s);

/* Autotuner (no reference counterpart): for each ef in efs (one per capacity
 * class 64 / 128 / 256 / 512 suffices), times the search of the nq sample
 * queries (host, k results) with HIP events for every launch shape in two
 * stages: wavefronts per CU {8, 12, 16, 20, 24} x LDS visited-hash words
 * 2^{10..13} x row loop (fixed-width for padded dim 128, or the generic one);
 * then, around the best of those, rows per memory round trip (16 or 32: the
 * gather width of the row loads, fixed-width loop only) and capacity class
 * (the ef's own, or the next larger: more register rows for the result and
 * candidate arrays) -- and keeps the fastest per class; searches then use it
 * ("use_tuned" = 1). One wavefront serves one query in every shape. Results
 * do not depend on the shape. The best shape depends on the batch size (a batch smaller than the
 * resident wavefronts is not bound by occupancy), so the sample should be as
 * large as the batches to be served. report (nullable) receives one line per
 * measurement. */
int hnswgpu_autotune(hnswgpu_index* h, const float* queries, uint64_t nq, uint64_t k, const uint64_t* efs,
                     uint64_t nefs, char* report, uint64_t report_len);

/* Tuning and test switches (no reference counterpart):
 *   "hash_log2"      visited-hash slots per wavefront (0 = automatic)
 *   "waves_per_cu"   resident wavefronts per CU of the persistent grid
 *   "visit_log2"     slots of each wavefront's HBM visited table
 *   "generic_gather" 1 = the generic row loop even for padded dim 128
 *   "row_trip"       16-row steps per memory round trip, fixed-width loop only
 *                    (0 = 2 for a bf16 / fp8 traversal copy, 1 for fp32 rows)
 *   "min_class"      smallest capacity class to run (0..3: ef <= 64/128/256/512)
 *   "wide_hash"      1 = 32-bit LDS visited entries (default: 16-bit when ids fit)
 *   "use_tuned"      1 = the autotuner's launch shapes (set by hnswgpu_autotune)
 *   "storage"        traversal copy, below
 *   "force_spill"    1 = every query through the general (HBM heap) kernel
 *   "deleted_path"   searches whose eligibility bitmap excludes nodes (soft
 *                    deletions, a filter or both): 0 = the filtered kernel,
 *                    1 = the unfiltered kernel with a side list for the
 *                    ineligible nodes it admits when at least 15 nodes in 16
 *                    are eligible by the host's estimate (deletions, and
 *                    filter words on the host), else the filtered kernel
 *                    (default), 2 = the side list always; a live candidate
 *                    pushed out of the 64-key side list hands its query to
 *                    the general kernel (same rows)
 *   "spill_waves"    wavefronts of the general kernel
 *   "max_contexts"   searches in flight on the handle at once (1..64, default 4)
 *   "collect_stats"  1 = record per-query work counts (hnswgpu_read_stats)
 *   "kernel_timing"  1 = bracket the fast kernel with events on its stream
 * and read-only values:
 *   "servers"        servers attached to the handle
 *   "contexts"       search contexts created so far
 *   "last_overflow"  queries the fast kernel handed to the general kernel
 *   "last_hbm_visited", "last_handoff_candidates", "last_handoff_visited"
 *   "last_kernel_ns" fast-kernel time (needs kernel_timing; waits for it)
 *   "tuned_waves_per_cu_<c>", "tuned_hash_log2_<c>", "tuned_generic_gather_<c>",
 *   "tuned_us_<c>"   the autotuner's choice for ef class c in {64, 128, 256, 512}
 *                    (-1 = not tuned). */
int hnswgpu_set_option(hnswgpu_index* h, const char* name, int64_t value);
int hnswgpu_get_option(hnswgpu_index* h, const char* name, int64_t* value);

/* Precision of the traversal (no reference counterpart): option "storage" =
 * 0 traverses the fp32 rows (exact ids); 1 keeps a bf16 copy of the rows and
 * 2 an OCP fp8 e4m3 copy with a per-vector fp32 scale (max|x| / 448), both
 * with |x^|^2 per row. Distances then accumulate in fp32 over the copy and the
 * final ef candidates are re-ranked with exact fp32 distances, so every
 * returned distance is exact; ids may differ from the fp32 traversal.
 * hnswgpu_get_row copies one stored row (storage 0: padded_dim fp32, 1:
 * padded_dim bf16, 2: padded_dim fp8 bytes) and, for 1 / 2, {scale, |x^|^2}. */
int hnswgpu_get_row(hnswgpu_index* h, uint32_t id, int storage, void* out, float* scale_norm);

/* Per-query work of the last search (collect_stats on): nq x 3 of
 * {distances computed, layer-0 lists read, upper-layer lists read}, the
 * counts from which algorithmic HBM bytes follow. No reference counterpart. */
int hnswgpu_read_stats(hnswgpu_index* h, uint32_t* out, uint64_t nq);

/* Device plumbing of this library's HIP runtime, for callers without a GPU
 * framework (no reference counterpart): device buffers (kind 0 = host to
 * device, 1 = device to host, 2 = device to device), non-blocking streams at
 * the default priority, events, and device 0's properties
 * (info[0..5] = CUs, global memory bytes, LDS bytes per CU, clock kHz,
 * memory clock kHz, memory bus width). */
int hnswgpu_malloc(uint64_t bytes, void** out);
int hnswgpu_mfree(void* p);
int hnswgpu_memcpy(void* dst, const void* src, uint64_t bytes, int kind);
/* A plain 16-byte-per-thread copy kernel (bytes % 16 == 0), synchronous: the
 * known-size reference for hardware byte counters. */
int hnswgpu_copy_kernel(void* dst, const void* src, uint64_t bytes);
int hnswgpu_stream_create(void** out);
int hnswgpu_stream_destroy(void* stream);
int hnswgpu_stream_synchronize(void* stream);
int hnswgpu_event_create(void** out);
int hnswgpu_event_destroy(void* event);
int hnswgpu_event_record(void* event, void* stream);
int hnswgpu_event_elapsed_ms(void* begin, void* end, float* ms);
int hnswgpu_device_info(char* name, int len, int64_t* info);

/* ------------------------------------------------------------------------ */
/* hnsw::Index surface                                                       */
/* ------------------------------------------------------------------------ */

typedef struct hnsw_index hnsw_index;

/* Replaces: Index(const Params&) (include/hnsw/index.hpp:14-21,39). */
int hnsw_index_create(uint64_t dim, int32_t metric, uint64_t M, uint64_t ef_construction, uint64_t max_elements,
                      uint64_t seed, hnsw_index** out);
/* Replaces: Index::load(path, mmap) (include/hnsw/index.hpp:86). */
int hnsw_index_load(const char* path, int mmap, hnsw_index** out);
void hnsw_index_free(hnsw_index* h);
/* Replaces: Index::save (include/hnsw/index.hpp:83). */
int hnsw_index_save(hnsw_index* h, const char* path);
/* Replaces: Index::add_batch / add (include/hnsw/index.hpp:47-50); CPU builder. */
int hnsw_index_add_batch(hnsw_index* h, const float* vecs, const uint64_t* labels, uint64_t n,
                         uint64_t num_threads);
/* What hnsw_index_add_batch_device did (hnsw::Index::DeviceBuildStats). */
typedef struct hnswgpu_build_stats {
  uint64_t waves;        /* insertion waves probed on the GPU */
  uint64_t device_nodes; /* nodes whose layer-0 insertion search ran on the GPU */
  uint64_t host_nodes;   /* nodes probed on the host (upper-level nodes, early small waves) */
  double device_ms;      /* GPU time of those searches */
  double refresh_ms;     /* host time validating and uploading the lists each wave changed */
} hnswgpu_build_stats;

/* No counterpart in the reference (its builder is CPU only): add_batch with the
 * insertion searches of level-0 nodes on the GPU. The graph is the one
 * hnsw_index_add_batch builds with the same num_threads: >= 2 inserts in waves
 * against the graph frozen at each wave's start (1/32 of the graph, at most
 * 16,384 nodes), 1 inserts one node at a time (the reference's order).
 * stats may be NULL. */
int hnsw_index_add_batch_device(hnsw_index* h, const float* vecs, const uint64_t* labels, uint64_t n,
                                uint64_t num_threads, hnswgpu_build_stats* stats);
/* Replaces: Index::search_batch, on the GPU (include/hnsw/index.hpp:59-60). */
int hnsw_index_search_batch(hnsw_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                            hnswgpu_neighbor* out, uint64_t num_threads, const uint64_t* filter_words,
                            uint64_t filter_size);
/* Replaces: Index::search_batch with a FunctionFilter (include/hnsw/filter.hpp:22-30):
 * allowed(label, ctx) != 0 admits a label. It is called once per node of the
 * index on the calling thread (the GPU takes the result as an id bitmap), so it
 * need not be thread-safe. allowed == NULL searches unfiltered. */
typedef int (*hnsw_label_predicate)(uint64_t label, void* ctx);
int hnsw_index_search_batch_pred(hnsw_index* h, const float* qs, uint64_t nq, uint64_t k, uint64_t ef,
                                 hnswgpu_neighbor* out, uint64_t num_threads, hnsw_label_predicate allowed,
                                 void* ctx);
/* Replaces: Index::search (include/hnsw/index.hpp:55-56): writes up to k rows
 * and their number to *n_out. */
int hnsw_index_search(hnsw_index* h, const float* q, uint64_t k, uint64_t ef, const uint64_t* filter_words,
                      uint64_t filter_size, hnswgpu_neighbor* out, uint64_t* n_out);
/* Replaces: Index::mark_deleted / is_deleted / contains (include/hnsw/index.hpp:64-66). */
int hnsw_index_mark_deleted(hnsw_index* h, uint64_t label);
int hnsw_index_is_deleted(hnsw_index* h, uint64_t label, int* out);
int hnsw_index_contains(hnsw_index* h, uint64_t label, int* out);
/* Replaces: Index::set_ef / ef (include/hnsw/index.hpp:73-74). */
int hnsw_index_set_ef(hnsw_index* h, uint64_t ef);
/* Replaces: Index::get_vector (include/hnsw/index.hpp:81). out has dim floats. */
int hnsw_index_get_vector(hnsw_index* h, uint64_t label, float* out);
/* info[0..9] = size, dim, metric, M, ef_construction, max_elements, ef, max_level, read_only,
 * memory_usage (include/hnsw/index.hpp:68-78). */
int hnsw_index_info(hnsw_index* h, int64_t* info);
/* The device layer handle of an index (uploaded on first use; refreshed after add). */
int hnsw_index_device(hnsw_index* h, hnswgpu_index** out);

/* ------------------------------------------------------------------------ */
/* BruteForceIndex (include/hnsw/index.hpp:101-118)                          */
/* ------------------------------------------------------------------------ */
typedef struct hnsw_brute_force hnsw_brute_force;
int hnsw_brute_force_create(uint64_t dim, int32_t metric, hnsw_brute_force** out);
void hnsw_brute_force_free(hnsw_brute_force* h);
int hnsw_brute_force_add(hnsw_brute_force* h, const float* vecs, const uint64_t* labels, uint64_t n);
int hnsw_brute_force_search_batch(hnsw_brute_force* h, const float* qs, uint64_t nq, uint64_t k,
                                  hnswgpu_neighbor* out, uint64_t num_threads, const uint64_t* filter_words,
                                  uint64_t filter_size);
/* GPU options of a BruteForceIndex; no counterpart in the reference.
 *   force_scan     1: always use the exact scan kernel (no MFMA prefilter)
 *   kernel_timing  1: time the device part of each search with HIP events
 *   chunk_rows     rows per MFMA row chunk (0: chosen per call)
 * read-only, about the last search:
 *   last_mfma      1 if the MFMA prefilter ran (k <= 32, padded dim <= 256)
 *   last_redo      queries it could not certify, re-run by the scan
 *   last_chunks    row chunks of the MFMA grid
 *   last_gemm_ns   MFMA kernel time; last_total_ns: device time of the search */
int hnsw_brute_force_set_option(hnsw_brute_force* h, const char* name, int64_t value);
int hnsw_brute_force_get_option(hnsw_brute_force* h, const char* name, int64_t* value);

#ifdef __cplusplus
}
#endif

#endif /* HNSWGPU_H_ */
