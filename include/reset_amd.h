/*
 * reset_amd.h -- C ABI of the MI355X-native RESET hash set-intersection path.
 *
 * This is the drop-in boundary.  The reference exposes this path as C++
 * functions, classes and CUDA kernels (no FFI layer); each entry point below
 * names the reference interface it replaces.  The C++ mirror of those classes
 * (HashTableBuilder, TCBaseline, SSSBaseline, IRBaseline, make_hash_table) is
 * in reset_amd.hpp, built on these functions.
 *
 * Conventions
 *  - Plain pointers and explicit int64 sizes; no torch or HIP types.
 *  - Pointers named d_* are device (HBM) pointers on the current HIP device;
 *    h_* are host pointers.  Outputs are caller-owned unless stated.
 *  - rs_stream is a hipStream_t (NULL = the legacy default stream).  *_async
 *    calls only enqueue work; other calls synchronize `stream` before return.
 *  - Elements are int32 bit patterns (uint32 ids < 2^31 in practice); the
 *    value -1 (0xFFFFFFFF) is the EMPTY slot marker and may not be an element,
 *    as in the reference (TC.cu:21).
 *  - Every function returns an rs_status; rs_last_error() gives the message of
 *    the calling thread's last failure.  Nothing exits the process.
 *  - Handles are not thread-safe: one host thread per handle at a time.
 */
#ifndef RESET_AMD_H
#define RESET_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rs_status;
#define RS_OK 0
#define RS_EINVAL 1     /* bad argument or malformed input                */
#define RS_ENOMEM 2     /* device allocation failed / over memory budget  */
#define RS_EOVERFLOW 3  /* bounded probe or insert hit a full table       */
#define RS_EDEVICE 4    /* HIP runtime error                              */
#define RS_EBUSY 5      /* serving queue full; retry after polling        */
#define RS_ECANCELED 6  /* serving: the request was cancelled while queued */
#define RS_ETIMEDOUT 7  /* serving: the request outwaited the queue timeout */

typedef void* rs_stream;

/* Probe/build layouts (TC.cuh:40-52). */
#define RS_NORMAL 0        /* bucket = x & (L-1)                          */
#define RS_HIERARCHICAL 1  /* bucket = (x & (M-1)) >> (log2 M - log2 L)   */
/* Layout masks for rs_tables_build. */
#define RS_BUILD_NORMAL 1
#define RS_BUILD_HIERARCHICAL 2
#define RS_BUILD_BOTH 3

const char* rs_version(void);
const char* rs_status_string(rs_status s);
const char* rs_last_error(void);

/* ------------------------------------------------------------------ a1, a2 */
/* triangle-count/hashtable.cu:20-36 calculate_length (host evaluation). */
int32_t rs_calculate_length(int32_t degree, float load_factor, int32_t bucket_size);

/* triangle-count/hashtable.cu:39-51 calculate_hash_length kernel:
 * d_lengths[i] = calculate_length(off[i+1]-off[i]); *h_max_length = max. */
rs_status rs_calculate_lengths(const int32_t* d_offsets, int64_t num_sets, float load_factor,
                               int32_t bucket_size, int32_t* d_lengths, int32_t* h_max_length,
                               rs_stream stream);

/* ------------------------------------------------------------- a3-a6, a13 */
typedef struct rs_tables rs_tables;

typedef struct rs_tables_desc {
  int64_t num_sets;             /* n                                        */
  int64_t num_elements;         /* E = offsets[n]                           */
  int64_t bucket_num;           /* BN = sum of L over sets (int64)          */
  int32_t max_length;           /* M = max L                                */
  int32_t bucket_size;          /* b (slots per bucket)                     */
  float load_factor;            /* alpha                                    */
  int32_t layouts;              /* RS_BUILD_* mask of the built tables      */
  int32_t* d_table_normal;      /* int32[BN*b], slot k of bucket j of table */
  int32_t* d_table_hierarchical;/*   t at [toff[t] + j + k*BN]; NULL if not built */
  int32_t* d_lengths;           /* int32[n]   (d_hash_length)               */
  int64_t* d_table_offsets;     /* int64[n+1] (d_hash_tables_offset)        */
  const int32_t* d_offsets;     /* borrowed CSR offsets int32[n+1]          */
  const int32_t* d_elements;    /* borrowed CSR elements int32[E]           */
  int64_t conflicts_normal;     /* failed slot attempts while inserting     */
  int64_t conflicts_hierarchical;
} rs_tables_desc;

/* Replaces make_hash_table(float, int) (triangle-count/hashtable.cu:103-184)
 * and HashTableBuilder::build_hash_tables (baseline/common/hash_table.cuh:39-42,
 * hash_table.cu:69-192): per-set lengths, int64 offset scan, -1 fill and the
 * atomic-CAS insert of every (set, element) in the chosen layouts.  The CSR
 * arrays are borrowed and must outlive the handle.  Synchronizes `stream`.
 * RS_EOVERFLOW if some table had no free slot (load_factor >= 1). */
rs_status rs_tables_build(const int32_t* d_offsets, int64_t num_sets, const int32_t* d_elements,
                          float load_factor, int32_t bucket_size, int32_t layouts,
                          rs_stream stream, rs_tables** out);
/* The build's fill and insert phases alone, over the handle's existing
 * tables: every slot of the `layouts` tables (a subset of the built ones) is
 * set to EMPTY, then every (set, element) of the borrowed CSR is inserted
 * again -- e.g. after the elements were rewritten in place with the same
 * offsets.  These are the phases the reference times per layout
 * (triangle-count/hashtable.cu:164-177, build_hash_table_normal /
 * _hierarchical after their memsets).  No allocation, no synchronization
 * (enqueue only): the lengths and offsets are kept, the tables must not be
 * in use by other work on the device, and a full table is reported by
 * rs_tables_status (RS_EOVERFLOW). */
rs_status rs_tables_rebuild_async(rs_tables* t, int32_t layouts, rs_stream stream);
/* Getters of HashTableBuilder (hash_table.cuh:75-80). */
rs_status rs_tables_describe(const rs_tables* t, rs_tables_desc* out);
/* Replaces ~HashTableBuilder / free_hash_tables (hash_table.cuh:22-72). */
void rs_tables_free(rs_tables* t);
/* Synchronizes `stream` and reports (then clears) device-side error bits set
 * by kernels that used this handle: RS_EOVERFLOW for an exhausted probe. */
rs_status rs_tables_status(rs_tables* t, rs_stream stream);

/* Kernel timing (no counterpart in the reference, which times with clock()
 * around a synchronize, TC.cu:299-306): when enabled, rs_tc_count_async and
 * the IR queries on this handle record HIP events on their stream around
 * their kernels (TC: the staged and tiled count kernels; IR: the segment
 * filter and gather); rs_tables_kernel_ms waits for and returns the last
 * one's duration.  Must be disabled while capturing a graph. */
rs_status rs_tables_timing(rs_tables* t, int32_t enable);
rs_status rs_tables_kernel_ms(rs_tables* t, float* h_ms);

/* Row expansion vertexs[e] = row of entry e (triangle-count/hashtable.cu:148-158,
 * TC.cu:277-282, done on the host there; a device kernel here). */
rs_status rs_expand_rows(const int32_t* d_offsets, int64_t num_sets, int64_t num_elements,
                         int32_t* d_rows, rs_stream stream);

/* ------------------------------------------------------------------------ a7 */
/* Replaces launch_extract_hashtable_to_csr (baseline/common/utils.cu:148-175)
 * and gpu_sort_csr_cols (TC.cu:107-126): writes every set's elements in the
 * (bucket, slot) order of its `layout` table -> d_out[E] (probe order). */
rs_status rs_tables_probe_order(const rs_tables* t, int32_t layout, int32_t* d_out,
                                rs_stream stream);

/* ------------------------------------------------------------------------ a9 */
/* Triangle counting over a degree- or id-oriented CSR whose tables `t` were
 * built over that same CSR.  Replaces dynamic_triangle_count (TC.cu:128-190)
 * and TCBaseline::run_normal / run_hierarchical (baseline/tc/tc.cuh:29-37).
 * d_probe_keys: the CSR elements in the order the probes read them (NULL =
 * the CSR itself; rs_tables_probe_order output for the hierarchical layout). */
typedef struct rs_tc_plan rs_tc_plan;
/* Edge schedule: per-edge row ids and the int64 prefix of probe counts
 * min(d+(u), d+(v)), which the tiled kernel splits into equal probe tiles
 * (hub edges are split across waves).  Reusable across layouts. */
rs_status rs_tc_plan_create(const rs_tables* t, rs_stream stream, rs_tc_plan** out);
void rs_tc_plan_free(rs_tc_plan* p);
/* Total probes (sum over edges of min(d+(u), d+(v))) and edges of the plan. */
rs_status rs_tc_plan_work(const rs_tc_plan* p, int64_t* h_probes, int64_t* h_edges);
/* The plan's LDS-staged share: probes whose table is copied into LDS once per
 * unit (tables probed >= 2x per slot), the number of units and the largest
 * staged table length.  Diagnostic; no counterpart in the reference (its
 * kernel reads every table slot from global memory, TC.cu:11-31). */
rs_status rs_tc_plan_staged(const rs_tc_plan* p, int64_t* h_staged_probes, int64_t* h_units,
                            int64_t* h_tables_max_len);
/* Tiled wave64 count into *d_count (uint64, overwritten).  The plan holds
 * the work-stealing counters of its kernels (and of the plain form), reset
 * by a kernel at the start of every count: counts with one plan must not
 * overlap on the device (use a plan per stream). */
rs_status rs_tc_count_async(const rs_tables* t, const rs_tc_plan* p, int32_t layout,
                            const int32_t* d_probe_keys, uint64_t* d_count, rs_stream stream);
/* The plain form: one wave per edge, CHUNK-edge work stealing, like the
 * reference kernel (TC.cu:141-183).  Same result. */
rs_status rs_tc_count_plain_async(const rs_tables* t, const rs_tc_plan* p, int32_t layout,
                                  const int32_t* d_probe_keys, uint64_t* d_count,
                                  rs_stream stream);
/* Convenience: plan (cached in t), count, synchronize, *h_count.  Calls on
 * one t from several host threads take turns (the cached plan's counters
 * serve one count at a time). */
rs_status rs_tc_count(rs_tables* t, int32_t layout, const int32_t* d_probe_keys,
                      uint64_t* h_count, rs_stream stream);

/* ----------------------------------------------------------------------- a10 */
/* Set-similarity verification.  Replaces set_similarity_search_kernel
 * (set-similarity-search/sss.cu:93-154) and SSSBaseline::run_* (baseline/sss/
 * sss.cuh:26-35): for candidate e = (u, v), r = |S_u cap S_v| by probing the
 * smaller set into the larger one's table; d_flags[e] = 1 iff
 * q*r >= p*(|S_u|+|S_v|-r) and |S_u|+|S_v|-r > 0 (exact form of the
 * reference's float `r / (float)(|u|+|v|-r) >= threshold`).  d_sizes[e] = r
 * (required, int32[P], used as the accumulator).  d_flags may be NULL. */
rs_status rs_sss_verify_async(const rs_tables* t, int32_t layout, const int32_t* d_probe_keys,
                              const int32_t* d_cand_u, const int32_t* d_cand_v,
                              int64_t num_candidates, int32_t p, int32_t q, int32_t* d_flags,
                              int32_t* d_sizes, rs_stream stream);

/* Prefix-filter self-join of a set collection at Jaccard >= p/q: global
 * token order by ascending frequency, prefix candidates with the size filter,
 * verification of every candidate by LDS hash tables (a table per build set
 * shared by its candidates; the batched pair kernel, indexed form, for build
 * sets too large for one workgroup's table) and the exact rule.
 * load_factor (0, 1] and bucket_size (1, 2 or 4) are the verification
 * tables' hyperparameters (0.25 and 4 are the tuned values; the reference's
 * 5-slot buckets have no LDS form here and are rejected, RS_EINVAL).  Sets
 * must hold strictly ascending, non-negative tokens: checked, RS_EINVAL.
 * The result is the sorted list of pairs (i < j) as int64 keys
 * (i << 32 | j), held by the handle. */
typedef struct rs_sss_result rs_sss_result;
rs_status rs_sss_selfjoin(const int32_t* d_offsets, int64_t num_sets, const int32_t* d_elements,
                          int32_t p, int32_t q, float load_factor, int32_t bucket_size,
                          rs_stream stream, rs_sss_result** out);
rs_status rs_sss_result_info(const rs_sss_result* r, int64_t* h_num_pairs,
                             int64_t* h_num_candidates, const int64_t** d_pairs);
/* Work of the verification stage: sum over candidates of the smaller set's
 * size, and of both sizes (the elements the batched kernel streams). */
rs_status rs_sss_result_work(const rs_sss_result* r, int64_t* h_verify_probes,
                             int64_t* h_verify_elements);
/* Stage timings of the last self-join (ms): token order, candidates, verify. */
rs_status rs_sss_result_timings(const rs_sss_result* r, float* h_ms3);
void rs_sss_result_free(rs_sss_result* r);

/* Similarity lookups against a resident collection (the lookup side of the
 * serving loop; the reference has no lookup form -- its sss.cu verifies a
 * given candidate list).  The index keeps the collection's global token order
 * (ascending frequency, ties by id) and its sorted prefix entries for the
 * threshold p/q; it borrows the collection CSR (ascending, distinct tokens,
 * ids < 2^22, < 2^21 sets).  A lookup batch gives Q query sets as a CSR
 * (ascending, distinct tokens; tokens outside the collection are allowed and
 * rank before all others) and returns every (query, set) with
 * J >= p/q exactly, as int64 keys (query << 32 | set) in ascending order,
 * held by an rs_sss_result (rs_sss_result_info / _work / _timings / _free). */
typedef struct rs_sss_index rs_sss_index;
rs_status rs_sss_index_build(const int32_t* d_offsets, int64_t num_sets, const int32_t* d_elements,
                             int32_t p, int32_t q, rs_stream stream, rs_sss_index** out);
void rs_sss_index_free(rs_sss_index* idx);
rs_status rs_sss_lookup(const rs_sss_index* idx, const int32_t* d_query_offsets,
                        int64_t num_queries, const int32_t* d_query_tokens, rs_stream stream,
                        rs_sss_result** out);

/* ----------------------------------------------------------------------- a11 */
/* Conjunctive queries.  Replaces ir_kernel (information-retrieval/ir.cu:94-176)
 * and IRBaseline::run_* (baseline/ir/ir.cuh:39-46).  Tables `t` are built over
 * the inverted index (sets = terms, elements = doc ids).  Query q has terms
 * d_query[d_query_offsets[q] .. d_query_offsets[q+1]) with term0 first (the
 * shortest list, as the preprocessing orders them).  Its result region is
 * d_result[d_result_offsets[q] ..) of |list(term0)| entries (ir.cu:324-330);
 * the result keeps term0's order as read from d_term0_docs (NULL = the index
 * elements, i.e. ascending doc ids).  d_result_count[q] = result size.
 * d_workspace: int32 scratch of d_result_offsets[Q] entries. */
rs_status rs_ir_query_async(const rs_tables* t, int32_t layout, const int32_t* d_term0_docs,
                            const int32_t* d_query, const int32_t* d_query_offsets,
                            int64_t num_queries, const int64_t* d_result_offsets,
                            int32_t* d_result, int32_t* d_result_count, int32_t* d_workspace,
                            rs_stream stream);

/* rs_ir_query_async without allocation.  No counterpart in the reference:
 * this entry point is for callers that cannot allocate device memory (tight
 * serving loops, code that forbids hipMalloc).  Same arguments, results and
 * errors as rs_ir_query_async; its scratch (segment offsets, counts, scan
 * sums and work counters, about 4.8 MB) is a fixed array in the library's
 * device memory, one per GPU, held until the process ends.  Limits per call:
 * at most 65,536 queries and 262,144 term0 segments of 4,096 entries
 * (RS_EINVAL above them; the segment count is checked after one
 * synchronization of `stream`, and nothing is written then).  Calls that
 * overlap on one GPU race on that array: issue them on one stream or one
 * after another.  Returns once the result kernels are enqueued. */
rs_status rs_ir_query_noalloc(const rs_tables* t, int32_t layout, const int32_t* d_term0_docs,
                              const int32_t* d_query, const int32_t* d_query_offsets,
                              int64_t num_queries, const int64_t* d_result_offsets,
                              int32_t* d_result, int32_t* d_result_count, int32_t* d_workspace,
                              rs_stream stream);

/* Batches of at least `segments` term0 segments (4096 entries each) are
 * split among waves by work stealing, smaller ones statically; -1 (the
 * default) means four runs of four segments per wave of the grid.  Results do
 * not depend on it; it lets a test run either schedule on any input.  Global
 * to the process; not synchronized with queries in flight. */
rs_status rs_ir_set_steal_threshold(int64_t segments);

/* Serving loop with continuous batching (no counterpart in the reference,
 * which runs one static batch) for two request kinds: IR conjunctive queries
 * against the resident index `t`, and -- once an rs_sss_index is attached --
 * set-similarity lookups against its resident collection.  A host thread
 * admits submitted IR queries while their result memory fits
 * `memory_budget_bytes`, at most `max_batch` per launch, keeping up to
 * `max_inflight` launches overlapping on their own streams; a second thread
 * batches pending lookups (same budget and batch limit) on its own stream,
 * so both kinds overlap on the device.  Each request completes when its
 * batch finishes; its result does not depend on what it was batched with.
 * Query ids are the caller's and must be unique across both kinds. */
typedef struct rs_server rs_server;
typedef struct rs_completion {
  int64_t query_id;
  int32_t count;          /* result size, or -1 on error                     */
  int32_t status;         /* RS_OK / error                                   */
  double submit_ms;       /* host clock at submit, ms                        */
  double done_ms;         /* host clock when the result was ready, ms        */
  int32_t kind;           /* RS_REQ_IR or RS_REQ_SSS                         */
  int32_t reserved;
} rs_completion;
#define RS_REQ_IR 0
#define RS_REQ_SSS 1
rs_status rs_server_create(const rs_tables* t, int32_t layout, int64_t memory_budget_bytes,
                              int32_t max_batch, int32_t max_inflight, rs_server** out);
/* Enqueue one query (terms on the host, term0 first).  RS_EBUSY if the
 * pending queue is full. */
rs_status rs_server_submit_ir(rs_server* s, int64_t query_id, const int32_t* h_terms,
                              int32_t num_terms);
/* Attach a similarity index (kept borrowed; it must outlive the server) and
 * start the lookup thread.  Once per server.  From then on IR batches use at
 * most GPU_MAX_HW_QUEUES - 1 (3 by default) of the max_inflight slots, so the
 * lookup stream keeps a hardware queue of its own. */
rs_status rs_server_attach_sss(rs_server* s, const rs_sss_index* idx);
/* Enqueue one similarity lookup: the query set's tokens on the host,
 * strictly ascending and non-negative.  Its result is the ascending list of
 * collection set ids with J >= the index's p/q.  RS_EINVAL without an
 * attached index, RS_EBUSY if the pending queue is full. */
rs_status rs_server_submit_sss(rs_server* s, int64_t query_id, const int32_t* h_tokens,
                               int32_t num_tokens);
/* Drain up to `cap` completions; *h_n receives how many were written. */
rs_status rs_server_poll(rs_server* s, rs_completion* h_out, int32_t cap, int32_t* h_n);
/* Copy a completed request's result (IR: doc ids in term0 order, ascending
 * if the index elements are sorted; lookup: ascending set ids) to h_docs (cap
 * entries); *h_count = its full size.  The result
 * is released once it was copied whole; call again with a larger buffer when
 * *h_count > cap. */
rs_status rs_server_result(rs_server* s, int64_t query_id, int32_t* h_docs, int64_t cap,
                              int64_t* h_count);
/* Block until every submitted query completed. */
rs_status rs_server_drain(rs_server* s);
/* Failure handling of the loop (no counterpart in the reference, which has
 * no serving loop).  Cancel a request that is still queued: it completes at
 * once with status RS_ECANCELED and count -1 and is never launched;
 * RS_EINVAL if the id is not queued (already admitted, completed or
 * unknown) -- an admitted request runs to completion. */
rs_status rs_server_cancel(rs_server* s, int64_t query_id);
/* Queue timeout: a request that has waited longer than `ms` when the server
 * comes to admit it completes with RS_ETIMEDOUT and count -1, unlaunched.
 * 0 (the default) disables the timeout. */
rs_status rs_server_set_queue_timeout(rs_server* s, double ms);
/* Pause (1) or resume (0) admission; batches already launched complete. */
rs_status rs_server_pause(rs_server* s, int32_t paused);
/* Launches (IR and lookup batches), completed requests and peak admitted
 * bytes so far. */
rs_status rs_server_stats(const rs_server* s, int64_t* h_launches, int64_t* h_queries,
                             int64_t* h_peak_bytes);
void rs_server_free(rs_server* s);

/* ----------------------------------------------------------------------- a12 */
/* Pairwise set intersection sizes for P pairs (A_i, B_i) given as two CSR
 * families with int64 offsets.  Elements within a set must be distinct.
 * Sets need not be sorted for the hash probe; the batched form requires
 * ascending sets when the smaller set exceeds the bin's chunk (below), and
 * reports sets out of order there (RS_EINVAL from rs_pairs_plan_status)
 * rather than counting them.  Bins with different bucket sizes or lanes per
 * pair run as consecutive kernels, one per run of equal settings. */
typedef struct rs_pairs_bin {
  float load_factor;   /* alpha of the per-pair LDS table                  */
  int32_t bucket_size; /* b: slots per bucket, 1, 2 or 4 (one LDS read)    */
  int32_t chunk;       /* build-side elements of one table, 64..1024; a   */
                       /* larger build side goes in chunks of <= 252      */
  int32_t unroll;      /* 16-byte loads per lane in flight (1, 2, 4) where */
                       /* |larger| >= 4 |smaller| (else 1); same in every  */
                       /* bin                                              */
  int32_t lanes;       /* probe width: lanes per pair, 64 (a wave), 32 or  */
                       /* 16 (2 or 4 pairs per wave at once, the smaller   */
                       /* set built whole); below 64 only for a bin other  */
                       /* than the last whose table of 2^bin - 1 keys fits */
                       /* 1/(64/lanes) of a wave's LDS region              */
} rs_pairs_bin;
#define RS_PAIRS_MAX_BINS 20
typedef struct rs_pairs_opts {
  int32_t num_bins;                 /* bins by floor(log2(|A|+|B|)), <= 20   */
  rs_pairs_bin bins[RS_PAIRS_MAX_BINS];
  int32_t waves_per_cu;             /* persistent grid occupancy (4..32)    */
  int32_t lds_slots_per_wave;       /* int32 LDS slots of a wave's table;   */
                                    /* (slots + 512 staging) x 4 B x waves  */
                                    /* per CU must fit the 160 KiB LDS      */
  int32_t schedule;                 /* 0 static, 1..64 work stealing        */
} rs_pairs_opts;
/* How waves of the batched launch take pairs from the size-sorted order:
 * schedule 0 (RS_PAIRS_STATIC), wave w takes positions w, w + W, ... (no
 * atomics); schedule k in 1..64, each wave takes runs of k consecutive
 * positions from a shared counter, the reference's work stealing
 * (TC.cu:141-142,176-182, CHUNK = k).  Runs are shortened to leave at least
 * 4 runs per wave of the grid; a batch too small for runs of
 * min(k, 4) is dealt statically.  The default is RS_PAIRS_STEAL, runs of 16. */
#define RS_PAIRS_STATIC 0
#define RS_PAIRS_STEAL 16
/* Defaults tuned for gfx950 (also the autotuner's starting point). */
void rs_pairs_default_opts(rs_pairs_opts* o);

/* The plain form of set_size/intersection.cu:128-181: for each pair, a
 * workspace table of the larger set is filled with -1, built and probed by
 * the smaller set, one launch sequence per pair on `stream`, which is then
 * synchronized (the reference synchronizes after each pair, :221-243).
 * d_workspace must hold bucket_size * calculate_length(max |B|) int32.
 * RS_EOVERFLOW, before any launch, if a set has more keys than its table has
 * slots (load_factor above 1), and after the launches if a probe walked a
 * whole (full) table without meeting an EMPTY slot. */
rs_status rs_pairs_intersect_plain(const int64_t* d_a_off, const int32_t* d_a,
                                   const int64_t* d_b_off, const int32_t* d_b,
                                   const int64_t* h_a_off, const int64_t* h_b_off, int64_t num_pairs,
                                   float load_factor, int32_t bucket_size, int32_t* d_workspace,
                                   int32_t* d_sizes, rs_stream stream);

/* The batched form: one persistent, size-binned launch whose waves take pairs
 * statically or by work stealing (opts.schedule).  A
 * plan bins the pairs (largest first) once per batch; the launch builds each
 * pair's smaller set into an LDS table and streams the larger set through
 * it; a smaller set above the bin's chunk is built chunk by chunk in value
 * order (ascending sets needed).
 * The plan + launch sequence contains no memset and can be replayed from a
 * hipGraph (rs_pairs_graph_*). */
typedef struct rs_pairs_plan rs_pairs_plan;
rs_status rs_pairs_plan_create(int64_t max_pairs, const rs_pairs_opts* opts, rs_pairs_plan** out);
void rs_pairs_plan_free(rs_pairs_plan* p);
rs_status rs_pairs_intersect_batched_async(rs_pairs_plan* plan, const int64_t* d_a_off,
                                           const int32_t* d_a, const int64_t* d_b_off,
                                           const int32_t* d_b, int64_t num_pairs, int32_t* d_sizes,
                                           rs_stream stream);
/* The batched form over pairs of one set family: pair i is (S_{u[i]}, S_{v[i]})
 * of the CSR (d_set_offsets, d_set_elements) -- e.g. the candidate list of a
 * similarity join.  Same plan, kernel and requirements as above. */
rs_status rs_pairs_intersect_indexed_async(rs_pairs_plan* plan, const int32_t* d_set_offsets,
                                           const int32_t* d_set_elements, const int32_t* d_u,
                                           const int32_t* d_v, int64_t num_pairs, int32_t* d_sizes,
                                           rs_stream stream);
/* The batched form over pairs drawn from two set families: pair i is
 * (A_{u[i]}, B_{v[i]}), e.g. lookup queries against a resident collection. */
rs_status rs_pairs_intersect_cross_async(rs_pairs_plan* plan, const int32_t* d_a_offsets,
                                         const int32_t* d_a_elements, const int32_t* d_b_offsets,
                                         const int32_t* d_b_elements, const int32_t* d_u,
                                         const int32_t* d_v, int64_t num_pairs, int32_t* d_sizes,
                                         rs_stream stream);
/* Capture the batched sequence for these buffers into a hipGraph held by the
 * plan, then replay it with rs_pairs_graph_launch. */
rs_status rs_pairs_graph_capture(rs_pairs_plan* plan, const int64_t* d_a_off, const int32_t* d_a,
                                 const int64_t* d_b_off, const int32_t* d_b, int64_t num_pairs,
                                 int32_t* d_sizes, rs_stream stream);
rs_status rs_pairs_graph_launch(rs_pairs_plan* plan, rs_stream stream);
/* Kernel timing: when enabled, rs_pairs_intersect_batched_async records HIP
 * events on its stream around the batched kernel (not the plan kernels);
 * rs_pairs_plan_kernel_ms waits for and returns the last launch's duration.
 * Must be disabled while capturing a graph. */
rs_status rs_pairs_plan_timing(rs_pairs_plan* p, int32_t enable);
rs_status rs_pairs_plan_kernel_ms(rs_pairs_plan* p, float* h_ms);
/* Device status of the plan's last launches (RS_EOVERFLOW on a full table). */
rs_status rs_pairs_plan_status(rs_pairs_plan* p, rs_stream stream);

/* Autotuner: times candidate (alpha, b, chunk, lanes) settings per size bin
 * on the given pairs (a sample of the workload) and writes the fastest into
 * *opts (the unroll stays the caller's); b in {1, 2, 4}, lanes per pair in
 * {64, 32, 16} where the bin allows it.  The analytic cost model of
 * hyperparameter_solver/src/solver.py ranks the candidates it tries first. */
rs_status rs_pairs_autotune(const int64_t* d_a_off, const int32_t* d_a, const int64_t* d_b_off,
                            const int32_t* d_b, int64_t num_pairs, int32_t reps,
                            rs_pairs_opts* opts, rs_stream stream);

/* ------------------------------------------------------ hyperparameter model */
/* The reference's analytic cost model (hyperparameter_solver/src/solver.py:
 * 132-387): expected probe steps x step cost of a bucketed table of arr_size
 * entries at load factor alpha with b-slot buckets, for warp width W,
 * transaction bytes S_trans and slot bytes S_slot (binomial bucket loads,
 * Markov overflow chain, occupancy by convolution).  Double precision. */
double rs_hyper_cost(int64_t arr_size, double alpha, int32_t b, int32_t W, int32_t S_trans,
                     int32_t S_slot);
/* solve_hyperparameters (c_api.py:288-338): the least-cost grid point,
 * alpha in {0.10, 0.15, ..., 1.00} and b in {1..9}. */
rs_status rs_hyper_solve(int64_t arr_size, int32_t W, int32_t S_trans, int32_t S_slot,
                         double* h_alpha, int32_t* h_b, double* h_cost);

/* -------------------------------------------------------- competitor tables */
/* The per-set tables the reference's baseline directory compares RESET
 * against, one flat region per set of the CSR they are built from (which
 * the handle borrows):
 *  - RS_COMP_LINEAR: cuCollections static_set with linear probing, 1-slot
 *    buckets and MurmurHash3 fmix32 (baseline/common/cuco.cuh:15-96,
 *    cuco.cu:35-80), capacity the next power of two of degree/alpha;
 *  - RS_COMP_CUCKOO: FlatCuckooHash (baseline/common/cuckoo_hash.cuh:
 *    37-135, cuckoo_hash.cu:40-228): two hashes, <= 6000 evictions, a
 *    16-entry stash per set;
 *  - RS_COMP_HOPSCOTCH: HopscotchHash (hopscotch_hash.cuh:35-122,
 *    hopscotch_hash.cu:27-236): H = 32 neighbourhoods, load factor capped
 *    at 0.1;
 *  - RS_COMP_ROARING: RoaringBitmap (roaring_bitmap.cuh:15-113,
 *    roaring_bitmap.cu:16-139): ids remapped to their rank, sorted 64-bit
 *    words per set (rows must be ascending; load_factor unused).
 * Elements must be >= 0 and != INT32_MAX (the EMPTY key).  A key that a
 * table cannot place (linear table full, cuckoo stash full, hopscotch
 * displacement failure) is dropped as in the reference and counted. */
#define RS_COMP_LINEAR 0
#define RS_COMP_CUCKOO 1
#define RS_COMP_HOPSCOTCH 2
#define RS_COMP_ROARING 3
typedef struct rs_comp rs_comp;
rs_status rs_comp_build(const int32_t* d_offsets, int64_t num_sets, const int32_t* d_elements,
                        int32_t kind, float load_factor, rs_stream stream, rs_comp** out);
/* Slots (words for roaring) and the number of dropped keys. */
rs_status rs_comp_info(const rs_comp* c, int64_t* h_slots, int64_t* h_failed);
/* The table's arrays, for inspection (0 bytes where a kind has no such
 * array): region offsets (int64 [n+1]), slots (int32), stash counts (int32
 * [n]) and stash (int32 [16 n]), hopscotch maps (uint32 per slot), roaring
 * remap (int32, -1 for ids outside the collection), node starts (int32
 * [n+1]), word indices (int32) and words (uint64). */
#define RS_COMP_ARR_OFFSETS 0
#define RS_COMP_ARR_SLOTS 1
#define RS_COMP_ARR_STASH_COUNT 2
#define RS_COMP_ARR_STASH 3
#define RS_COMP_ARR_BITMAP 4
#define RS_COMP_ARR_REMAP 5
#define RS_COMP_ARR_NODE_START 6
#define RS_COMP_ARR_WORD_INDEX 7
#define RS_COMP_ARR_WORD_DATA 8
rs_status rs_comp_array_bytes(const rs_comp* c, int32_t array, int64_t* h_bytes);
rs_status rs_comp_copy_out(const rs_comp* c, int32_t array, void* h_dst, int64_t bytes);
/* Triangle counting over the CSR the tables were built from: for every
 * entry (u, v), |N(s) n N(g)| probed through g's table with s the endpoint
 * of smaller degree (baseline/tc/tc.cu:35-108, tc_cuco.cu:19-100; the w > v
 * filter of tc_cuco.cu:81-86 is left out, so degree-oriented CSR counts
 * too).  Synchronizes the stream. */
rs_status rs_comp_tc_count(rs_comp* c, uint64_t* h_count, rs_stream stream);
/* SSS verification of candidate pairs (u_k, v_k) of the table's sets
 * (sss/sss_cuckoo.cu:19-107, sss_hopscotch.cu, sss_roaring.cu, sss_cuco.cu):
 * sizes r and flags q*r >= p*(|u|+|v|-r).  Enqueue only. */
rs_status rs_comp_sss_verify_async(rs_comp* c, const int32_t* d_u, const int32_t* d_v,
                                   int64_t num_pairs, int32_t p, int32_t q, int32_t* d_flags,
                                   int32_t* d_sizes, rs_stream stream);
/* IR conjunctive queries (ir/ir_cuco.cu:11-110) over tables built on an
 * inverted index (sets = terms, elements = doc ids): query q's terms are
 * d_query[d_query_offsets[q] .. d_query_offsets[q+1]), term0 first; its
 * survivors, in term0's order, go to d_result[d_result_offsets[q] ..) (a
 * region of |list(term0)| entries) and their number to d_result_count[q].
 * Enqueue only. */
rs_status rs_comp_ir_query_async(rs_comp* c, const int32_t* d_query, const int32_t* d_query_offsets,
                                 int64_t num_queries, const int64_t* d_result_offsets,
                                 int32_t* d_result, int32_t* d_result_count, rs_stream stream);
/* Bad candidate ids or query terms from earlier calls (RS_EINVAL); synchronizes. */
rs_status rs_comp_status(rs_comp* c, rs_stream stream);
void rs_comp_free(rs_comp* c);

/* --------------------------------------------------------------- generators */
/* Synthetic inputs (not on the measured path): counter-based, seeded,
 * reproducible.  Pair sets: A_i = sorted(pi_i[0 .. |A_i|)),
 * B_i = sorted(pi_i[|A_i|-k_i .. |A_i|-k_i+|B_i|)) for a keyed bijection pi_i
 * of [0, 2^24), so |A_i cap B_i| = k_i exactly. */
rs_status rs_gen_pair_sets_async(const int64_t* d_a_off, const int64_t* d_b_off,
                                 const int32_t* d_k, int64_t num_pairs, uint64_t seed,
                                 int32_t* d_a, int32_t* d_b, rs_stream stream);
/* R-MAT edges (Chakrabarti et al. 2004) with integer quadrant thresholds
 * a,b,c (of 2^32) and vertex labels scrambled by a keyed bijection of
 * [0, 2^scale): d_edges[2*i], d_edges[2*i+1] for i < num_edges. */
rs_status rs_gen_rmat_async(int32_t scale, int64_t num_edges, uint32_t a, uint32_t b, uint32_t c,
                            uint64_t seed, int32_t* d_edges, rs_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* RESET_AMD_H */
