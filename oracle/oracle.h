/*
 * oracle.h -- CPU restatement of RESET's hash-based set-intersection path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke() entry of
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load this
 * library, and only as the checker (or as the timed CPU baseline).  Nothing in
 * reset_amd/ links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Parity of each function is pinned by tests/test_oracle.py against closed
 * forms, brute force and textbook identities (known-answer graphs, sorted
 * merge intersections, exhaustive float32 checks); the reference ships no
 * golden vectors for this path.
 *
 * Conventions: elements are int32 bit patterns, EMPTY slot = -1 (0xFFFFFFFF),
 * so no element may equal -1.  Table offsets and the global bucket count BN
 * are int64 (the reference's int BN overflows at R-MAT scale 24).
 */
#ifndef RESET_ORACLE_H
#define RESET_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_EOVERFLOW 3

#define ORC_LAYOUT_NORMAL 0
#define ORC_LAYOUT_HIERARCHICAL 1

/* triangle-count/hashtable.cu:20-36 (same in set_size/intersection.cu:72-87) */
int32_t orc_calculate_length(int32_t degree, float load_factor, int32_t bucket_size);

/* triangle-count/TC.cuh:48-52 and TC.cuh:40-46 */
uint32_t orc_hash_normal(uint32_t x, uint32_t L);
uint32_t orc_hash_hierarchical(uint32_t x, uint32_t L, uint32_t M);

/* triangle-count/hashtable.cu:103-130: per-set bucket counts L[n], exclusive
 * int64 offsets toff[n+1] with toff[n] = BN, and M = max L.  Returns ORC_OK. */
int orc_table_geometry(const int32_t* off, int64_t n, float load_factor,
                       int32_t bucket_size, int32_t* L, int64_t* toff,
                       int32_t* max_len);

/* triangle-count/hashtable.cu:53-101: sequential insert of every (row, elem)
 * in CSR order into the slot-plane table T[toff[t] + j + k*BN] (T must hold
 * BN*b int32 and is filled with -1 here).  Bounded: ORC_EOVERFLOW if a row's
 * table has no free slot.  *conflicts = failed slot attempts. */
int orc_table_build(const int32_t* off, const int32_t* elems, int64_t n,
                    const int32_t* L, const int64_t* toff, int32_t bucket_size,
                    int32_t max_len, int32_t layout, int32_t* T,
                    int64_t* conflicts);

/* triangle-count/TC.cu:11-31 (bounded at L*b slots).  Tt points at table t's
 * first bucket.  Returns 1 found, 0 absent, -1 probe bound exhausted. */
int orc_table_search(const int32_t* Tt, int64_t BN, uint32_t bucket, int32_t L,
                     int32_t bucket_size, int32_t key);

/* Invariant check of a table built by anyone (schedule-dependent placement is
 * allowed): every slot lies in its row's region, every stored element belongs
 * to its row, row multiplicities match, each stored element is reachable
 * from its home bucket before an EMPTY slot, and the occupied slots of every
 * bucket form a prefix of it (the property a bucket-at-a-time probe relies
 * on; reachability implies it).  Returns the number of violations. */
int64_t orc_table_check(const int32_t* off, const int32_t* elems, int64_t n,
                        const int32_t* L, const int64_t* toff,
                        int32_t bucket_size, int32_t max_len, int32_t layout,
                        const int32_t* T);

/* baseline/common/utils.cu:76-146: each row's elements in (bucket, slot) order
 * of its table. */
int orc_extract_probe_order(const int32_t* off, int64_t n, const int64_t* toff,
                            int32_t bucket_size, const int32_t* T, int32_t* out);

/* triangle-count/TC.cu:128-190: sum over CSR edges (u, cols[e]) of the number
 * of keys of the smaller out-list found in the larger's table. */
int orc_tc_hash(const int32_t* off, const int32_t* cols, int64_t n,
                const int32_t* L, const int64_t* toff, int32_t bucket_size,
                int32_t max_len, int32_t layout, const int32_t* T,
                int nthreads, uint64_t* out);
/* orc_tc_hash over the edges of source rows [row_begin, row_end) only. */
int orc_tc_hash_rows(const int32_t* off, const int32_t* cols, int64_t n, int64_t row_begin,
                     int64_t row_end, const int32_t* L, const int64_t* toff,
                     int32_t bucket_size, int32_t max_len, int32_t layout, const int32_t* T,
                     int nthreads, uint64_t* out);

/* Independent textbook check: sum over edges of |N+(u) cap N+(v)| by sorted
 * merge (rows must be sorted ascending). */
uint64_t orc_tc_merge(const int32_t* off, const int32_t* cols, int64_t n,
                      int nthreads);
/* The same count restricted to source rows [row_begin, row_end): in an
 * acyclic orientation every triangle has one source (the vertex with both of
 * its other vertices as out-neighbours), counted at its first out-edge, so
 * the rows' count is the number of triangles whose source lies in the range
 * and the counts of disjoint ranges add up to orc_tc_merge. */
uint64_t orc_tc_merge_rows(const int32_t* off, const int32_t* cols, int64_t row_begin,
                           int64_t row_end, int nthreads);

/* set_size/intersection.cu:128-181 for many pairs: per pair, a table of the
 * larger set (length rule of hashtable.cu:20-36, single-table BN = L) probed
 * by the smaller.  Pairs are CSR slices a[a_off[i]..a_off[i+1]). */
int orc_pairs_hash(const int64_t* a_off, const int32_t* a, const int64_t* b_off,
                   const int32_t* b, int64_t P, float load_factor,
                   int32_t bucket_size, int nthreads, int32_t* sizes);
void orc_pairs_merge(const int64_t* a_off, const int32_t* a,
                     const int64_t* b_off, const int32_t* b, int64_t P,
                     int nthreads, int32_t* sizes);

/* set-similarity-search/sss.cu:93-154 with the float compare replaced by the
 * exact rule q*r >= p*(|u|+|v|-r).  flags[e] = 1 if candidate e passes;
 * sizes[e] = r (either output may be NULL). */
int orc_sss_verify_hash(const int32_t* off, const int32_t* cols, int64_t n,
                        const int32_t* L, const int64_t* toff,
                        int32_t bucket_size, int32_t max_len, int32_t layout,
                        const int32_t* T, const int32_t* cand_u,
                        const int32_t* cand_v, int64_t P, int32_t p, int32_t q,
                        int nthreads, int32_t* flags, int32_t* sizes);
/* The reference's own float decision (sss.cu:130-133) for one pair. */
int orc_sss_float_rule(int32_t r, int32_t size_u, int32_t size_v, float threshold);

/* All pairs i<j of rows with J >= p/q, by sorted merge over every pair
 * (tiny inputs only).  Writes keys (i<<32|j) ascending; returns the count or
 * -1 if cap is too small. */
int64_t orc_sss_bruteforce(const int32_t* off, const int32_t* cols, int64_t n,
                           int32_t p, int32_t q, int64_t* out, int64_t cap);

/* Every (query a, row j) with J(Q_a, S_j) >= p/q, by sorted merge of each
 * query against every row (tiny inputs only); keys (a<<32|j) ascending;
 * returns the count or -1 if cap is too small. */
int64_t orc_sss_lookup_bruteforce(const int32_t* off, const int32_t* cols, int64_t n,
                                  const int32_t* qoff, const int32_t* qtok, int64_t nq,
                                  int32_t p, int32_t q, int64_t* out, int64_t cap);

/* Prefix-filter self-join (Chaudhuri et al. 2006 / Bayardo et al. 2007):
 * global order = ascending token frequency (ties by token id), prefix length
 * s - ceil(p*s/q) + 1, size filter, candidate dedup, exact verification by
 * sorted merge.  Rows must be sorted ascending.  Same output convention as
 * orc_sss_bruteforce; *num_candidates receives the deduplicated candidate count. */
int64_t orc_sss_prefix_join(const int32_t* off, const int32_t* cols, int64_t n,
                            int32_t p, int32_t q, int nthreads, int64_t* out,
                            int64_t cap, int64_t* num_candidates);

/* information-retrieval/ir.cu:94-176: per query, copy term0's list (in index
 * order) to result[roff[q]..], then keep the survivors found in each later
 * term's table (stable).  count[q] = survivors. */
int orc_ir_hash(const int32_t* idx_off, const int32_t* docs, int64_t num_terms,
                const int64_t* toff, int32_t bucket_size, int32_t max_len,
                int32_t layout, const int32_t* T, const int32_t* query,
                const int32_t* qoff, int64_t Q, const int64_t* roff,
                int nthreads, int32_t* result, int32_t* count);
/* Independent check: chained sorted-merge intersection (lists sorted). */
void orc_ir_merge(const int32_t* idx_off, const int32_t* docs,
                  const int32_t* query, const int32_t* qoff, int64_t Q,
                  const int64_t* roff, int nthreads, int32_t* result,
                  int32_t* count);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
