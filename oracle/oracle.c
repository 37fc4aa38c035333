/*
 * oracle.c -- CPU restatement of RESET's hash-based set-intersection path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE: loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker
 * and as the timed CPU baseline ("port" of the reference algorithm: same
 * length rule, hashes, slot-plane layout, smaller-probes-larger and dynamic
 * chunk-4 scheduling, with OpenMP threads standing in for GPU warps).
 *
 * Parity of every function here is pinned in tests/test_oracle.py (closed
 * forms, brute force, known-answer graphs, sorted-merge cross checks and an
 * exhaustive float32 check of the length rule).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* triangle-count/hashtable.cu:20-36.  `int total_size = degrees / load_factor`
 * is an int->float conversion, an IEEE float32 divide and a truncating
 * float->int conversion; the bucket count is then the next power of two
 * strictly above total_size / bucket_size, at least 8 (8 when that is 0).
 * Out-of-range inputs (quotients the reference would overflow on) return -1. */
int32_t orc_calculate_length(int32_t degree, float load_factor, int32_t bucket_size) {
  if (degree < 0 || !(load_factor > 0.0f) || bucket_size <= 0) return -1;
  volatile float q = (float)degree / load_factor;
  if (!(q < 2147483648.0f)) return -1;
  int32_t total = (int32_t)q;
  int32_t len = total / bucket_size;
  if (len == 0) return 8;
  if (len >= (1 << 30)) return -1;
  uint32_t x = (uint32_t)len;
  x |= x >> 1;
  x |= x >> 2;
  x |= x >> 4;
  x |= x >> 8;
  x |= x >> 16;
  x++;
  if (x < 8) x = 8;
  return (int32_t)x;
}

/* TC.cuh:48-52 */
uint32_t orc_hash_normal(uint32_t x, uint32_t L) { return x & (L - 1u); }

/* TC.cuh:40-46: (x & (M-1)) >> (ffs(M) - ffs(L)); both are powers of two so
 * the shift is log2(M) - log2(L). */
uint32_t orc_hash_hierarchical(uint32_t x, uint32_t L, uint32_t M) {
  int shift = __builtin_ffs((int)M) - __builtin_ffs((int)L);
  return (x & (M - 1u)) >> shift;
}

static inline uint32_t home_bucket(int32_t layout, int32_t v, int32_t L, int32_t M) {
  return layout == ORC_LAYOUT_HIERARCHICAL
             ? orc_hash_hierarchical((uint32_t)v, (uint32_t)L, (uint32_t)M)
             : orc_hash_normal((uint32_t)v, (uint32_t)L);
}

/* hashtable.cu:103-130 (calculate_hash_length kernel + CUB exclusive scan,
 * bucket_num = last offset + last length). */
int orc_table_geometry(const int32_t* off, int64_t n, float load_factor,
                       int32_t bucket_size, int32_t* L, int64_t* toff,
                       int32_t* max_len) {
  int32_t m = 0;
  int64_t acc = 0;
  for (int64_t i = 0; i < n; i++) {
    int32_t l = orc_calculate_length(off[i + 1] - off[i], load_factor, bucket_size);
    if (l < 0) return ORC_EINVAL;
    L[i] = l;
    toff[i] = acc;
    acc += l;
    if (l > m) m = l;
  }
  toff[n] = acc;
  *max_len = m;
  return ORC_OK;
}

/* hashtable.cu:53-101: for each (row u, element v): starting at slot k = 0 of
 * the home bucket, try slots k = 0..b-1 of the bucket (slot k of bucket j of
 * table t lives at T[toff[t] + j + k*BN]), then move to bucket (j+1) mod L.
 * The reference inserts with atomicCAS(-1 -> v) from many threads; inserting
 * sequentially in CSR order is one of its admissible schedules. */
int orc_table_build(const int32_t* off, const int32_t* elems, int64_t n,
                    const int32_t* L, const int64_t* toff, int32_t bucket_size,
                    int32_t max_len, int32_t layout, int32_t* T,
                    int64_t* conflicts) {
  const int64_t BN = toff[n];
  const int64_t total = BN * (int64_t)bucket_size;
  for (int64_t i = 0; i < total; i++) T[i] = -1;
  int64_t c = 0;
  for (int64_t t = 0; t < n; t++) {
    const int32_t len = L[t];
    int32_t* Tt = T + toff[t];
    const int64_t bound = (int64_t)len * bucket_size;
    for (int32_t e = off[t]; e < off[t + 1]; e++) {
      const int32_t v = elems[e];
      uint32_t bucket = home_bucket(layout, v, len, max_len);
      int32_t k = 0;
      int64_t tries = 0;
      while (Tt[bucket + (int64_t)k * BN] != -1) {
        if (++tries >= bound) return ORC_EOVERFLOW;
        k++;
        c++;
        if (k == bucket_size) {
          k = 0;
          bucket = (bucket + 1u) % (uint32_t)len;
        }
      }
      Tt[bucket + (int64_t)k * BN] = v;
    }
  }
  if (conflicts) *conflicts = c;
  return ORC_OK;
}

/* TC.cu:11-31: compare with the key first, then with EMPTY; after b slots move
 * to the next bucket (mod L).  The reference loops without bound; here the
 * walk stops after L*b slots and reports -1. */
int orc_table_search(const int32_t* Tt, int64_t BN, uint32_t bucket, int32_t L,
                     int32_t bucket_size, int32_t key) {
  const int64_t bound = (int64_t)L * bucket_size;
  int32_t index = 0;
  for (int64_t step = 0; step < bound; step++) {
    const int32_t val = Tt[bucket + (int64_t)index * BN];
    if (val == key) return 1;
    if (val == -1) return 0;
    index++;
    if (index == bucket_size) {
      index = 0;
      bucket = (bucket + 1u) & (uint32_t)(L - 1);
    }
  }
  return -1;
}

static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int64_t orc_table_check(const int32_t* off, const int32_t* elems, int64_t n,
                        const int32_t* L, const int64_t* toff,
                        int32_t bucket_size, int32_t max_len, int32_t layout,
                        const int32_t* T) {
  const int64_t BN = toff[n];
  int64_t bad = 0;
  int32_t* a = NULL;
  int32_t* b = NULL;
  int64_t cap = 0;
  for (int64_t t = 0; t < n; t++) {
    const int32_t deg = off[t + 1] - off[t];
    const int64_t slots = (int64_t)L[t] * bucket_size;
    if (toff[t + 1] - toff[t] != L[t]) bad++;
    if (slots > cap || deg > cap) {
      cap = slots > deg ? slots : deg;
      a = (int32_t*)realloc(a, sizeof(int32_t) * (size_t)cap);
      b = (int32_t*)realloc(b, sizeof(int32_t) * (size_t)cap);
    }
    const int32_t* Tt = T + toff[t];
    int64_t w = 0;
    for (int32_t j = 0; j < L[t]; j++)
      for (int32_t k = 0, empty_seen = 0; k < bucket_size; k++) {
        const int32_t v = Tt[j + (int64_t)k * BN];
        if (v == -1) {
          empty_seen = 1;
          continue;
        }
        /* inserts take the first EMPTY slot of their walk and never clear
           one, so a bucket's occupied slots are a prefix of it */
        if (empty_seen) bad++;
        if (w < cap) a[w] = v;
        w++;
        if (orc_table_search(Tt, BN, home_bucket(layout, v, L[t], max_len), L[t],
                             bucket_size, v) != 1)
          bad++;
      }
    if (w != deg) {
      bad++;
      continue;
    }
    memcpy(b, elems + off[t], sizeof(int32_t) * (size_t)deg);
    qsort(a, (size_t)deg, sizeof(int32_t), cmp_i32);
    qsort(b, (size_t)deg, sizeof(int32_t), cmp_i32);
    if (memcmp(a, b, sizeof(int32_t) * (size_t)deg) != 0) bad++;
  }
  free(a);
  free(b);
  return bad;
}

/* baseline/common/utils.cu:76-146: slot index s = bucket*b + slot walks the
 * table bucket-major; non-empty slots are appended to the row in that order. */
int orc_extract_probe_order(const int32_t* off, int64_t n, const int64_t* toff,
                            int32_t bucket_size, const int32_t* T, int32_t* out) {
  const int64_t BN = toff[n];
  for (int64_t t = 0; t < n; t++) {
    const int32_t len = (int32_t)(toff[t + 1] - toff[t]);
    const int32_t* Tt = T + toff[t];
    int32_t w = off[t];
    for (int32_t j = 0; j < len; j++)
      for (int32_t k = 0; k < bucket_size; k++) {
        const int32_t v = Tt[j + (int64_t)k * BN];
        if (v != -1) {
          if (w >= off[t + 1]) return ORC_EINVAL;
          out[w++] = v;
        }
      }
    if (w != off[t + 1]) return ORC_EINVAL;
  }
  return ORC_OK;
}

static int32_t* expand_rows(const int32_t* off, int64_t n) {
  const int64_t E = off[n];
  int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
  for (int64_t t = 0; t < n; t++)
    for (int32_t e = off[t]; e < off[t + 1]; e++) rows[e] = (int32_t)t;
  return rows;
}

/* Probe every key of row u into row v's table; returns hits or -1. */
static int64_t probe_row(const int32_t* off, const int32_t* keys, int32_t u,
                         int32_t v, const int32_t* L, const int64_t* toff,
                         int64_t BN, int32_t bucket_size, int32_t max_len,
                         int32_t layout, const int32_t* T) {
  const int32_t* Tt = T + toff[v];
  const int32_t len = L[v];
  int64_t hits = 0;
  for (int32_t j = off[u]; j < off[u + 1]; j++) {
    const int32_t key = keys[j];
    const int r = orc_table_search(Tt, BN, home_bucket(layout, key, len, max_len),
                                   len, bucket_size, key);
    if (r < 0) return -1;
    hits += r;
  }
  return hits;
}

/* TC.cu:128-190 over the edges [e_begin, e_end): edge e = (rows[e - e_begin],
 * cols[e]); swap so the smaller out-list probes the larger one's table
 * (TC.cu:146-155); chunk-4 dynamic schedule (TC.cu:141-182). */
static int tc_hash_edges(const int32_t* off, const int32_t* cols, const int32_t* rows,
                         int64_t e_begin, int64_t e_end, int64_t BN, const int32_t* L,
                         const int64_t* toff, int32_t bucket_size, int32_t max_len,
                         int32_t layout, const int32_t* T, uint64_t* out) {
  uint64_t total = 0;
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total) reduction(| : failed)
  for (int64_t e = e_begin; e < e_end; e++) {
    int32_t u = rows[e - e_begin], v = cols[e];
    if (off[u + 1] - off[u] > off[v + 1] - off[v]) {
      const int32_t t = u;
      u = v;
      v = t;
    }
    const int64_t h = probe_row(off, cols, u, v, L, toff, BN, bucket_size,
                                max_len, layout, T);
    if (h < 0)
      failed = 1;
    else
      total += (uint64_t)h;
  }
  *out = total;
  return failed ? ORC_EOVERFLOW : ORC_OK;
}

int orc_tc_hash(const int32_t* off, const int32_t* cols, int64_t n,
                const int32_t* L, const int64_t* toff, int32_t bucket_size,
                int32_t max_len, int32_t layout, const int32_t* T,
                int nthreads, uint64_t* out) {
  int32_t* rows = expand_rows(off, n);
  set_threads(nthreads);
  const int rc = tc_hash_edges(off, cols, rows, 0, off[n], toff[n], L, toff, bucket_size,
                               max_len, layout, T, out);
  free(rows);
  return rc;
}

/* orc_tc_hash restricted to the edges of source rows [row_begin, row_end)
 * (a bounded sample of a large graph; every table is still needed). */
int orc_tc_hash_rows(const int32_t* off, const int32_t* cols, int64_t n, int64_t row_begin,
                     int64_t row_end, const int32_t* L, const int64_t* toff,
                     int32_t bucket_size, int32_t max_len, int32_t layout, const int32_t* T,
                     int nthreads, uint64_t* out) {
  if (row_begin < 0 || row_end > n || row_begin > row_end) return ORC_EINVAL;
  const int64_t e0 = off[row_begin], e1 = off[row_end];
  int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * (size_t)(e1 > e0 ? e1 - e0 : 1));
  for (int64_t t = row_begin; t < row_end; t++)
    for (int32_t e = off[t]; e < off[t + 1]; e++) rows[e - e0] = (int32_t)t;
  set_threads(nthreads);
  const int rc = tc_hash_edges(off, cols, rows, e0, e1, toff[n], L, toff, bucket_size,
                               max_len, layout, T, out);
  free(rows);
  return rc;
}

static int64_t merge_count(const int32_t* a, int64_t na, const int32_t* b, int64_t nb) {
  int64_t i = 0, j = 0, c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j])
      i++;
    else if (a[i] > b[j])
      j++;
    else {
      c++;
      i++;
      j++;
    }
  }
  return c;
}

uint64_t orc_tc_merge(const int32_t* off, const int32_t* cols, int64_t n, int nthreads) {
  uint64_t total = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int64_t u = 0; u < n; u++)
    for (int32_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = cols[e];
      total += (uint64_t)merge_count(cols + off[u], off[u + 1] - off[u],
                                     cols + off[v], off[v + 1] - off[v]);
    }
  return total;
}

uint64_t orc_tc_merge_rows(const int32_t* off, const int32_t* cols, int64_t row_begin,
                           int64_t row_end, int nthreads) {
  uint64_t total = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int64_t u = row_begin; u < row_end; u++)
    for (int32_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = cols[e];
      total += (uint64_t)merge_count(cols + off[u], off[u + 1] - off[u],
                                     cols + off[v], off[v + 1] - off[v]);
    }
  return total;
}

/* set_size/intersection.cu:128-181, one pair at a time: build the larger set's
 * table (normal hash; with one table per launch BN = L), probe the smaller. */
int orc_pairs_hash(const int64_t* a_off, const int32_t* a, const int64_t* b_off,
                   const int32_t* b, int64_t P, float load_factor,
                   int32_t bucket_size, int nthreads, int32_t* sizes) {
  int failed = 0;
  set_threads(nthreads);
#pragma omp parallel reduction(| : failed)
  {
    int32_t* tab = NULL;
    int64_t cap = 0;
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < P; i++) {
      const int32_t* x = a + a_off[i];
      const int32_t* y = b + b_off[i];
      int64_t nx = a_off[i + 1] - a_off[i], ny = b_off[i + 1] - b_off[i];
      if (nx > ny) {
        const int32_t* t = x;
        x = y;
        y = t;
        const int64_t tn = nx;
        nx = ny;
        ny = tn;
      }
      const int32_t len = orc_calculate_length((int32_t)ny, load_factor, bucket_size);
      if (len < 0) {
        failed = 1;
        sizes[i] = -1;
        continue;
      }
      const int64_t slots = (int64_t)len * bucket_size;
      if (slots > cap) {
        cap = slots;
        tab = (int32_t*)realloc(tab, sizeof(int32_t) * (size_t)cap);
      }
      for (int64_t s = 0; s < slots; s++) tab[s] = -1;
      int bad = 0;
      for (int64_t j = 0; j < ny && !bad; j++) {
        uint32_t bucket = orc_hash_normal((uint32_t)y[j], (uint32_t)len);
        int32_t k = 0;
        int64_t tries = 0;
        while (tab[bucket + (int64_t)k * len] != -1) {
          if (++tries >= slots) {
            bad = 1;
            break;
          }
          if (++k == bucket_size) {
            k = 0;
            bucket = (bucket + 1u) & (uint32_t)(len - 1);
          }
        }
        if (!bad) tab[bucket + (int64_t)k * len] = y[j];
      }
      int32_t hits = 0;
      for (int64_t j = 0; j < nx && !bad; j++) {
        const int r = orc_table_search(tab, len, orc_hash_normal((uint32_t)x[j], (uint32_t)len),
                                       len, bucket_size, x[j]);
        if (r < 0)
          bad = 1;
        else
          hits += r;
      }
      if (bad) {
        failed = 1;
        sizes[i] = -1;
      } else {
        sizes[i] = hits;
      }
    }
    free(tab);
  }
  return failed ? ORC_EOVERFLOW : ORC_OK;
}

void orc_pairs_merge(const int64_t* a_off, const int32_t* a,
                     const int64_t* b_off, const int32_t* b, int64_t P,
                     int nthreads, int32_t* sizes) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < P; i++)
    sizes[i] = (int32_t)merge_count(a + a_off[i], a_off[i + 1] - a_off[i],
                                    b + b_off[i], b_off[i + 1] - b_off[i]);
}

/* Exact Jaccard rule: J = r / U >= p / q  <=>  q*r >= p*U, with U = |u|+|v|-r.
 * U = 0 (two empty sets) is not flagged, as the reference's 0/0 = NaN compare
 * is false. */
static inline int jaccard_pass(int64_t r, int64_t su, int64_t sv, int32_t p, int32_t q) {
  const int64_t U = su + sv - r;
  return U > 0 && (int64_t)q * r >= (int64_t)p * U;
}

/* set-similarity-search/sss.cu:93-154 */
int orc_sss_verify_hash(const int32_t* off, const int32_t* cols, int64_t n,
                        const int32_t* L, const int64_t* toff,
                        int32_t bucket_size, int32_t max_len, int32_t layout,
                        const int32_t* T, const int32_t* cand_u,
                        const int32_t* cand_v, int64_t P, int32_t p, int32_t q,
                        int nthreads, int32_t* flags, int32_t* sizes) {
  const int64_t BN = toff[n];
  int failed = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(| : failed)
  for (int64_t e = 0; e < P; e++) {
    int32_t u = cand_u[e], v = cand_v[e];
    int32_t su = off[u + 1] - off[u], sv = off[v + 1] - off[v];
    if (su > sv) {
      const int32_t t = u;
      u = v;
      v = t;
      const int32_t ts = su;
      su = sv;
      sv = ts;
    }
    const int64_t r = probe_row(off, cols, u, v, L, toff, BN, bucket_size,
                                max_len, layout, T);
    if (r < 0) failed = 1;
    if (flags) flags[e] = r >= 0 && jaccard_pass(r, su, sv, p, q);
    if (sizes) sizes[e] = (int32_t)r;
  }
  return failed ? ORC_EOVERFLOW : ORC_OK;
}

/* sss.cu:130-133: float jaccard = r / (float)(|u| + |v| - r); jaccard >= t */
int orc_sss_float_rule(int32_t r, int32_t size_u, int32_t size_v, float threshold) {
  volatile float den = (float)(size_u + size_v - r);
  volatile float j = (float)r / den;
  return j >= threshold;
}

static int cmp_i64(const void* a, const void* b) {
  const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

int64_t orc_sss_bruteforce(const int32_t* off, const int32_t* cols, int64_t n,
                           int32_t p, int32_t q, int64_t* out, int64_t cap) {
  int64_t w = 0;
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = i + 1; j < n; j++) {
      const int64_t si = off[i + 1] - off[i], sj = off[j + 1] - off[j];
      const int64_t r = merge_count(cols + off[i], si, cols + off[j], sj);
      if (jaccard_pass(r, si, sj, p, q)) {
        if (w >= cap) return -1;
        out[w++] = (i << 32) | j;
      }
    }
  return w;
}

int64_t orc_sss_lookup_bruteforce(const int32_t* off, const int32_t* cols, int64_t n,
                                  const int32_t* qoff, const int32_t* qtok, int64_t nq,
                                  int32_t p, int32_t q, int64_t* out, int64_t cap) {
  int64_t w = 0;
  for (int64_t a = 0; a < nq; a++) {
    const int64_t sa = qoff[a + 1] - qoff[a];
    for (int64_t j = 0; j < n; j++) {
      const int64_t sj = off[j + 1] - off[j];
      const int64_t r = merge_count(qtok + qoff[a], sa, cols + off[j], sj);
      if (jaccard_pass(r, sa, sj, p, q)) {
        if (w >= cap) return -1;
        out[w++] = (a << 32) | j;
      }
    }
  }
  return w;
}

typedef struct {
  int64_t freq;
  int32_t tok;
} tokfreq;

static int cmp_tokfreq(const void* a, const void* b) {
  const tokfreq* x = (const tokfreq*)a;
  const tokfreq* y = (const tokfreq*)b;
  if (x->freq != y->freq) return (x->freq > y->freq) - (x->freq < y->freq);
  return (x->tok > y->tok) - (x->tok < y->tok);
}

int64_t orc_sss_prefix_join(const int32_t* off, const int32_t* cols, int64_t n,
                            int32_t p, int32_t q, int nthreads, int64_t* out,
                            int64_t cap, int64_t* num_candidates) {
  const int64_t E = off[n];
  int32_t maxtok = -1;
  for (int64_t e = 0; e < E; e++)
    if (cols[e] > maxtok) maxtok = cols[e];
  const int64_t V = (int64_t)maxtok + 1;
  /* global order: ascending frequency, ties by token id */
  tokfreq* tf = (tokfreq*)calloc((size_t)(V > 0 ? V : 1), sizeof(tokfreq));
  for (int64_t t = 0; t < V; t++) tf[t].tok = (int32_t)t;
  for (int64_t e = 0; e < E; e++) tf[cols[e]].freq++;
  qsort(tf, (size_t)V, sizeof(tokfreq), cmp_tokfreq);
  int32_t* rank = (int32_t*)malloc(sizeof(int32_t) * (size_t)(V > 0 ? V : 1));
  for (int64_t r = 0; r < V; r++) rank[tf[r].tok] = (int32_t)r;
  free(tf);
  /* each row in rank order; prefix length s - ceil(p*s/q) + 1 */
  int32_t* ranked = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
  int32_t* plen = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int64_t* pcount = (int64_t*)calloc((size_t)(V + 1), sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) {
    const int32_t s = off[i + 1] - off[i];
    for (int32_t e = off[i]; e < off[i + 1]; e++) ranked[e] = rank[cols[e]];
    qsort(ranked + off[i], (size_t)s, sizeof(int32_t), cmp_i32);
    int64_t pl = (int64_t)s - ((int64_t)p * s + q - 1) / q + 1;
    if (pl > s) pl = s;
    if (pl < 0) pl = 0;
    plen[i] = (int32_t)pl;
    for (int32_t k = 0; k < pl; k++) pcount[ranked[off[i] + k] + 1]++;
  }
  for (int64_t r = 0; r < V; r++) pcount[r + 1] += pcount[r];
  int32_t* plist = (int32_t*)malloc(sizeof(int32_t) * (size_t)(pcount[V] > 0 ? pcount[V] : 1));
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(V > 0 ? V : 1));
  for (int64_t r = 0; r < V; r++) fill[r] = pcount[r];
  for (int64_t i = 0; i < n; i++)
    for (int32_t k = 0; k < plen[i]; k++) plist[fill[ranked[off[i] + k]]++] = (int32_t)i;
  free(fill);
  /* candidate pairs sharing a prefix token and passing the size filter */
  int64_t ncand = 0, ccap = 1024;
  int64_t* cand = (int64_t*)malloc(sizeof(int64_t) * (size_t)ccap);
  for (int64_t r = 0; r < V; r++)
    for (int64_t x = pcount[r]; x < pcount[r + 1]; x++)
      for (int64_t y = x + 1; y < pcount[r + 1]; y++) {
        int64_t i = plist[x], j = plist[y];
        if (i > j) {
          const int64_t t = i;
          i = j;
          j = t;
        }
        int64_t si = off[i + 1] - off[i], sj = off[j + 1] - off[j];
        const int64_t lo = si < sj ? si : sj, hi = si < sj ? sj : si;
        if ((int64_t)p * hi > (int64_t)q * lo) continue;
        if (ncand == ccap) {
          ccap *= 2;
          cand = (int64_t*)realloc(cand, sizeof(int64_t) * (size_t)ccap);
        }
        cand[ncand++] = (i << 32) | j;
      }
  qsort(cand, (size_t)ncand, sizeof(int64_t), cmp_i64);
  int64_t u = 0;
  for (int64_t k = 0; k < ncand; k++)
    if (k == 0 || cand[k] != cand[k - 1]) cand[u++] = cand[k];
  if (num_candidates) *num_candidates = u;
  char* pass = (char*)malloc((size_t)(u > 0 ? u : 1));
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t k = 0; k < u; k++) {
    const int64_t i = cand[k] >> 32, j = cand[k] & 0xffffffffLL;
    const int64_t si = off[i + 1] - off[i], sj = off[j + 1] - off[j];
    const int64_t r = merge_count(cols + off[i], si, cols + off[j], sj);
    pass[k] = (char)jaccard_pass(r, si, sj, p, q);
  }
  int64_t w = 0;
  for (int64_t k = 0; k < u; k++)
    if (pass[k]) {
      if (w >= cap) {
        w = -1;
        break;
      }
      out[w++] = cand[k];
    }
  free(pass);
  free(cand);
  free(plist);
  free(pcount);
  free(plen);
  free(ranked);
  free(rank);
  return w;
}

/* information-retrieval/ir.cu:114-165.  Table length of a term is
 * toff[term+1] - toff[term] (ir.cu:133-134); BN = toff[num_terms]. */
int orc_ir_hash(const int32_t* idx_off, const int32_t* docs, int64_t num_terms,
                const int64_t* toff, int32_t bucket_size, int32_t max_len,
                int32_t layout, const int32_t* T, const int32_t* query,
                const int32_t* qoff, int64_t Q, const int64_t* roff,
                int nthreads, int32_t* result, int32_t* count) {
  const int64_t BN = toff[num_terms];
  int failed = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 2) reduction(| : failed)
  for (int64_t qi = 0; qi < Q; qi++) {
    const int32_t qs = qoff[qi], qe = qoff[qi + 1];
    if (qe <= qs) {
      count[qi] = 0;
      continue;
    }
    const int32_t t0 = query[qs];
    const int32_t n0 = idx_off[t0 + 1] - idx_off[t0];
    if (roff[qi + 1] - roff[qi] != n0) {
      failed = 1;
      count[qi] = -1;
      continue;
    }
    int32_t* res = result + roff[qi];
    memcpy(res, docs + idx_off[t0], sizeof(int32_t) * (size_t)n0);
    int32_t cur = n0;
    for (int32_t i = qs + 1; i < qe; i++) {
      const int32_t term = query[i];
      const int32_t len = (int32_t)(toff[term + 1] - toff[term]);
      const int32_t* Tt = T + toff[term];
      int32_t w = 0;
      for (int32_t j = 0; j < cur; j++) {
        const int32_t key = res[j];
        const int r = orc_table_search(Tt, BN, home_bucket(layout, key, len, max_len),
                                       len, bucket_size, key);
        if (r < 0) failed = 1;
        if (r == 1) res[w++] = key;
      }
      cur = w;
    }
    count[qi] = cur;
  }
  return failed ? ORC_EOVERFLOW : ORC_OK;
}

void orc_ir_merge(const int32_t* idx_off, const int32_t* docs,
                  const int32_t* query, const int32_t* qoff, int64_t Q,
                  const int64_t* roff, int nthreads, int32_t* result,
                  int32_t* count) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 2)
  for (int64_t qi = 0; qi < Q; qi++) {
    const int32_t qs = qoff[qi], qe = qoff[qi + 1];
    if (qe <= qs) {
      count[qi] = 0;
      continue;
    }
    const int32_t t0 = query[qs];
    const int32_t n0 = idx_off[t0 + 1] - idx_off[t0];
    int32_t* res = result + roff[qi];
    memcpy(res, docs + idx_off[t0], sizeof(int32_t) * (size_t)n0);
    int32_t cur = n0;
    for (int32_t i = qs + 1; i < qe; i++) {
      const int32_t term = query[i];
      const int32_t* lst = docs + idx_off[term];
      const int32_t nl = idx_off[term + 1] - idx_off[term];
      int32_t a = 0, b = 0, w = 0;
      while (a < cur && b < nl) {
        if (res[a] < lst[b])
          a++;
        else if (res[a] > lst[b])
          b++;
        else {
          res[w++] = res[a];
          a++;
          b++;
        }
      }
      cur = w;
    }
    count[qi] = cur;
  }
}
