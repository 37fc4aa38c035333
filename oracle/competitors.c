/*
 * competitors.c -- CPU restatement of the competitor tables of RESET's
 * baseline directory: cuCollections-style linear probing, cuckoo hashing
 * with a stash, hopscotch hashing and the per-set sparse ("roaring") bitmap,
 * and of their TC / SSS consumers.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE: loaded only by tests/ as the checker.
 *
 * Every table is one region per set (flat storage, 64-bit region offsets;
 * the reference keeps hopscotch offsets in int, which overflows past 2^31
 * slots).  The cuckoo, hopscotch and roaring builds are serial per set in
 * the reference, so their final contents are deterministic and the GPU
 * build is compared with these arrays slot for slot, failed inserts
 * included.  The linear-probing build inserts in parallel (placement depends
 * on the schedule); it is checked through orc_comp_linear_check instead.
 */
#include "competitors.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define COMP_EMPTY 0x7fffffff /* std::numeric_limits<int>::max() */

/* ------------------------------------------------------------ linear probing */
/* MurmurHash3's 32-bit finalizer (Appleby, public domain), seed 0: the hash
 * cuco::murmurhash3_fmix_32 applies (baseline/common/cuco.cuh:25). */
uint32_t orc_fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

/* baseline/common/cuco.cu:69-80: capacity = next power of two of
 * (size_t)(degree / load_factor) (float divide, truncated), at least 1. */
int64_t orc_comp_linear_geometry(const int32_t* off, int64_t n, float lf, int64_t* toff) {
  toff[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t d = off[i + 1] - off[i];
    volatile float q = (float)d / lf;
    uint64_t req = (uint64_t)q;
    uint64_t cap = 1;
    while (cap < req) cap <<= 1;
    toff[i + 1] = toff[i] + (int64_t)cap;
  }
  return toff[n];
}

/* baseline/common/cuco.cu:35-66, one thread per set in row order (the
 * reference runs 64 threads per set with atomicCAS; any schedule yields a
 * valid linear-probing table, which orc_comp_linear_check verifies).  A key
 * that finds no slot is dropped; returns how many were (the same for every
 * schedule: a region keeps min(capacity, distinct keys) of them). */
int64_t orc_comp_linear_build(const int32_t* off, const int32_t* el, int64_t n, const int64_t* toff,
                              int32_t* slots) {
  int64_t failed = 0;
  for (int64_t t = 0; t < toff[n]; ++t) slots[t] = COMP_EMPTY;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t cap = toff[i + 1] - toff[i];
    const uint32_t mask = (uint32_t)(cap - 1);
    int32_t* s = slots + toff[i];
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      const int32_t key = el[e];
      const uint32_t h = orc_fmix32((uint32_t)key);
      int64_t p = 0;
      for (; p < cap; ++p) {
        int32_t* a = s + ((h + (uint32_t)p) & mask);
        if (*a == COMP_EMPTY) {
          *a = key;
          break;
        }
        if (*a == key) break;
      }
      failed += p == cap;
    }
  }
  return failed;
}

/* cuco::static_set_ref::contains with linear_probing<1>: from hash % cap,
 * step 1, until the key (found) or an EMPTY slot (absent); cap slots at most. */
static int linear_contains(const int32_t* slots, const int64_t* toff, int32_t set, int32_t key) {
  const int64_t cap = toff[set + 1] - toff[set];
  const uint32_t mask = (uint32_t)(cap - 1);
  const int32_t* s = slots + toff[set];
  const uint32_t h = orc_fmix32((uint32_t)key);
  for (int64_t p = 0; p < cap; ++p) {
    const int32_t v = s[(h + (uint32_t)p) & mask];
    if (v == key) return 1;
    if (v == COMP_EMPTY) return 0;
  }
  return 0;
}

/* Invariants of a linear-probing table built from the rows: every slot holds
 * EMPTY or an element of its row, each distinct element is stored exactly
 * once, and no EMPTY slot lies between an element's home slot and its slot.
 * Returns the number of violations. */
int64_t orc_comp_linear_check(const int32_t* off, const int32_t* el, int64_t n,
                              const int64_t* toff, const int32_t* slots) {
  int64_t bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t cap = toff[i + 1] - toff[i];
    const uint32_t mask = (uint32_t)(cap - 1);
    const int32_t* s = slots + toff[i];
    int64_t stored = 0;
    for (int64_t p = 0; p < cap; ++p) {
      const int32_t v = s[p];
      if (v == COMP_EMPTY) continue;
      ++stored;
      int in_row = 0;
      for (int32_t e = off[i]; e < off[i + 1] && !in_row; ++e) in_row = el[e] == v;
      if (!in_row) ++bad;
      const uint32_t h = orc_fmix32((uint32_t)v) & mask;
      for (uint32_t q = h; q != (uint32_t)p; q = (q + 1u) & mask)
        if (s[q] == COMP_EMPTY) {
          ++bad;
          break;
        }
    }
    int64_t distinct = 0;
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      int seen = 0;
      for (int32_t f = off[i]; f < e && !seen; ++f) seen = el[f] == el[e];
      distinct += !seen;
      if (!linear_contains(slots, toff, (int32_t)i, el[e])) ++bad;
    }
    if (stored != distinct) ++bad;
  }
  return bad;
}

/* ------------------------------------------------------------------ cuckoo */
/* baseline/common/utils.cuh:123-134 */
static int32_t std_hash(int32_t key, int32_t set, int32_t cap) {
  return (key ^ (key >> 8) ^ set) & (cap - 1);
}
static int32_t std_hash2(int32_t key, int32_t set, int32_t cap) {
  return (int32_t)(((uint32_t)key ^ (uint32_t)(key >> 16) ^ ((uint32_t)set << 1) ^ 0xAAAAAAAAu) &
                   (uint32_t)(cap - 1));
}

/* baseline/common/cuckoo_hash.cu:150-170: capacity = smallest power of two
 * >= (int)ceil(degree / load_factor) (float divide), at least 1. */
int64_t orc_comp_cuckoo_geometry(const int32_t* off, int64_t n, float lf, int64_t* toff) {
  toff[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t d = off[i + 1] - off[i];
    volatile float q = (float)d / lf;
    const int32_t req = (int32_t)ceilf(q);
    int64_t cap = 1;
    while (cap < req) cap <<= 1;
    toff[i + 1] = toff[i] + cap;
  }
  return toff[n];
}

/* baseline/common/cuckoo_hash.cu:40-140, one set at a time, its row in
 * order: skip a key already at one of its positions; else take the first
 * EMPTY of pos1, pos2 (pos3 = pos2); else kick, starting at the position
 * key % 3 picks, moving each evicted key to a position of its own other than
 * the one it left (the choice rule of the reference), for at most
 * ORC_CUCKOO_KICKS evictions; the key left over goes to the set's 16-entry
 * stash, or is lost (counted as failed).  Returns the failed count. */
int64_t orc_comp_cuckoo_build(const int32_t* off, const int32_t* el, int64_t n,
                              const int64_t* toff, int32_t* table, int32_t* stash_count,
                              int32_t* stash) {
  int64_t failed = 0;
  for (int64_t t = 0; t < toff[n]; ++t) table[t] = COMP_EMPTY;
  for (int64_t t = 0; t < n * ORC_CUCKOO_STASH; ++t) stash[t] = COMP_EMPTY;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t set = (int32_t)i;
    const int32_t cap = (int32_t)(toff[i + 1] - toff[i]);
    int32_t* T = table + toff[i];
    int32_t* st = stash + i * ORC_CUCKOO_STASH;
    int32_t ns = 0;
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      const int32_t key = el[e];
      const int32_t p1 = std_hash(key, set, cap), p2 = std_hash2(key, set, cap), p3 = p2;
      if (T[p1] == key || T[p2] == key || T[p3] == key) continue;
      if (T[p1] == COMP_EMPTY) {
        T[p1] = key;
        continue;
      }
      if (T[p2] == COMP_EMPTY) {
        T[p2] = key;
        continue;
      }
      int32_t cur_key = key;
      const int32_t r = key % 3;
      int32_t cur = r == 0 ? p1 : (r == 1 ? p2 : p3);
      int done = 0;
      for (int it = 0; it < ORC_CUCKOO_KICKS; ++it) {
        const int32_t old = T[cur];
        T[cur] = cur_key;
        if (old == COMP_EMPTY) {
          done = 1;
          break;
        }
        const int32_t o1 = std_hash(old, set, cap), o2 = std_hash2(old, set, cap), o3 = o2;
        int32_t cand[3];
        int cnt = 0;
        if (o1 != cur) cand[cnt++] = o1;
        if (o2 != cur) cand[cnt++] = o2;
        if (o3 != cur) cand[cnt++] = o3;
        cur_key = old;
        if (cnt == 1)
          cur = cand[0];
        else if (cnt == 2)
          cur = (old & 1) ? cand[0] : cand[1];
        else
          cur = (old % 3 == 0) ? o1 : (old % 3 == 1 ? o2 : o3);
      }
      if (!done) {
        if (ns < ORC_CUCKOO_STASH)
          st[ns++] = cur_key;
        else
          ++failed;
      }
    }
    stash_count[i] = ns;
  }
  return failed;
}

/* baseline/common/cuckoo_hash.cuh:86-112 */
static int cuckoo_contains(const int32_t* table, const int64_t* toff, const int32_t* stash_count,
                           const int32_t* stash, int32_t set, int32_t key) {
  const int32_t cap = (int32_t)(toff[set + 1] - toff[set]);
  const int32_t* T = table + toff[set];
  if (T[std_hash(key, set, cap)] == key || T[std_hash2(key, set, cap)] == key) return 1;
  const int32_t* st = stash + (int64_t)set * ORC_CUCKOO_STASH;
  for (int32_t k = 0; k < stash_count[set]; ++k)
    if (st[k] == key) return 1;
  return 0;
}

/* --------------------------------------------------------------- hopscotch */
/* baseline/common/hopscotch_hash.cu:183-202: load factor min(lf, 0.1);
 * capacity = smallest power of two >= degree / lf (double divide), at least
 * H; the region is capacity + H slots. */
int64_t orc_comp_hopscotch_geometry(const int32_t* off, int64_t n, float lf, int64_t* toff) {
  const float a = lf < 0.1f ? lf : 0.1f;
  toff[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double req = (double)(off[i + 1] - off[i]) / (double)a;
    int64_t cap = 1;
    while ((double)cap < req) cap <<= 1;
    if (cap < ORC_HOP_H) cap = ORC_HOP_H;
    toff[i + 1] = toff[i] + cap + ORC_HOP_H;
  }
  return toff[n];
}

/* baseline/common/hopscotch_hash.cu:27-166, one set at a time, its row in
 * order (no duplicate check, as in the reference).  A key takes the first
 * EMPTY slot of its neighbourhood [home, home+H); else the first EMPTY slot
 * at or after home+H is moved back, each step swapping in the element at the
 * highest neighbourhood offset of the latest home (scanning homes downwards
 * from the free slot) that lies before the free slot.  A moved element's old
 * slot keeps its value until it is overwritten.  Returns the failed count. */
int64_t orc_comp_hopscotch_build(const int32_t* off, const int32_t* el, int64_t n,
                                 const int64_t* toff, int32_t* table, uint32_t* bitmap) {
  int64_t failed = 0;
  for (int64_t t = 0; t < toff[n]; ++t) {
    table[t] = COMP_EMPTY;
    bitmap[t] = 0u;
  }
  for (int64_t i = 0; i < n; ++i) {
    const int32_t total = (int32_t)(toff[i + 1] - toff[i]);
    const int32_t cap = total - ORC_HOP_H;
    int32_t* T = table + toff[i];
    uint32_t* bm = bitmap + toff[i];
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      const int32_t key = el[e];
      const int32_t home = std_hash(key, (int32_t)i, cap);
      int placed = 0;
      for (int j = 0; j < ORC_HOP_H; ++j) {
        const int32_t pos = home + j;
        if (pos < total && T[pos] == COMP_EMPTY) {
          T[pos] = key;
          bm[home] |= 1u << j;
          placed = 1;
          break;
        }
      }
      if (placed) continue;
      int32_t empty = home + ORC_HOP_H;
      while (empty < total && T[empty] != COMP_EMPTY) ++empty;
      if (empty >= total) {
        ++failed;
        continue;
      }
      while (empty >= home + ORC_HOP_H) {
        const int32_t cur = empty;
        const int32_t lo = cur - ORC_HOP_H + 1 > 0 ? cur - ORC_HOP_H + 1 : 0;
        int found = 0, fj = -1, fh = -1;
        for (int32_t h = cur; h >= lo && !found; --h) {
          const uint32_t b = bm[h];
          for (int j = ORC_HOP_H - 1; j >= 0; --j)
            if ((b >> j) & 1u) {
              if (h + j < cur) {
                fj = j;
                fh = h;
                found = 1;
              }
              if (found) break;
            }
        }
        if (!found) break;
        const int32_t sp = fh + fj;
        T[cur] = T[sp];
        bm[fh] = (bm[fh] & ~(1u << fj)) | (1u << (cur - fh));
        empty = sp;
      }
      if (empty >= home && empty < home + ORC_HOP_H) {
        T[empty] = key;
        bm[home] |= 1u << (empty - home);
      } else {
        ++failed;
      }
    }
  }
  return failed;
}

/* baseline/common/hopscotch_hash.cuh:82-104 */
static int hopscotch_contains(const int32_t* table, const uint32_t* bitmap, const int64_t* toff,
                              int32_t set, int32_t key) {
  const int32_t total = (int32_t)(toff[set + 1] - toff[set]);
  const int32_t home = std_hash(key, set, total - ORC_HOP_H);
  uint32_t b = bitmap[toff[set] + home];
  while (b) {
    const int j = __builtin_ctz(b);
    b &= b - 1u;
    if (home + j < total && table[toff[set] + home + j] == key) return 1;
  }
  return 0;
}

/* ----------------------------------------------------------------- roaring */
/* baseline/common/roaring_bitmap.cu:16-36: ids = {0..n-1} u elements, sorted
 * and unique; remap[id] = its rank.  remap has max_id + 1 entries; ids that
 * are not in the collection get -1 (the reference leaves 0 there).  Returns
 * the number of distinct ids, or -1 for a negative element. */
int64_t orc_comp_roaring_remap(const int32_t* off, const int32_t* el, int64_t n, int32_t* remap,
                               int64_t remap_len) {
  const int64_t E = off[n];
  for (int64_t k = 0; k < remap_len; ++k) remap[k] = 0;
  for (int64_t k = 0; k < n && k < remap_len; ++k) remap[k] = 1;
  for (int64_t e = 0; e < E; ++e) {
    if (el[e] < 0 || el[e] >= remap_len) return -1;
    remap[el[e]] = 1;
  }
  int64_t rank = 0;
  for (int64_t k = 0; k < remap_len; ++k) remap[k] = remap[k] ? (int32_t)rank++ : -1;
  return rank;
}

/* baseline/common/roaring_bitmap.cu:38-66: per set, the 64-bit words
 * (remapped id >> 6) its elements touch, ascending, each with the bits of
 * its elements.  node_start[n+1] indexes word_index / word_data.  Returns the
 * total word count (the caller sizes the outputs by orc_comp_roaring_words). */
int64_t orc_comp_roaring_words(const int32_t* off, const int32_t* el, int64_t n,
                               const int32_t* remap) {
  int64_t words = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t last = -1;
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      const int32_t w = remap[el[e]] >> 6;
      /* rows are ascending here, so a set's words come in order */
      if (e == off[i] || w != last) ++words;
      last = w;
    }
  }
  return words;
}

int64_t orc_comp_roaring_build(const int32_t* off, const int32_t* el, int64_t n,
                               const int32_t* remap, int32_t* node_start, int32_t* word_index,
                               uint64_t* word_data) {
  int64_t w = 0;
  node_start[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t last = -1;
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      const int32_t r = remap[el[e]];
      if (e == off[i] || (r >> 6) != last) {
        word_index[w] = r >> 6;
        word_data[w] = 0;
        ++w;
      }
      word_data[w - 1] |= 1ull << (r & 63);
      last = r >> 6;
    }
    node_start[i + 1] = (int32_t)w;
  }
  return w;
}

/* baseline/common/roaring_bitmap.cuh:56-86: binary search of the set's words */
static int roaring_contains(const orc_comp* c, int32_t set, int32_t key) {
  if (key < 0 || key >= c->remap_len) return 0;
  const int32_t r = c->remap[key];
  if (r < 0) return 0;
  const int32_t wi = r >> 6;
  int32_t lo = c->node_start[set], hi = c->node_start[set + 1] - 1;
  while (lo <= hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    const int32_t m = c->word_index[mid];
    if (m == wi) return (c->word_data[mid] >> (r & 63)) & 1ull ? 1 : 0;
    if (m < wi)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return 0;
}

/* ---------------------------------------------------------------- consumers */
int orc_comp_contains(const orc_comp* c, int32_t set, int32_t key) {
  switch (c->kind) {
    case ORC_COMP_LINEAR: return linear_contains(c->slots, c->toff, set, key);
    case ORC_COMP_CUCKOO:
      return cuckoo_contains(c->slots, c->toff, c->stash_count, c->stash, set, key);
    case ORC_COMP_HOPSCOTCH: return hopscotch_contains(c->slots, c->bitmap, c->toff, set, key);
    default: return roaring_contains(c, set, key);
  }
}

/* |N(s) n N(g)| through g's table, s the endpoint with the smaller row (the
 * CSR source on ties), as baseline/tc/tc.cu:35-108 and sss/sss_cuckoo.cu:
 * 34-97 pair their endpoints. */
static int64_t probe_count(const orc_comp* c, const int32_t* off, const int32_t* el, int32_t u,
                           int32_t v) {
  int32_t s = u, g = v;
  if (off[u + 1] - off[u] > off[v + 1] - off[v]) {
    s = v;
    g = u;
  }
  int64_t r = 0;
  for (int32_t e = off[s]; e < off[s + 1]; ++e) r += orc_comp_contains(c, g, el[e]);
  return r;
}

uint64_t orc_comp_tc(const orc_comp* c, const int32_t* off, const int32_t* cols, int64_t n) {
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int64_t u = 0; u < n; ++u)
    for (int32_t e = off[u]; e < off[u + 1]; ++e)
      total += (uint64_t)probe_count(c, off, cols, (int32_t)u, cols[e]);
  return total;
}

/* flag = q*r >= p*(|u| + |v| - r): J >= p/q decided in integers */
void orc_comp_sss(const orc_comp* c, const int32_t* off, const int32_t* el, const int32_t* cu,
                  const int32_t* cv, int64_t P, int32_t p, int32_t q, int32_t* flags,
                  int32_t* sizes) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t k = 0; k < P; ++k) {
    const int64_t r = probe_count(c, off, el, cu[k], cv[k]);
    const int64_t uni = (int64_t)(off[cu[k] + 1] - off[cu[k]]) + (off[cv[k] + 1] - off[cv[k]]) - r;
    sizes[k] = (int32_t)r;
    flags[k] = (int64_t)q * r >= (int64_t)p * uni ? 1 : 0;
  }
}
