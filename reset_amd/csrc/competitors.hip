// competitors.hip -- the competitor tables of the reference's baseline
// directory on gfx950, with their TC and SSS consumers.
//
// Four per-set table kinds, each one flat region per set:
//  * linear: cuCollections-style open addressing (baseline/common/cuco.cu,
//    cuco.cuh): 1-slot buckets, MurmurHash3 fmix32, linear probing,
//    EMPTY = INT_MAX; one wave inserts a set with CAS (cuco.cu:35-66);
//  * cuckoo: two hashes (pos3 = pos2), kick chains of at most 6000
//    evictions, a 16-entry stash per set (baseline/common/cuckoo_hash.cu);
//  * hopscotch: H = 32 neighbourhoods with a 32-bit map per home slot and
//    backward displacement (baseline/common/hopscotch_hash.cu);
//  * roaring: ids remapped to their rank in {0..n-1} u elements, each set a
//    sorted list of 64-bit words (baseline/common/roaring_bitmap.cu).
// The cuckoo and hopscotch inserts are serial per set in the reference
// (their final contents depend on the order), so here too one lane builds
// one set, in row order, and the tables come out identical to the
// reference's.  Region offsets are 64-bit (the reference's hopscotch int
// offsets overflow past 2^31 slots).  Consumers: TC over the CSR the tables
// were built from and SSS verification of candidate pairs, a wave per edge
// or pair with the smaller set's lanes probing the larger set's table
// (baseline/tc/tc.cu:35-108, tc_cuco.cu, sss/sss_cuckoo.cu:19-107), and IR
// conjunctive queries over tables of an inverted index (ir/ir_cuco.cu).
#include "common.hpp"

namespace rs {

constexpr int32_t kCompEmpty = 0x7fffffff;
constexpr int32_t kCuckooStash = 16;
constexpr int32_t kCuckooKicks = 6000;
constexpr int32_t kHopH = 32;
constexpr int kCompBlock = 256;

// ---------------------------------------------------------------- hashes
// MurmurHash3 32-bit finalizer (cuco::murmurhash3_fmix_32, seed 0).
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
// baseline/common/utils.cuh:123-134 (capacity a power of two)
__device__ __forceinline__ int32_t std_hash(int32_t key, int32_t set, int32_t cap) {
  return (key ^ (key >> 8) ^ set) & (cap - 1);
}
__device__ __forceinline__ int32_t std_hash2(int32_t key, int32_t set, int32_t cap) {
  return (int32_t)(((uint32_t)key ^ (uint32_t)(key >> 16) ^ ((uint32_t)set << 1) ^ 0xAAAAAAAAu) &
                   (uint32_t)(cap - 1));
}

// ---------------------------------------------------------------- geometry
// Region size of set i (cuco.cu:69-80, cuckoo_hash.cu:150-170,
// hopscotch_hash.cu:183-202).  A set whose degree / load factor exceeds
// 2^30 slots is flagged (bad input) and given one slot.
__global__ void comp_capacity_kernel(const int32_t* __restrict__ off, int64_t n, float lf, int32_t kind,
                                     int64_t* __restrict__ cap, int32_t* status) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = off[i + 1] - off[i];
    int64_t c = 1;
    if (!((double)d / (double)(lf < 0.1f || kind != RS_COMP_HOPSCOTCH ? lf : 0.1f) <= 1073741824.0)) {
      bad = true;
      cap[i] = 1;
      continue;
    }
    if (kind == RS_COMP_LINEAR) {
      const uint64_t req = (uint64_t)((float)d / lf);
      while ((uint64_t)c < req) c <<= 1;
    } else if (kind == RS_COMP_CUCKOO) {
      const int32_t req = (int32_t)ceilf((float)d / lf);
      while (c < req) c <<= 1;
    } else {
      const float a = lf < 0.1f ? lf : 0.1f;
      const double req = (double)d / (double)a;
      while ((double)c < req) c <<= 1;
      if (c < kHopH) c = kHopH;
      c += kHopH;
    }
    cap[i] = c;
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

template <typename T>
__global__ void comp_fill_kernel(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------- builds
// linear: a wave per set, lanes over its row, CAS along the probe sequence
// (cuco.cu:35-66); a key already present ends its walk.  A key that finds
// no slot (capacity below the degree, load factor > 1) is dropped and
// counted as failed.
__global__ void __launch_bounds__(kCompBlock) comp_linear_insert_kernel(
    const int32_t* __restrict__ off, const int32_t* __restrict__ el, int64_t n,
    const int64_t* __restrict__ toff, int32_t* slots, unsigned long long* failed) {
  const int lane = __lane_id();
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  uint32_t lost = 0;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; i < n; i += waves) {
    const int64_t base = toff[i];
    const int64_t cap = toff[i + 1] - base;
    const uint32_t mask = (uint32_t)(cap - 1);
    for (int32_t e = off[i] + lane; e < off[i + 1]; e += kWave) {
      const int32_t key = el[e];
      const uint32_t h = fmix32((uint32_t)key);
      int64_t p = 0;
      for (; p < cap; ++p) {
        const int32_t old = atomicCAS(slots + base + ((h + (uint32_t)p) & mask), kCompEmpty, key);
        if (old == kCompEmpty || old == key) break;
      }
      lost += p == cap;  // no free slot: the key is dropped, as in the reference
    }
  }
  lost = wave_sum(lost);
  if (lane == 0 && lost) atomicAdd(failed, (unsigned long long)lost);
}

// cuckoo: one lane per set, its row in order (cuckoo_hash.cu:40-140).
__global__ void __launch_bounds__(kCompBlock) comp_cuckoo_insert_kernel(
    const int32_t* __restrict__ off, const int32_t* __restrict__ el, int64_t n,
    const int64_t* __restrict__ toff, int32_t* table, int32_t* stash_count, int32_t* stash,
    unsigned long long* failed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t set = (int32_t)i;
    const int32_t cap = (int32_t)(toff[i + 1] - toff[i]);
    int32_t* T = table + toff[i];
    int32_t* st = stash + i * kCuckooStash;
    int32_t ns = 0;
    uint32_t lost = 0;
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      const int32_t key = el[e];
      const int32_t p1 = std_hash(key, set, cap), p2 = std_hash2(key, set, cap);
      if (T[p1] == key || T[p2] == key) continue;
      if (T[p1] == kCompEmpty) {
        T[p1] = key;
        continue;
      }
      if (T[p2] == kCompEmpty) {
        T[p2] = key;
        continue;
      }
      int32_t cur_key = key;
      const int32_t r = key % 3;
      int32_t cur = r == 0 ? p1 : p2;  // (pos3 = pos2)
      bool done = false;
      for (int32_t it = 0; it < kCuckooKicks; ++it) {
        const int32_t old = T[cur];
        T[cur] = cur_key;
        if (old == kCompEmpty) {
          done = true;
          break;
        }
        const int32_t o1 = std_hash(old, set, cap), o2 = std_hash2(old, set, cap);
        // the reference lists the evicted key's positions other than the
        // slot it left, in the order o1, o2, o3 (= o2): one candidate (o1)
        // is taken; two (o2 twice) give o2; three, or none, are chosen by
        // old % 3 (o1 for 0, else o2)
        const bool c1 = o1 != cur, c2 = o2 != cur;
        cur_key = old;
        if (c1 && !c2)
          cur = o1;
        else if (!c1 && c2)
          cur = o2;
        else
          cur = (old % 3 == 0) ? o1 : o2;
      }
      if (!done) {
        if (ns < kCuckooStash)
          st[ns++] = cur_key;
        else
          ++lost;
      }
    }
    stash_count[i] = ns;
    if (lost) atomicAdd(failed, (unsigned long long)lost);
  }
}

// hopscotch: one lane per set, its row in order (hopscotch_hash.cu:27-166).
__global__ void __launch_bounds__(kCompBlock) comp_hopscotch_insert_kernel(
    const int32_t* __restrict__ off, const int32_t* __restrict__ el, int64_t n,
    const int64_t* __restrict__ toff, int32_t* table, uint32_t* bitmap,
    unsigned long long* failed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t total = (int32_t)(toff[i + 1] - toff[i]);
    const int32_t cap = total - kHopH;
    int32_t* T = table + toff[i];
    uint32_t* bm = bitmap + toff[i];
    uint32_t lost = 0;
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      const int32_t key = el[e];
      const int32_t home = std_hash(key, (int32_t)i, cap);
      bool placed = false;
      for (int32_t j = 0; j < kHopH; ++j) {
        const int32_t pos = home + j;
        if (pos < total && T[pos] == kCompEmpty) {
          T[pos] = key;
          bm[home] |= 1u << j;
          placed = true;
          break;
        }
      }
      if (placed) continue;
      int32_t empty = home + kHopH;
      while (empty < total && T[empty] != kCompEmpty) ++empty;
      if (empty >= total) {
        ++lost;
        continue;
      }
      while (empty >= home + kHopH) {
        const int32_t cur = empty;
        const int32_t lo = cur - kHopH + 1 > 0 ? cur - kHopH + 1 : 0;
        int32_t fh = -1, fj = -1;
        // latest home first; within it the highest offset before `cur`
        for (int32_t h = cur; h >= lo && fh < 0; --h) {
          const uint32_t b = bm[h];
          for (int32_t j = kHopH - 1; j >= 0; --j)
            if (((b >> j) & 1u) && h + j < cur) {
              fh = h;
              fj = j;
              break;
            }
        }
        if (fh < 0) break;
        const int32_t sp = fh + fj;
        T[cur] = T[sp];
        bm[fh] = (bm[fh] & ~(1u << fj)) | (1u << (cur - fh));
        empty = sp;
      }
      if (empty >= home && empty < home + kHopH) {
        T[empty] = key;
        bm[home] |= 1u << (empty - home);
      } else {
        ++lost;
      }
    }
    if (lost) atomicAdd(failed, (unsigned long long)lost);
  }
}

// roaring: ids < n and every element are flagged; remap = rank of the id.
__global__ void comp_max_kernel(const int32_t* __restrict__ el, int64_t E, int32_t* out,
                                int32_t* status) {
  int32_t m = -1;
  bool bad = false;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    m = max(m, el[e]);
    bad |= el[e] < 0;
  }
  for (int d = kWave / 2; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, kWave));
  if (__lane_id() == 0) atomicMax(out, m);
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

__global__ void comp_flag_ids_kernel(const int32_t* __restrict__ el, int64_t E, int64_t n,
                                     int32_t* __restrict__ flag, int64_t len) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < E + n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t id = k < n ? k : el[k - n];
    if (id < len) flag[id] = 1;
  }
}

__global__ void comp_remap_kernel(const int32_t* __restrict__ flag, const int64_t* __restrict__ rank,
                                  int64_t len, int32_t* __restrict__ remap) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < len;
       k += (int64_t)gridDim.x * blockDim.x)
    remap[k] = flag[k] ? (int32_t)rank[k] : -1;
}

// f[e] = 1 where e opens a word of its set (first element of the row, or a
// remapped id in a later 64-id word than its predecessor's); rows must be
// ascending (an element below its predecessor is flagged as bad input).
__global__ void comp_word_starts_kernel(const int32_t* __restrict__ off, const int32_t* __restrict__ rows,
                                        const int32_t* __restrict__ el, const int32_t* __restrict__ remap,
                                        int64_t E, int32_t* __restrict__ f, int32_t* status) {
  bool bad = false;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const bool first = e == off[rows[e]];
    int32_t s = 1;
    if (!first) {
      bad |= el[e] < el[e - 1];
      s = (remap[el[e]] >> 6) != (remap[el[e - 1]] >> 6);
    }
    f[e] = s;
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

__global__ void comp_words_kernel(const int32_t* __restrict__ el, const int32_t* __restrict__ remap,
                                  const int32_t* __restrict__ f, const int64_t* __restrict__ wpos,
                                  int64_t E, int32_t* __restrict__ word_index,
                                  unsigned long long* word_data) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = remap[el[e]];
    const int64_t w = wpos[e] + f[e] - 1;  // inclusive count of word starts, minus one
    if (f[e]) word_index[w] = r >> 6;
    atomicOr(word_data + w, 1ull << (r & 63));
  }
}

__global__ void comp_node_start_kernel(const int32_t* __restrict__ off, int64_t n,
                                       const int64_t* __restrict__ wpos, int32_t* __restrict__ node_start) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
       i += (int64_t)gridDim.x * blockDim.x)
    node_start[i] = (int32_t)wpos[off[i]];
}

// ---------------------------------------------------------------- lookups
struct CompView {
  const int64_t* toff;
  const int32_t* slots;
  const int32_t* stash_count;
  const int32_t* stash;
  const uint32_t* bitmap;
  const int32_t* remap;
  int64_t remap_len;
  const int32_t* node_start;
  const int32_t* word_index;
  const unsigned long long* word_data;
};

template <int KIND>
__device__ __forceinline__ bool comp_contains(const CompView& c, int32_t set, int32_t key) {
  if constexpr (KIND == RS_COMP_LINEAR) {
    // cuco static_set_ref::contains, linear probing from hash % capacity
    const int64_t base = c.toff[set];
    const int64_t cap = c.toff[set + 1] - base;
    const uint32_t mask = (uint32_t)(cap - 1);
    const uint32_t h = fmix32((uint32_t)key);
    for (int64_t p = 0; p < cap; ++p) {
      const int32_t v = c.slots[base + ((h + (uint32_t)p) & mask)];
      if (v == key) return true;
      if (v == kCompEmpty) return false;
    }
    return false;
  } else if constexpr (KIND == RS_COMP_CUCKOO) {
    // cuckoo_hash.cuh:86-112
    const int64_t base = c.toff[set];
    const int32_t cap = (int32_t)(c.toff[set + 1] - base);
    if (c.slots[base + std_hash(key, set, cap)] == key) return true;
    if (c.slots[base + std_hash2(key, set, cap)] == key) return true;
    const int32_t* st = c.stash + (int64_t)set * kCuckooStash;
    const int32_t ns = c.stash_count[set];
    for (int32_t k = 0; k < ns; ++k)
      if (st[k] == key) return true;
    return false;
  } else if constexpr (KIND == RS_COMP_HOPSCOTCH) {
    // hopscotch_hash.cuh:82-104
    const int64_t base = c.toff[set];
    const int32_t total = (int32_t)(c.toff[set + 1] - base);
    const int32_t home = std_hash(key, set, total - kHopH);
    uint32_t b = c.bitmap[base + home];
    while (b) {
      const int32_t j = __builtin_ctz(b);
      b &= b - 1u;
      if (home + j < total && c.slots[base + home + j] == key) return true;
    }
    return false;
  } else {
    // roaring_bitmap.cuh:56-86: binary search of the set's word indices
    if (key < 0 || key >= c.remap_len) return false;
    const int32_t r = c.remap[key];
    if (r < 0) return false;
    const int32_t wi = r >> 6;
    int32_t lo = c.node_start[set], hi = c.node_start[set + 1] - 1;
    while (lo <= hi) {
      const int32_t mid = lo + ((hi - lo) >> 1);
      const int32_t m = c.word_index[mid];
      if (m == wi) return (c.word_data[mid] >> (r & 63)) & 1ull;
      if (m < wi)
        lo = mid + 1;
      else
        hi = mid - 1;
    }
    return false;
  }
}

// |S_s n S_g| through g's table, s the endpoint with the smaller set (the
// first endpoint on ties): a wave, lanes over S_s, counted by ballots.
template <int KIND>
__device__ __forceinline__ uint32_t comp_pair_count(const CompView& c, const int32_t* __restrict__ off,
                                                    const int32_t* __restrict__ el, int32_t u, int32_t v,
                                                    int32_t& su, int32_t& sv) {
  su = off[u + 1] - off[u];
  sv = off[v + 1] - off[v];
  int32_t s = u, g = v, ns = su;
  if (su > sv) {
    s = v;
    g = u;
    ns = sv;
  }
  const int32_t base = off[s];
  uint32_t r = 0;
  for (int32_t j0 = 0; j0 < ns; j0 += kWave) {
    const int32_t j = j0 + __lane_id();
    const bool hit = j < ns && comp_contains<KIND>(c, g, el[base + j]);
    r += (uint32_t)__popcll(__ballot(hit));
  }
  return r;
}

template <int KIND>
__global__ void __launch_bounds__(kCompBlock) comp_tc_kernel(const CompView c, const int32_t* __restrict__ off,
                                                             const int32_t* __restrict__ rows,
                                                             const int32_t* __restrict__ cols, int64_t E,
                                                             unsigned long long* out) {
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  unsigned long long total = 0;
  for (int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; e < E; e += waves) {
    int32_t su, sv;
    total += comp_pair_count<KIND>(c, off, cols, rows[e], cols[e], su, sv);
  }
  if (__lane_id() == 0 && total) atomicAdd(out, total);
}

template <int KIND>
__global__ void __launch_bounds__(kCompBlock) comp_sss_kernel(const CompView c, const int32_t* __restrict__ off,
                                                              const int32_t* __restrict__ el,
                                                              const int32_t* __restrict__ cu,
                                                              const int32_t* __restrict__ cv, int64_t P,
                                                              int32_t p, int32_t q, int64_t n,
                                                              int32_t* __restrict__ flags,
                                                              int32_t* __restrict__ sizes, int32_t* status) {
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  bool bad = false;
  for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; k < P; k += waves) {
    const int32_t u = cu[k], v = cv[k];
    if (u < 0 || v < 0 || u >= n || v >= n) {  // wave-uniform
      bad = true;
      if (__lane_id() == 0) {
        flags[k] = 0;
        sizes[k] = 0;
      }
      continue;
    }
    int32_t su, sv;
    const int64_t r = comp_pair_count<KIND>(c, off, el, u, v, su, sv);
    if (__lane_id() == 0) {
      sizes[k] = (int32_t)r;
      flags[k] = (int64_t)q * r >= (int64_t)p * ((int64_t)su + sv - r) ? 1 : 0;
    }
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

// IR conjunctive queries over tables of an inverted index (ir/ir_cuco.cu:
// 11-110): a wave per query copies term0's list into its result region and
// keeps, in order, the documents every later term's table holds (stable
// in-place compaction: a write never passes the read position).
template <int KIND>
__global__ void __launch_bounds__(kCompBlock) comp_ir_kernel(
    const CompView c, const int32_t* __restrict__ off, const int32_t* __restrict__ docs, int64_t n,
    const int32_t* __restrict__ query, const int32_t* __restrict__ qoff, int64_t Q,
    const int64_t* __restrict__ roff, int32_t* result, int32_t* __restrict__ count, int32_t* status) {
  const int lane = __lane_id();
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  bool bad = false;
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; q < Q; q += waves) {
    const int32_t qs = qoff[q], qe = qoff[q + 1];
    bool ok = qe > qs;
    for (int32_t i = qs; i < qe; ++i) ok = ok && query[i] >= 0 && query[i] < n;
    if (!ok) {  // an empty query, or a term outside the index (flagged)
      bad |= qe > qs;
      if (lane == 0) count[q] = 0;
      continue;
    }
    const int32_t t0 = query[qs];
    const int32_t n0 = off[t0 + 1] - off[t0];
    const int32_t* src = docs + off[t0];
    int32_t* dst = result + roff[q];
    int32_t cur = n0;
    if (qe - qs == 1)
      for (int32_t j = lane; j < n0; j += kWave) dst[j] = src[j];
    for (int32_t i = qs + 1; i < qe && cur > 0; ++i) {
      const int32_t term = query[i];
      int32_t w = 0;
      for (int32_t j0 = 0; j0 < cur; j0 += kWave) {
        const int32_t j = j0 + lane;
        int32_t key = 0;
        bool hit = false;
        if (j < cur) {
          key = src[j];
          hit = comp_contains<KIND>(c, term, key);
        }
        const uint64_t m = __ballot(hit);
        if (hit) dst[w + __popcll(m & lanemask_lt())] = key;
        w += __popcll(m);
      }
      cur = w;
      src = dst;
    }
    if (lane == 0) count[q] = cur;
  }
  if (__any(bad) && lane == 0) atomicOr(status, kErrBadInput);
}

}  // namespace rs

using namespace rs;

struct rs_comp {
  int32_t kind = 0;
  int64_t n = 0, E = 0;
  const int32_t* d_off = nullptr;  // borrowed
  const int32_t* d_el = nullptr;
  int64_t slots_total = 0, words = 0, remap_len = 0;
  unsigned long long failed = 0;
  DevBuf toff, slots, stash_count, stash, bitmap;      // linear / cuckoo / hopscotch
  DevBuf remap, node_start, word_index, word_data;     // roaring
  DevBuf rows;                                         // int32[E], made on first TC count
  DevBuf scratch;                                      // status word, counters, scan temp
  DevBuf tmp;
  CompView view() const {
    CompView v{};
    v.toff = toff.as<int64_t>();
    v.slots = slots.as<int32_t>();
    v.stash_count = stash_count.as<int32_t>();
    v.stash = stash.as<int32_t>();
    v.bitmap = bitmap.as<uint32_t>();
    v.remap = remap.as<int32_t>();
    v.remap_len = remap_len;
    v.node_start = node_start.as<int32_t>();
    v.word_index = word_index.as<int32_t>();
    v.word_data = word_data.as<unsigned long long>();
    return v;
  }
  // scratch: [0] status bits (int32), [8] failed inserts, [16] TC count,
  // [24] largest element (int32)
  int32_t* status() const { return scratch.as<int32_t>(); }
  unsigned long long* counter() const { return reinterpret_cast<unsigned long long*>(scratch.as<char>() + 8); }
  unsigned long long* tc_out() const { return reinterpret_cast<unsigned long long*>(scratch.as<char>() + 16); }
  int32_t* max_el() const { return reinterpret_cast<int32_t*>(scratch.as<char>() + 24); }
};

namespace {

rs_status check_status(const rs_comp* c, hipStream_t st, const char* what) {
  int32_t s = 0;
  RS_HIP(hipMemcpyAsync(&s, c->status(), sizeof(s), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (s & kErrBadInput) {
    set_error("%s: bad input (negative element, unsorted row for roaring, or a candidate id out of range)",
              what);
    return RS_EINVAL;
  }
  if (s & kErrTableFull) {
    set_error("%s: a table had no free slot", what);
    return RS_EOVERFLOW;
  }
  return RS_OK;
}

rs_status build_regions(rs_comp* c, float lf, hipStream_t st) {
  DevBuf cap;
  RS_TRY(cap.alloc(sizeof(int64_t) * (size_t)(c->n + 1)));
  const int g = grid_for(c->n, kCompBlock);
  comp_capacity_kernel<<<g, kCompBlock, 0, st>>>(c->d_off, c->n, lf, c->kind, cap.as<int64_t>(), c->status());
  RS_TRY(check_status(c, st, "comp build (a set needs more than 2^30 slots at this load factor)"));
  RS_TRY(c->toff.alloc(sizeof(int64_t) * (size_t)(c->n + 1)));
  RS_TRY(exclusive_scan_i64(cap.as<int64_t>(), c->n, c->toff.as<int64_t>(), c->tmp, st));
  RS_HIP(hipMemcpyAsync(&c->slots_total, c->toff.as<int64_t>() + c->n, sizeof(int64_t),
                        hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  RS_TRY(c->slots.alloc(sizeof(int32_t) * (size_t)c->slots_total));
  comp_fill_kernel<int32_t><<<grid_for(c->slots_total, kCompBlock), kCompBlock, 0, st>>>(
      c->slots.as<int32_t>(), c->slots_total, kCompEmpty);
  return RS_OK;
}

rs_status build_roaring(rs_comp* c, hipStream_t st) {
  const int64_t n = c->n, E = c->E;
  int32_t mx = -1;
  int32_t* d_max = c->max_el();
  RS_HIP(hipMemcpyAsync(d_max, &mx, sizeof(mx), hipMemcpyHostToDevice, st));
  if (E > 0) comp_max_kernel<<<grid_for(E, kCompBlock), kCompBlock, 0, st>>>(c->d_el, E, d_max, c->status());
  RS_HIP(hipMemcpyAsync(&mx, d_max, sizeof(mx), hipMemcpyDeviceToHost, st));
  RS_TRY(check_status(c, st, "roaring"));
  c->remap_len = n > (int64_t)mx + 1 ? n : (int64_t)mx + 1;
  DevBuf flag, rank;
  RS_TRY(flag.alloc(sizeof(int32_t) * (size_t)c->remap_len));
  RS_TRY(rank.alloc(sizeof(int64_t) * (size_t)(c->remap_len + 1)));
  comp_fill_kernel<int32_t><<<grid_for(c->remap_len, kCompBlock), kCompBlock, 0, st>>>(
      flag.as<int32_t>(), c->remap_len, 0);
  comp_flag_ids_kernel<<<grid_for(E + n, kCompBlock), kCompBlock, 0, st>>>(c->d_el, E, n, flag.as<int32_t>(),
                                                                           c->remap_len);
  RS_TRY(exclusive_scan_i32_to_i64(flag.as<int32_t>(), c->remap_len, rank.as<int64_t>(), c->tmp, st));
  RS_TRY(c->remap.alloc(sizeof(int32_t) * (size_t)c->remap_len));
  comp_remap_kernel<<<grid_for(c->remap_len, kCompBlock), kCompBlock, 0, st>>>(
      flag.as<int32_t>(), rank.as<int64_t>(), c->remap_len, c->remap.as<int32_t>());
  flag.release();
  rank.release();
  // word starts -> word positions
  DevBuf f, wpos;
  RS_TRY(c->rows.alloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)));
  RS_TRY(expand_rows(c->d_off, n, E, c->rows.as<int32_t>(), st));
  RS_TRY(f.alloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)));
  RS_TRY(wpos.alloc(sizeof(int64_t) * (size_t)(E + 1)));
  if (E > 0)
    comp_word_starts_kernel<<<grid_for(E, kCompBlock), kCompBlock, 0, st>>>(
        c->d_off, c->rows.as<int32_t>(), c->d_el, c->remap.as<int32_t>(), E, f.as<int32_t>(), c->status());
  RS_TRY(check_status(c, st, "roaring"));
  RS_TRY(exclusive_scan_i32_to_i64(f.as<int32_t>(), E, wpos.as<int64_t>(), c->tmp, st));
  RS_HIP(hipMemcpyAsync(&c->words, wpos.as<int64_t>() + E, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  RS_REQUIRE(c->words < (1ll << 31), "roaring: %lld words exceed int32 word offsets", (long long)c->words);
  RS_TRY(c->word_index.alloc(sizeof(int32_t) * (size_t)(c->words > 0 ? c->words : 1)));
  RS_TRY(c->word_data.alloc(sizeof(uint64_t) * (size_t)(c->words > 0 ? c->words : 1)));
  RS_TRY(c->node_start.alloc(sizeof(int32_t) * (size_t)(n + 1)));
  comp_fill_kernel<unsigned long long><<<grid_for(c->words, kCompBlock), kCompBlock, 0, st>>>(
      c->word_data.as<unsigned long long>(), c->words, 0ull);
  if (E > 0)
    comp_words_kernel<<<grid_for(E, kCompBlock), kCompBlock, 0, st>>>(
        c->d_el, c->remap.as<int32_t>(), f.as<int32_t>(), wpos.as<int64_t>(), E, c->word_index.as<int32_t>(),
        c->word_data.as<unsigned long long>());
  comp_node_start_kernel<<<grid_for(n + 1, kCompBlock), kCompBlock, 0, st>>>(c->d_off, n, wpos.as<int64_t>(),
                                                                             c->node_start.as<int32_t>());
  RS_HIP(hipStreamSynchronize(st));
  return RS_OK;
}

template <int KIND>
void launch_tc(const rs_comp* c, unsigned long long* out, hipStream_t st) {
  const int g = grid_for(c->E * kWave, kCompBlock);
  comp_tc_kernel<KIND><<<g, kCompBlock, 0, st>>>(c->view(), c->d_off, c->rows.as<int32_t>(), c->d_el, c->E, out);
}

template <int KIND>
void launch_sss(const rs_comp* c, const int32_t* u, const int32_t* v, int64_t P, int32_t p, int32_t q,
                int32_t* flags, int32_t* sizes, hipStream_t st) {
  const int g = grid_for(P * kWave, kCompBlock);
  comp_sss_kernel<KIND><<<g, kCompBlock, 0, st>>>(c->view(), c->d_off, c->d_el, u, v, P, p, q, c->n, flags, sizes,
                                                  c->status());
}

template <int KIND>
void launch_ir(const rs_comp* c, const int32_t* query, const int32_t* qoff, int64_t Q, const int64_t* roff,
               int32_t* result, int32_t* count, hipStream_t st) {
  const int g = grid_for(Q * kWave, kCompBlock);
  comp_ir_kernel<KIND><<<g, kCompBlock, 0, st>>>(c->view(), c->d_off, c->d_el, c->n, query, qoff, Q, roff, result,
                                                 count, c->status());
}

}  // namespace

extern "C" {

rs_status rs_comp_build(const int32_t* d_offsets, int64_t num_sets, const int32_t* d_elements, int32_t kind,
                        float load_factor, rs_stream stream, rs_comp** out) {
  RS_REQUIRE(out != nullptr, "comp: null output handle");
  *out = nullptr;
  RS_REQUIRE(d_offsets != nullptr && num_sets >= 0 && num_sets < (1ll << 31), "comp: bad set count");
  RS_REQUIRE(kind >= RS_COMP_LINEAR && kind <= RS_COMP_ROARING, "comp: unknown table kind %d", kind);
  RS_REQUIRE(kind == RS_COMP_ROARING || (load_factor > 0.0f && load_factor <= 64.0f),
             "comp: load factor must be in (0, 64]");
  hipStream_t st = as_stream(stream);
  auto* c = new rs_comp();
  c->kind = kind;
  c->n = num_sets;
  c->d_off = d_offsets;
  c->d_el = d_elements;
  rs_status s = [&]() -> rs_status {
    int32_t last = 0;
    if (num_sets > 0)
      RS_HIP(hipMemcpyAsync(&last, d_offsets + num_sets, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    c->E = last;
    RS_REQUIRE(c->E >= 0 && (c->E == 0 || d_elements != nullptr), "comp: bad element count");
    RS_TRY(c->scratch.alloc(64));
    RS_HIP(hipMemsetAsync(c->scratch.p, 0, 64, st));
    if (kind == RS_COMP_ROARING) return build_roaring(c, st);
    RS_TRY(build_regions(c, load_factor, st));
    const int g = grid_for(c->n, kCompBlock);
    if (kind == RS_COMP_LINEAR) {
      const int gw = grid_for(c->n * kWave, kCompBlock);
      comp_linear_insert_kernel<<<gw, kCompBlock, 0, st>>>(c->d_off, c->d_el, c->n, c->toff.as<int64_t>(),
                                                          c->slots.as<int32_t>(), c->counter());
    } else if (kind == RS_COMP_CUCKOO) {
      RS_TRY(c->stash_count.alloc(sizeof(int32_t) * (size_t)(c->n > 0 ? c->n : 1)));
      RS_TRY(c->stash.alloc(sizeof(int32_t) * (size_t)(c->n > 0 ? c->n * kCuckooStash : 1)));
      comp_fill_kernel<int32_t><<<grid_for(c->n * kCuckooStash, kCompBlock), kCompBlock, 0, st>>>(
          c->stash.as<int32_t>(), c->n * kCuckooStash, kCompEmpty);
      comp_cuckoo_insert_kernel<<<g, kCompBlock, 0, st>>>(c->d_off, c->d_el, c->n, c->toff.as<int64_t>(),
                                                         c->slots.as<int32_t>(), c->stash_count.as<int32_t>(),
                                                         c->stash.as<int32_t>(), c->counter());
    } else {
      RS_TRY(c->bitmap.alloc(sizeof(uint32_t) * (size_t)c->slots_total));
      comp_fill_kernel<uint32_t><<<grid_for(c->slots_total, kCompBlock), kCompBlock, 0, st>>>(
          c->bitmap.as<uint32_t>(), c->slots_total, 0u);
      comp_hopscotch_insert_kernel<<<g, kCompBlock, 0, st>>>(c->d_off, c->d_el, c->n, c->toff.as<int64_t>(),
                                                            c->slots.as<int32_t>(), c->bitmap.as<uint32_t>(),
                                                            c->counter());
    }
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(&c->failed, c->counter(), sizeof(c->failed), hipMemcpyDeviceToHost, st));
    return check_status(c, st, "comp build");
  }();
  if (s != RS_OK) {
    delete c;
    return s;
  }
  *out = c;
  return RS_OK;
}

rs_status rs_comp_info(const rs_comp* c, int64_t* h_slots, int64_t* h_failed) {
  RS_REQUIRE(c != nullptr, "comp: null handle");
  if (h_slots) *h_slots = c->kind == RS_COMP_ROARING ? c->words : c->slots_total;
  if (h_failed) *h_failed = (int64_t)c->failed;
  return RS_OK;
}

static const DevBuf* comp_array(const rs_comp* c, int32_t which, int64_t* bytes) {
  switch (which) {
    case RS_COMP_ARR_OFFSETS: *bytes = c->kind == RS_COMP_ROARING ? 0 : 8 * (c->n + 1); return &c->toff;
    case RS_COMP_ARR_SLOTS: *bytes = c->kind == RS_COMP_ROARING ? 0 : 4 * c->slots_total; return &c->slots;
    case RS_COMP_ARR_STASH_COUNT: *bytes = c->kind == RS_COMP_CUCKOO ? 4 * c->n : 0; return &c->stash_count;
    case RS_COMP_ARR_STASH: *bytes = c->kind == RS_COMP_CUCKOO ? 4 * c->n * kCuckooStash : 0; return &c->stash;
    case RS_COMP_ARR_BITMAP: *bytes = c->kind == RS_COMP_HOPSCOTCH ? 4 * c->slots_total : 0; return &c->bitmap;
    case RS_COMP_ARR_REMAP: *bytes = c->kind == RS_COMP_ROARING ? 4 * c->remap_len : 0; return &c->remap;
    case RS_COMP_ARR_NODE_START: *bytes = c->kind == RS_COMP_ROARING ? 4 * (c->n + 1) : 0; return &c->node_start;
    case RS_COMP_ARR_WORD_INDEX: *bytes = c->kind == RS_COMP_ROARING ? 4 * c->words : 0; return &c->word_index;
    case RS_COMP_ARR_WORD_DATA: *bytes = c->kind == RS_COMP_ROARING ? 8 * c->words : 0; return &c->word_data;
    default: *bytes = -1; return nullptr;
  }
}

rs_status rs_comp_array_bytes(const rs_comp* c, int32_t array, int64_t* h_bytes) {
  RS_REQUIRE(c != nullptr && h_bytes != nullptr, "comp: null argument");
  comp_array(c, array, h_bytes);
  RS_REQUIRE(*h_bytes >= 0, "comp: unknown array %d", array);
  return RS_OK;
}

rs_status rs_comp_copy_out(const rs_comp* c, int32_t array, void* h_dst, int64_t bytes) {
  RS_REQUIRE(c != nullptr, "comp: null handle");
  int64_t have = 0;
  const DevBuf* b = comp_array(c, array, &have);
  RS_REQUIRE(b != nullptr && have >= 0, "comp: unknown array %d", array);
  RS_REQUIRE(bytes == have, "comp: array %d holds %lld bytes, %lld requested", array, (long long)have,
             (long long)bytes);
  if (bytes > 0) RS_HIP(hipMemcpy(h_dst, b->p, (size_t)bytes, hipMemcpyDeviceToHost));
  return RS_OK;
}

rs_status rs_comp_tc_count(rs_comp* c, uint64_t* h_count, rs_stream stream) {
  RS_REQUIRE(c != nullptr && h_count != nullptr, "comp: null argument");
  hipStream_t st = as_stream(stream);
  if (!c->rows.p) {
    RS_TRY(c->rows.alloc(sizeof(int32_t) * (size_t)(c->E > 0 ? c->E : 1)));
    RS_TRY(expand_rows(c->d_off, c->n, c->E, c->rows.as<int32_t>(), st));
  }
  unsigned long long* out = c->tc_out();
  RS_HIP(hipMemsetAsync(out, 0, sizeof(*out), st));
  if (c->E > 0) {
    switch (c->kind) {
      case RS_COMP_LINEAR: launch_tc<RS_COMP_LINEAR>(c, out, st); break;
      case RS_COMP_CUCKOO: launch_tc<RS_COMP_CUCKOO>(c, out, st); break;
      case RS_COMP_HOPSCOTCH: launch_tc<RS_COMP_HOPSCOTCH>(c, out, st); break;
      default: launch_tc<RS_COMP_ROARING>(c, out, st); break;
    }
    RS_HIP(hipGetLastError());
  }
  unsigned long long cnt = 0;
  RS_HIP(hipMemcpyAsync(&cnt, out, sizeof(cnt), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  *h_count = cnt;
  return RS_OK;
}

rs_status rs_comp_sss_verify_async(rs_comp* c, const int32_t* d_u, const int32_t* d_v, int64_t num_pairs,
                                   int32_t p, int32_t q, int32_t* d_flags, int32_t* d_sizes, rs_stream stream) {
  RS_REQUIRE(c != nullptr && num_pairs >= 0, "comp: bad arguments");
  RS_REQUIRE(num_pairs == 0 || (d_u && d_v && d_flags && d_sizes), "comp: null candidate or output buffer");
  RS_REQUIRE(p >= 0 && q > 0, "comp: threshold p/q needs q > 0 and p >= 0");
  hipStream_t st = as_stream(stream);
  if (num_pairs == 0) return RS_OK;
  switch (c->kind) {
    case RS_COMP_LINEAR: launch_sss<RS_COMP_LINEAR>(c, d_u, d_v, num_pairs, p, q, d_flags, d_sizes, st); break;
    case RS_COMP_CUCKOO: launch_sss<RS_COMP_CUCKOO>(c, d_u, d_v, num_pairs, p, q, d_flags, d_sizes, st); break;
    case RS_COMP_HOPSCOTCH:
      launch_sss<RS_COMP_HOPSCOTCH>(c, d_u, d_v, num_pairs, p, q, d_flags, d_sizes, st);
      break;
    default: launch_sss<RS_COMP_ROARING>(c, d_u, d_v, num_pairs, p, q, d_flags, d_sizes, st); break;
  }
  RS_HIP(hipGetLastError());
  return RS_OK;
}

rs_status rs_comp_ir_query_async(rs_comp* c, const int32_t* d_query, const int32_t* d_query_offsets,
                                 int64_t num_queries, const int64_t* d_result_offsets, int32_t* d_result,
                                 int32_t* d_result_count, rs_stream stream) {
  RS_REQUIRE(c != nullptr && num_queries >= 0, "comp: bad arguments");
  RS_REQUIRE(num_queries == 0 || (d_query_offsets && d_result_offsets && d_result_count),
             "comp: null query or output buffer");
  hipStream_t st = as_stream(stream);
  if (num_queries == 0) return RS_OK;
  switch (c->kind) {
    case RS_COMP_LINEAR:
      launch_ir<RS_COMP_LINEAR>(c, d_query, d_query_offsets, num_queries, d_result_offsets, d_result,
                                d_result_count, st);
      break;
    case RS_COMP_CUCKOO:
      launch_ir<RS_COMP_CUCKOO>(c, d_query, d_query_offsets, num_queries, d_result_offsets, d_result,
                                d_result_count, st);
      break;
    case RS_COMP_HOPSCOTCH:
      launch_ir<RS_COMP_HOPSCOTCH>(c, d_query, d_query_offsets, num_queries, d_result_offsets, d_result,
                                   d_result_count, st);
      break;
    default:
      launch_ir<RS_COMP_ROARING>(c, d_query, d_query_offsets, num_queries, d_result_offsets, d_result,
                                 d_result_count, st);
      break;
  }
  RS_HIP(hipGetLastError());
  return RS_OK;
}

rs_status rs_comp_status(rs_comp* c, rs_stream stream) {
  RS_REQUIRE(c != nullptr, "comp: null handle");
  return check_status(c, as_stream(stream), "comp");
}

void rs_comp_free(rs_comp* c) { delete c; }

}  // extern "C"
