// pairs.hip -- intersection sizes of many set pairs (set_size/intersection.cu).
//
// Plain form (intersection.cu:128-181, one pair per launch sequence): fill a
// workspace table with -1, build the larger set into it, probe the smaller.
//
// Batched form: one persistent launch over all pairs.  Each wave owns a
// private LDS table region and takes the pairs of a size-binned plan (largest
// bins first) by work stealing, in runs of consecutive pairs (or round-robin,
// on request).  Per pair it builds the SMALLER set into an LDS
// table (bucket-major, the reference's bucket rule) and streams the LARGER
// set through it once, so HBM sees each input element once (the algorithmic
// minimum 4(|A|+|B|) bytes).  A smaller set above the bin's chunk is built
// chunk by chunk in value order (both sets ascending), each chunk answering
// for the larger set's elements up to its own last element; its elements are
// staged through LDS ahead of use.  Where the sets are within 8x of each
// other a chunk's table is built without atomics: its keys go in set order
// to the slots a sequential insert would give them, found by a prefix max
// over the wave (pair_ordered_wave).  The per-pair count is reduced across
// the wave and written once.
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace rs {

constexpr int kPairBlock = 256;
constexpr int kWavesPerBlock = kPairBlock / kWave;
constexpr int kPlanBlock = 256;

struct PairBinDev {
  float alpha;
  int32_t b;
  int32_t chunk;
  int32_t unroll;
};

struct PairsConst {
  PairBinDev bins[RS_PAIRS_MAX_BINS];
  int32_t num_bins;
  int32_t steal;  // > 0: waves take this many positions a time from the shared counter
};

__device__ __forceinline__ int pair_bin(int64_t total, int num_bins) {
  int bin = total <= 1 ? 0 : 63 - __builtin_clzll((unsigned long long)total);
  return bin >= num_bins ? num_bins - 1 : bin;
}

// ---------------------------------------------------------------- pair sources
// Pair i is (A_i, B_i).  CSR form: A_i = a[a_off[i] .. a_off[i+1]), B_i the
// same in b.  Indexed form: A_i = S_{u[i]}, B_i = S_{v[i]} of one set family
// (off, el) -- the candidate lists of a set-similarity join.
struct PairSrcCsr {
  const int64_t* a_off;
  const int32_t* a;
  const int64_t* b_off;
  const int32_t* b;
  __device__ __forceinline__ void get(int64_t i, const int32_t*& A, int32_t& na, const int32_t*& B,
                                      int32_t& nb) const {
    const int64_t a0 = a_off[i], b0 = b_off[i];
    A = a + a0;
    na = (int32_t)(a_off[i + 1] - a0);
    B = b + b0;
    nb = (int32_t)(b_off[i + 1] - b0);
  }
  __device__ __forceinline__ int64_t total(int64_t i) const {
    return (a_off[i + 1] - a_off[i]) + (b_off[i + 1] - b_off[i]);
  }
};

// Pair i = (A_{u[i]}, B_{v[i]}) of two CSR families with int32 offsets (one
// family twice for a self-join's candidates).
struct PairSrcIdx {
  const int32_t* a_off;
  const int32_t* a_el;
  const int32_t* b_off;
  const int32_t* b_el;
  const int32_t* u;
  const int32_t* v;
  __device__ __forceinline__ void get(int64_t i, const int32_t*& A, int32_t& na, const int32_t*& B,
                                      int32_t& nb) const {
    const int32_t x = u[i], y = v[i];
    const int32_t a0 = a_off[x], b0 = b_off[y];
    A = a_el + a0;
    na = a_off[x + 1] - a0;
    B = b_el + b0;
    nb = b_off[y + 1] - b0;
  }
  __device__ __forceinline__ int64_t total(int64_t i) const {
    const int32_t x = u[i], y = v[i];
    return (int64_t)(a_off[x + 1] - a_off[x]) + (b_off[y + 1] - b_off[y]);
  }
};

// ---------------------------------------------------------------- plan
// Layout of the plan's control block (int32): [0..kMaxBins) histogram,
// [kMaxBins..2kMaxBins) cursors, [2kMaxBins..3kMaxBins) the work-stealing
// counters of the launch groups (one per group of bins), [3kMaxBins] error
// bits.  Reset by a kernel each launch (no memset inside graphs).
constexpr int kCtlHist = 0;
constexpr int kCtlCursor = RS_PAIRS_MAX_BINS;
constexpr int kCtlNext = 2 * RS_PAIRS_MAX_BINS;
constexpr int kCtlStatus = 3 * RS_PAIRS_MAX_BINS;
constexpr int kCtlWords = 3 * RS_PAIRS_MAX_BINS + 1;

// The positions of the size-sorted order one launch works on: the bins
// bin_lo..bin_hi of the plan (read from the control block on the device:
// after the scatter, cursor[bin] is one past the bin's last position), or
// [begin, end) when bin_hi < 0.  `counter` is the launch's work-stealing word.
struct PairRange {
  int64_t begin, end;
  int32_t bin_lo, bin_hi;
  int32_t counter;
};

__device__ __forceinline__ void pair_range(const PairRange& rg, const int32_t* ctl, int64_t& b,
                                           int64_t& e) {
  if (rg.bin_hi < 0) {
    b = rg.begin;
    e = rg.end;
  } else {
    e = ctl[kCtlCursor + rg.bin_lo];
    b = (int64_t)ctl[kCtlCursor + rg.bin_hi] - ctl[kCtlHist + rg.bin_hi];
  }
}

__global__ void plan_reset(int32_t* ctl) {
  if (threadIdx.x < kCtlWords - 1) ctl[threadIdx.x] = 0;  // keeps the error word
}

template <class Src>
__global__ void plan_hist(const Src src, int64_t P, int num_bins, int32_t* ctl) {
  __shared__ int32_t h[RS_PAIRS_MAX_BINS];
  if (threadIdx.x < RS_PAIRS_MAX_BINS) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P;
       i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&h[pair_bin(src.total(i), num_bins)], 1);
  __syncthreads();
  if (threadIdx.x < num_bins && h[threadIdx.x]) atomicAdd(&ctl[kCtlHist + threadIdx.x], h[threadIdx.x]);
}

// cursors[bin] = number of pairs in bins above `bin` (largest bins first)
__global__ void plan_bins_scan(int num_bins, int32_t* ctl) {
  if (threadIdx.x == 0) {
    int32_t run = 0;
    for (int bin = num_bins - 1; bin >= 0; --bin) {
      ctl[kCtlCursor + bin] = run;
      run += ctl[kCtlHist + bin];
    }
  }
}

// Block-aggregated scatter: ranks inside the block come from LDS atomics,
// then one global atomic per (block, bin) reserves the block's range.
template <class Src>
__global__ void __launch_bounds__(kPlanBlock) plan_scatter(const Src src, int64_t P, int num_bins,
                                                           int32_t* ctl, int32_t* __restrict__ order) {
  __shared__ int32_t cnt[RS_PAIRS_MAX_BINS];
  __shared__ int32_t base[RS_PAIRS_MAX_BINS];
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < P; i0 += (int64_t)gridDim.x * blockDim.x) {
    if (threadIdx.x < RS_PAIRS_MAX_BINS) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = i0 + threadIdx.x;
    int bin = 0, rank = 0;
    if (i < P) {
      bin = pair_bin(src.total(i), num_bins);
      rank = atomicAdd(&cnt[bin], 1);
    }
    __syncthreads();
    if (threadIdx.x < num_bins && cnt[threadIdx.x] > 0)
      base[threadIdx.x] = atomicAdd(&ctl[kCtlCursor + threadIdx.x], cnt[threadIdx.x]);
    __syncthreads();
    if (i < P) order[base[bin] + rank] = (int32_t)i;
    __syncthreads();
  }
}

// ---------------------------------------------------------------- LDS tables
// A wave's transient table lives in its private LDS region, bucket-major:
// slot k of bucket j at tab[j*B + k], so one B-slot bucket is one 4/8/16-byte
// LDS read.  B (slots per bucket) is the reference's bucket_size and the
// bucket count L follows its length rule (hashtable.cu:20-36); home bucket
// x & (L-1), then the following buckets (hashtable.cu:53-101, TC.cu:11-31).
typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));

// Insert; returns the slot index taken (and sets `moved` if v had to leave
// its home bucket because it was full), or -1 if the table had no free slot.
template <int B>
__device__ __forceinline__ int32_t lds_insert(int32_t* tab, uint32_t L, int32_t v, bool& moved) {
  uint32_t bucket = (uint32_t)v & (L - 1u);
  for (uint32_t step = 0; step < L; ++step) {
#pragma unroll
    for (int k = 0; k < B; ++k)
      if (atomicCAS(&tab[bucket * B + k], kEmpty, v) == kEmpty) {
        moved |= step != 0;
        return (int32_t)(bucket * B + k);
      }
    bucket = (bucket + 1u) & (L - 1u);
  }
  return -1;
}

template <int B>
__device__ __forceinline__ void lds_read_bucket(const int32_t* tab, uint32_t bucket, int32_t (&o)[B]) {
  if constexpr (B == 4) {
    const v4i v = *reinterpret_cast<const v4i*>(tab + bucket * 4);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else if constexpr (B == 2) {
    const v2i v = *reinterpret_cast<const v2i*>(tab + bucket * 2);
    o[0] = v.x; o[1] = v.y;
  } else {
    o[0] = tab[bucket];
  }
}

constexpr int kMaxChunkPerLane = 16;  // chunk <= 1024 build-side elements
constexpr int32_t kWideRatio = 4;     // |G| >= 4 |S|: stream U blocks per lane
// Chunked pairs with |G| >= 8 |S| rebuild rarely and take the hashed chunk
// tables (one AND per lookup); below that the ordered chunk tables, whose
// build needs no atomics, win (1M pairs of 16-65536: 11.59 -> 11.03 ms).
constexpr int32_t kOrderedRatio = 8;
static_assert(kOrderedRatio >= kWideRatio, "the hashed chunk path runs wide only");

// One block of the larger set: 4*U elements per lane, 256*U per wave, from a
// 16-byte aligned position `pos` (which may lie up to 3 elements before the
// set's first element: those lanes are inactive).  Element e is valid if
// first <= e < n.  Every lane issues one aligned 16-byte load per unroll
// step, with its address clamped to the aligned group holding element n-1:
// an aligned 16-byte group that holds an element of the set lies in one
// page, so the load is in bounds even where part of it is not.  No branch
// separates the loads, so the wait before the probe counts them exactly.
template <int U>
struct GBlock {
  int32_t x[4 * U];
  bool act[4 * U];
};

template <int U>
__device__ __forceinline__ void load_block(const int32_t* __restrict__ G, int32_t pos, int32_t first,
                                           int32_t n, GBlock<U>& blk) {
  const int lane = __lane_id();
  const int32_t last4 = pos + ((n - 1 - pos) & ~3);  // aligned group of element n-1
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int32_t e = pos + 4 * (u * kWave + lane);  // first element of this group
    const int32_t el = e < last4 ? e : last4;
    const v4i v = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(G + el));
    blk.x[4 * u] = v.x;
    blk.x[4 * u + 1] = v.y;
    blk.x[4 * u + 2] = v.z;
    blk.x[4 * u + 3] = v.w;
#pragma unroll
    for (int c = 0; c < 4; ++c) blk.act[4 * u + c] = e + c >= first && e + c < n;
  }
}

// Inserts chunk [c0, c0+cl) of the build side, held in registers
// (sv[r] = S[c0 + r*64 + lane]), into the wave's LDS table, which is all
// EMPTY on entry; pos[r] records the slot each key took, so clear_chunk can
// restore the all-EMPTY state by touching only those slots.  Returns whether
// some key had to leave its home bucket.
template <int B, int CPL>
__device__ __forceinline__ bool build_chunk(int32_t* tab, uint32_t L, const int32_t (&sv)[CPL],
                                            int32_t cl, int32_t (&pos)[CPL], bool& bad) {
  const int lane = __lane_id();
  bool moved = false;
#pragma unroll
  for (int r = 0; r < CPL; ++r) {
    pos[r] = -1;
    if (r * kWave + lane < cl) {
      pos[r] = lds_insert<B>(tab, L, sv[r], moved);
      bad |= pos[r] < 0;
    }
  }
  __builtin_amdgcn_wave_barrier();
  return __any(moved);
}

template <int CPL>
__device__ __forceinline__ void clear_chunk(int32_t* tab, const int32_t (&pos)[CPL]) {
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int r = 0; r < CPL; ++r)
    if (pos[r] >= 0) tab[pos[r]] = kEmpty;
  __builtin_amdgcn_wave_barrier();
}

template <int CPL>
__device__ __forceinline__ void load_chunk(const int32_t* __restrict__ S, int32_t c0, int32_t cl,
                                           int32_t (&sv)[CPL]) {
  const int lane = __lane_id();
#pragma unroll
  for (int r = 0; r < CPL; ++r) {
    const int32_t j = r * kWave + lane;
    sv[r] = j < cl ? S[c0 + j] : kEmpty;
  }
}

// Whether x is among the B slots of a bucket: min(slot ^ x) == 0, on the
// vector ALU only.  (Counting the matches on the scalar unit instead, from
// compare masks and a ballot's population count, measured slower: 11.3 ms
// against 11.0 ms on the 1M-pair batch.)
template <int B>
__device__ __forceinline__ bool bucket_has(const int32_t (&v)[B], int32_t x) {
  uint32_t m = (uint32_t)(v[0] ^ x);
#pragma unroll
  for (int k = 1; k < B; ++k) m = min(m, (uint32_t)(v[k] ^ x));
  return m == 0u;
}

// Lookups of N keys per lane.  Buckets fill as prefixes (an insert takes the
// first EMPTY slot of its walk and no slot is emptied while the table lives),
// so one bucket read decides a key unless that bucket is full and some key of
// the table was placed outside its home bucket (`displaced`); only such keys
// walk on, bucket by bucket, until they meet themselves or a bucket with an
// EMPTY slot.  Returns the lane's hit count.
template <int B, int N>
__device__ __forceinline__ uint32_t lds_lookup(const int32_t* tab, uint32_t L, bool displaced,
                                               const int32_t (&x)[N], const bool (&act)[N],
                                               bool& bad) {
  const uint32_t mask = L - 1u;
  uint32_t hits = 0;
  int32_t v[N][B];
  uint32_t j[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    j[n] = (uint32_t)x[n] & mask;
    lds_read_bucket<B>(tab, j[n], v[n]);
  }
  if (!displaced) {  // (wave-uniform) every key sits in its home bucket
#pragma unroll
    for (int n = 0; n < N; ++n) hits += (uint32_t)(bucket_has<B>(v[n], x[n]) & act[n]);
    return hits;
  }
  bool pend[N];
  bool any = false;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const bool hit = bucket_has<B>(v[n], x[n]);
    hits += (uint32_t)(hit && act[n]);
    pend[n] = act[n] && !hit && v[n][B - 1] != kEmpty;
    any |= pend[n];
  }
  if (__any(any)) {
#pragma unroll
    for (int n = 0; n < N; ++n) {
      bool p = pend[n];
      uint32_t jj = j[n];
      for (uint32_t step = 1; step < L && __any(p); ++step) {
        jj = (jj + 1u) & mask;
        int32_t w[B];
        lds_read_bucket<B>(tab, jj, w);
        const bool hit = bucket_has<B>(w, x[n]);
        hits += (uint32_t)(p && hit);
        p = p && !hit && w[B - 1] != kEmpty;
      }
      bad |= p;
    }
  }
  return hits;
}

// Aligned start of G: at most 3 elements before G[0], in the same allocation.
__device__ __forceinline__ int32_t block_origin(const int32_t* G) {
  return -(int32_t)((reinterpret_cast<uintptr_t>(G) >> 2) & 3u);
}

// A pair whose smaller set S fits the bin's table (ns <= chunk): S is built
// once and the larger set G streams through it in aligned blocks (the next
// block is loaded while the current one is probed).  S and G need not be
// sorted.
template <int B, int U, int CPL>
__device__ __forceinline__ uint32_t pair_resident_wave(const int32_t* __restrict__ S, int32_t ns,
                                                       const int32_t* __restrict__ G, int32_t ng,
                                                       int32_t* tab, float alpha, bool& bad) {
  constexpr int32_t kBlock = 4 * U * kWave;
  int32_t sv[CPL], spos[CPL];
  load_chunk<CPL>(S, 0, ns, sv);
  const uint32_t L = (uint32_t)calc_length(ns, alpha, B);
  // three block buffers in fixed roles (the loop is unrolled by three): each
  // step loads the block two ahead into the buffer probed in the step before
  // and probes the block loaded two steps ago, so two blocks are in flight
  // while one is probed and no loaded register is ever copied; loads past
  // the end are clamped (all lanes inactive)
  int32_t pos = block_origin(G);
  GBlock<U> b0, b1, b2;
  load_block<U>(G, pos, 0, ng, b0);
  load_block<U>(G, pos + kBlock, 0, ng, b1);
  const bool displaced = build_chunk<B, CPL>(tab, L, sv, ns, spos, bad);
  uint32_t hits = 0;
  while (true) {
    load_block<U>(G, pos + 2 * kBlock, 0, ng, b2);
    hits += lds_lookup<B, 4 * U>(tab, L, displaced, b0.x, b0.act, bad);
    pos += kBlock;
    if (pos >= ng) break;
    load_block<U>(G, pos + 2 * kBlock, 0, ng, b0);
    hits += lds_lookup<B, 4 * U>(tab, L, displaced, b1.x, b1.act, bad);
    pos += kBlock;
    if (pos >= ng) break;
    load_block<U>(G, pos + 2 * kBlock, 0, ng, b1);
    hits += lds_lookup<B, 4 * U>(tab, L, displaced, b2.x, b2.act, bad);
    pos += kBlock;
    if (pos >= ng) break;
  }
  clear_chunk<CPL>(tab, spos);
  return hits;
}

// Staging of the smaller set: groups of 256 elements, each read by one
// 16-byte load per lane from the set's 16-byte aligned origin, kept in a
// two-slot ring in LDS behind each wave's table, so element idx of a staged
// group sits at ring[(idx - origin) % 512].
constexpr int32_t kGroupElems = 4 * kWave;
constexpr int32_t kRing = 2 * kGroupElems;  // int32 per wave

// A pair whose smaller set S exceeds the bin's chunk (both sets ascending).
// S is cut into chunks aligned to the staging groups: each chunk is one
// group (c = 256 elements) when the bin's table holds 256 keys at load <= 1/2,
// else c = the largest power of two <= the bin's chunk, so no chunk spans two
// groups; chunk k's table
// answers for the values up to its last element, so a block element x is
// probed against the first chunk whose last element is >= x (the last chunk
// takes everything above).  When a block reaches past the current chunk the
// table is emptied (its keys un-inserted), rebuilt from the next chunk and
// the rest of the block probed again.  G streams once in aligned blocks.
//
// Loads: the next block of G and the next unstaged group of S are loaded at
// the top of every block, unconditionally (indices clamped to the set, lanes
// past it masked), and the group is written to the ring at the bottom of the
// block -- when the block registers turn over anyway -- if its slot is free.
// Chunks are built from the ring, so neither a probe nor a rebuild waits on
// a load issued after the block it works on; only a rebuild that outruns the
// staging loads its group itself.
template <int B, int U>
__device__ __forceinline__ uint32_t pair_chunked_wave(const int32_t* __restrict__ S, int32_t ns,
                                                      const int32_t* __restrict__ G, int32_t ng,
                                                      int32_t* tab, int32_t* ring, float alpha,
                                                      int32_t chunk, bool& bad, uint32_t& unordered) {
  constexpr int32_t kBlock = 4 * U * kWave;
  const int lane = __lane_id();
  // the bin's table: calculate_length(chunk) buckets (fits the wave's region)
  const uint32_t L = (uint32_t)calc_length(chunk, alpha, B);
  int32_t c = kGroupElems;
  if (L * (uint32_t)B < 2u * kGroupElems)
    while (c > chunk) c >>= 1;
  // S group g covers element indices [s0 + 256g, s0 + 256g + 256)
  const int32_t s0 = block_origin(S);
  const int32_t s_last4 = s0 + ((ns - 1 - s0) & ~3);
  // (a plain load: a group is read again each block until a ring slot frees
  // up, and those reads should hit L2)
  auto group_load = [&](int32_t g) {
    const int32_t e = s0 + g * kGroupElems + 4 * lane;
    return *reinterpret_cast<const v4i*>(S + (e < s_last4 ? e : s_last4));
  };
  auto group_store = [&](int32_t g, const v4i& v) {
    *reinterpret_cast<v4i*>(ring + (g & 1) * kGroupElems + 4 * lane) = v;
  };
  auto group_of = [&](int32_t idx) { return (idx - s0) >> 8; };
  auto ring_at = [&](int32_t idx) { return ring[(idx - s0) & (kRing - 1)]; };
  {
    const v4i g0 = group_load(0), g1 = group_load(1);
    group_store(0, g0);
    group_store(1, g1);
  }
  int32_t sc = 2;  // next group to stage
  // the current chunk [cb, ce) = element indices [s0 + c k, s0 + c (k+1)) of
  // the set; its table keys' slots in spos
  int32_t cb = 0, ce = s0 + c < ns ? s0 + c : ns;
  bool last = ce >= ns;
  int32_t spos[4];
  // The current chunk answers for the values [lo, top]: lo is one above the
  // previous chunk's last element (0 for the first chunk), top its own last
  // element (all values for the last chunk).  A build key outside [lo, top]
  // or a probe key below lo would be answered by the wrong chunk: that is
  // input out of ascending order, flagged (kErrBadInput) rather than counted.
  uint32_t lo = 0u, top = 0u;
  auto build = [&]() {
    // stage whatever the chunk needs that is not staged yet
    const int32_t gh = group_of(ce - 1);
    while (sc <= gh) {
      __builtin_amdgcn_wave_barrier();
      group_store(sc, group_load(sc));
      ++sc;
    }
    __builtin_amdgcn_wave_barrier();
    const uint32_t last_el = (uint32_t)__builtin_amdgcn_readfirstlane(ring_at(ce - 1));
    const uint32_t hi = last ? 0xFFFFFFFFu : last_el;
    bool moved = false;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int32_t idx = cb + r * kWave + lane;
      spos[r] = -1;
      if (idx < ce) {
        const int32_t v = ring_at(idx);
        unordered |= (uint32_t)((uint32_t)v < lo) | (uint32_t)((uint32_t)v > hi);
        spos[r] = lds_insert<B>(tab, L, v, moved);
        bad |= spos[r] < 0;
      }
    }
    __builtin_amdgcn_wave_barrier();
    top = hi;
    return __any(moved);
  };
  bool displaced = build();
  int32_t pos = block_origin(G);
  // G streams through three block buffers in fixed roles, as in the
  // resident path: each step loads the block two ahead into the buffer
  // probed in the step before, so two blocks are in flight while one is
  // probed.  The step's S group is loaded before its block, so that waiting
  // for the group when it is staged at the end of the step leaves the block
  // loads behind it in flight.
  GBlock<U> b0, b1, b2;
  load_block<U>(G, pos, 0, ng, b0);
  load_block<U>(G, pos + kBlock, 0, ng, b1);
  uint32_t hits = 0;
  auto step = [&](GBlock<U>& cur, GBlock<U>& ahead) -> bool {
    const int32_t staged_g = sc;
    const v4i staged = group_load(staged_g);          // unconditional, as above
    load_block<U>(G, pos + 2 * kBlock, 0, ng, ahead);  // unconditional, as above
    bool done = false;
    if (pos >= 0 && pos + kBlock <= ng) {
      // a block wholly inside G: when its smallest and largest elements
      // both lie in [lo, top] (most blocks: G is 8x denser than S), all of
      // it is this chunk's, with no per-element routing or order check
      uint32_t mn = (uint32_t)cur.x[0], mx = mn;
#pragma unroll
      for (int i = 1; i < 4 * U; ++i) {
        mn = min(mn, (uint32_t)cur.x[i]);
        mx = max(mx, (uint32_t)cur.x[i]);
      }
      if (wave_reduce_u32<false>(mn) >= lo && wave_reduce_u32<true>(mx) <= top) {
        bool all[4 * U];
#pragma unroll
        for (int i = 0; i < 4 * U; ++i) all[i] = true;
        hits += lds_lookup<B, 4 * U>(tab, L, displaced, cur.x, all, bad);
        done = true;
      }
    }
    // an element below lo can only be met on the block's first pass: the
    // elements left for a later chunk lie above this chunk's top, which is
    // one below the next chunk's lo
    if (!done) {
#pragma unroll
      for (int i = 0; i < 4 * U; ++i) unordered |= (uint32_t)cur.act[i] & (uint32_t)((uint32_t)cur.x[i] < lo);
    }
    while (!done) {
      bool in[4 * U];
      bool rest = false;
#pragma unroll
      for (int i = 0; i < 4 * U; ++i) {
        const bool below = (uint32_t)cur.x[i] <= top;
        in[i] = cur.act[i] && below;
        cur.act[i] = cur.act[i] && !below;
        rest |= cur.act[i];
      }
      hits += lds_lookup<B, 4 * U>(tab, L, displaced, cur.x, in, bad);
      if (last || !__any(rest)) break;
      // the block reaches past this chunk: rebuild from the next one
      clear_chunk<4>(tab, spos);
      lo = top + 1u;
      cb = ce;
      ce = ns - cb < c ? ns : cb + c;  // (cb is on the chunk grid from here on)
      last = ce >= ns;
      displaced = build();
    }
    if (pos + kBlock >= ng) return false;
    // stage the group loaded above if it is still the next one and its ring
    // slot is free (the group two before it lies below the current chunk)
    if (staged_g == sc && sc <= group_of(cb) + 1) {
      __builtin_amdgcn_wave_barrier();
      group_store(sc, staged);
      ++sc;
    }
    pos += kBlock;
    return true;
  };
  while (step(b0, b2) && step(b1, b0) && step(b2, b1)) {
  }
  clear_chunk<4>(tab, spos);
  __builtin_amdgcn_wave_barrier();
  return hits;
}

// ---------------------------------------------------------------- ordered chunk tables
// Order-preserving bucket of a chunk table: the chunk's value range
// [first, last] spread over nb buckets, h(x) = clamp(x * scale + bias) with
// one fused multiply-add.  Every step is monotone (conversion, a multiply-add
// with positive scale rounded to nearest, a clamp, truncation), so ascending
// keys have non-decreasing buckets; keys outside the range land in bucket 0
// or nb - 1, where they cannot match (they are not in the chunk).
struct OrderedHash {
  float scale, bias, top_f;  // top_f = nb - 1
};

__device__ __forceinline__ uint32_t ordered_bucket(uint32_t x, const OrderedHash& oh) {
  const float f = __builtin_fmaf((float)x, oh.scale, oh.bias);
  return (uint32_t)__builtin_amdgcn_fmed3f(f, 0.f, oh.top_f);
}

// Lookups of N keys per lane in an ordered chunk table: the home bucket, then
// the next `D` buckets (D = the chunk's largest displacement, wave-uniform,
// 0 for most chunks), the last clamped to the region's last bucket `rmax`.
// Slots hold the current chunk's keys or keys of earlier chunks of the same
// pair, which all lie below every key probed here, so a slot equal to x is x.
template <int B, int N>
__device__ __forceinline__ uint32_t ordered_lookup(const int32_t* tab, const OrderedHash& oh, int32_t D,
                                                   uint32_t rmax, const int32_t (&x)[N],
                                                   const bool (&in)[N]) {
  int32_t v[N][B];
  uint32_t j[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    j[n] = ordered_bucket((uint32_t)x[n], oh);
    lds_read_bucket<B>(tab, j[n], v[n]);
  }
  uint32_t hits = 0;
#pragma unroll
  for (int n = 0; n < N; ++n) hits += (uint32_t)(bucket_has<B>(v[n], x[n]) & in[n]);
  for (int32_t d = 1; d <= D; ++d) {
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const uint32_t jj = min(j[n] + (uint32_t)d, rmax);
      int32_t w[B];
      lds_read_bucket<B>(tab, jj, w);
      hits += (uint32_t)(bucket_has<B>(w, x[n]) & in[n]);
    }
  }
  return hits;
}

// A pair whose smaller set S exceeds the bin's chunk, both sets ascending:
// the same stream as pair_chunked_wave (S staged in 256-element groups
// through the ring, G in aligned blocks, a block reaching past the current
// chunk probed again after the switch), with each group of S as one chunk
// and its table built without atomics.  The chunk's keys get ordered buckets
// (ordered_bucket over the chunk's range, nb buckets of B slots), and key i
// (in set order) takes slot pos_i = max(B*h_i, pos_{i-1} + 1): where a
// sequential linear-probing insert in set order would put it.  That is
// pos_i = i + max_{j<=i}(B*h_j - j), a prefix max over the wave.  No chunk is
// cleared before the next is built: the slots it leaves hold keys below
// everything the next chunk is asked about.  The region (nb*B + 256 slots)
// is filled with EMPTY once, when the pair ends, for the next pair.
template <int B, int U>
__device__ __forceinline__ uint32_t pair_ordered_wave(const int32_t* __restrict__ S, int32_t ns,
                                                      const int32_t* __restrict__ G, int32_t ng,
                                                      int32_t* tab, int32_t* ring, uint32_t nb,
                                                      uint32_t& unordered) {
  constexpr int32_t kBlock = 4 * U * kWave;
  const int lane = __lane_id();
  const uint32_t rmax = nb + (uint32_t)(kGroupElems / B) - 1u;  // last bucket of the region
  const int32_t s0 = block_origin(S);
  const int32_t s_last4 = s0 + ((ns - 1 - s0) & ~3);
  auto group_load = [&](int32_t g) {
    const int32_t e = s0 + g * kGroupElems + 4 * lane;
    return *reinterpret_cast<const v4i*>(S + (e < s_last4 ? e : s_last4));
  };
  auto group_store = [&](int32_t g, const v4i& v) {
    *reinterpret_cast<v4i*>(ring + (g & 1) * kGroupElems + 4 * lane) = v;
  };
  auto ring_at = [&](int32_t idx) { return ring[(idx - s0) & (kRing - 1)]; };
  {
    const v4i g0 = group_load(0), g1 = group_load(1);
    group_store(0, g0);
    group_store(1, g1);
  }
  int32_t sc = 2;  // next group to stage
  // the current chunk: group gk, element indices [cb, ce)
  int32_t gk = 0;
  int32_t cb = 0, ce = s0 + kGroupElems < ns ? s0 + kGroupElems : ns;
  bool last = ce >= ns;
  // the chunk answers for the values [lo, top] (see pair_chunked_wave)
  uint32_t lo = 0u, top = 0u;
  OrderedHash oh{0.f, 0.f, (float)(nb - 1u)};
  int32_t D = 0;
  auto build = [&]() {
    while (sc <= gk) {
      __builtin_amdgcn_wave_barrier();
      group_store(sc, group_load(sc));
      ++sc;
    }
    __builtin_amdgcn_wave_barrier();
    const v4i kv = *reinterpret_cast<const v4i*>(ring + (gk & 1) * kGroupElems + 4 * lane);
    const uint32_t first = (uint32_t)__builtin_amdgcn_readfirstlane(ring_at(cb));
    const uint32_t last_el = (uint32_t)__builtin_amdgcn_readfirstlane(ring_at(ce - 1));
    // nb buckets over [first, last_el]; a range that wraps (keys out of
    // order, flagged below) gets some valid table
    const float span = (float)(last_el - first) + 1.0f;
    oh.scale = (float)nb * __builtin_amdgcn_rcpf(span);  // (any scale works: build and lookups share it)
    oh.bias = -(float)first * oh.scale;
    const int32_t g0 = s0 + gk * kGroupElems;  // element index of the group's first slot
    const int32_t e0 = g0 + 4 * lane;           // element index of kv.x
    const int32_t k[4] = {kv.x, kv.y, kv.z, kv.w};
    // FULL: every element of the group is in the chunk (all chunks but a
    // pair's first and last), so no element needs a mask
    auto place = [&](auto full) {
      constexpr bool FULL = decltype(full)::value;
      bool act[4];
      int32_t h[4], p[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        act[c] = FULL || (e0 + c >= cb && e0 + c < ce);
        h[c] = (int32_t)ordered_bucket((uint32_t)k[c], oh);
        const int32_t t = act[c] ? B * h[c] - (4 * lane + c) : INT32_MIN / 2;
        p[c] = c == 0 ? t : max(p[c - 1], t);
      }
      const int32_t before = wave_shr1(wave_prefix_max(p[3]), INT32_MIN / 2);
      int32_t disp = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int32_t pos = 4 * lane + c + max(before, p[c]);
        if (act[c]) {
          tab[pos] = k[c];
          disp = max(disp, (int32_t)((uint32_t)pos / B) - h[c]);
        }
      }
      D = __any(disp > 0) ? __builtin_amdgcn_readlane(wave_prefix_max(disp), kWave - 1) : 0;
      // strictly ascending keys, above the previous chunk's range
      const uint32_t prev = (uint32_t)wave_shr1(kv.w, 0);
      uint32_t bad_order = (uint32_t)(act[0] && e0 - 1 >= cb && (uint32_t)k[0] <= prev);
#pragma unroll
      for (int c = 1; c < 4; ++c)
        bad_order |= (uint32_t)(act[c] && act[c - 1] && (uint32_t)k[c] <= (uint32_t)k[c - 1]);
      unordered |= bad_order;
    };
    if (g0 >= 0 && ce - g0 == kGroupElems)
      place(std::true_type{});
    else
      place(std::false_type{});
    unordered |= (uint32_t)(first < lo);
    __builtin_amdgcn_wave_barrier();
    top = last ? 0xFFFFFFFFu : last_el;
  };
  build();
  int32_t pos = block_origin(G);
  GBlock<U> b0, b1, b2;
  load_block<U>(G, pos, 0, ng, b0);
  load_block<U>(G, pos + kBlock, 0, ng, b1);
  uint32_t hits = 0;
  auto step = [&](GBlock<U>& cur, GBlock<U>& ahead) -> bool {
    const int32_t staged_g = sc;
    const v4i staged = group_load(staged_g);          // unconditional, as in pair_chunked_wave
    load_block<U>(G, pos + 2 * kBlock, 0, ng, ahead);  // unconditional
#pragma unroll
    for (int i = 0; i < 4 * U; ++i) unordered |= (uint32_t)cur.act[i] & (uint32_t)((uint32_t)cur.x[i] < lo);
    while (true) {
      bool in[4 * U];
      bool rest = false, any_in = false;
#pragma unroll
      for (int i = 0; i < 4 * U; ++i) {
        const bool below = (uint32_t)cur.x[i] <= top;
        in[i] = cur.act[i] && below;
        cur.act[i] = cur.act[i] && !below;
        rest |= cur.act[i];
        any_in |= in[i];
      }
      if (__any(any_in)) hits += ordered_lookup<B, 4 * U>(tab, oh, D, rmax, cur.x, in);
      if (last || !__any(rest)) break;
      // the block reaches past this chunk: build the next one
      lo = top + 1u;
      ++gk;
      cb = s0 + gk * kGroupElems;
      ce = ns - cb < kGroupElems ? ns : cb + kGroupElems;
      last = ce >= ns;
      build();
    }
    if (pos + kBlock >= ng) return false;
    if (staged_g == sc && sc <= gk + 1) {
      __builtin_amdgcn_wave_barrier();
      group_store(sc, staged);
      ++sc;
    }
    pos += kBlock;
    return true;
  };
  while (step(b0, b2) && step(b1, b0) && step(b2, b1)) {
  }
  // leave the region all EMPTY for the next pair
  __builtin_amdgcn_wave_barrier();
  const int32_t region4 = (int32_t)((nb * (uint32_t)B + kGroupElems + 3u) >> 2);
  for (int32_t s4 = lane; s4 < region4; s4 += kWave)
    *reinterpret_cast<v4i*>(tab + 4 * s4) = v4i{kEmpty, kEmpty, kEmpty, kEmpty};
  __builtin_amdgcn_wave_barrier();
  return hits;
}

// Buckets of an ordered chunk table for a bin: calculate_length(256, alpha,
// B), capped so that nb*B + 256 slots fit the wave's table region.
__device__ __forceinline__ uint32_t ordered_nb(float alpha, int32_t B, int32_t slots_per_wave) {
  const int32_t want = calc_length(kGroupElems, alpha, B);
  const int32_t cap = (slots_per_wave - kGroupElems) / B;
  return (uint32_t)(want > 0 && want < cap ? want : cap);
}

// Persistent kernel over the size-sorted order (largest first).  Static
// schedule: wave w takes the positions w, w + W, w + 2W, ..., so every wave
// gets a near-equal mix and no global atomic is needed.  Work stealing
// (TC.cu:141-142,176-182 takes chunks of edges the same way): lane 0 of a
// wave takes the next run of K consecutive positions from the control
// block's counter when it hands out the last position of its current run,
// a pair before the new run is needed.  Runs are cut so that the batch
// holds at least 4 runs per wave (longer runs leave waves with a whole run
// where the static deal gives them a few pairs); a batch too small for runs
// of 4 (or of the requested length, if shorter) is dealt statically.
// Either way the next pair's offsets are fetched before the current
// pair is processed.
// The default launch keeps 16 waves on a CU (4 per SIMD, the LDS split), so
// the kernel is held to 4 waves per SIMD worth of registers (128 VGPRs);
// with U = 4 it needs more and is held to 2.
template <int B, int U, int CPL, class Src>
__global__ void __launch_bounds__(kPairBlock, (U > 2 ? 2 : 4)) pairs_batched_kernel(
    const Src src, const PairRange rg, const int32_t* __restrict__ order_all, int32_t* ctl,
    int32_t* __restrict__ sizes, const PairsConst cfg, int32_t slots_per_wave) {
  extern __shared__ __attribute__((aligned(16))) int32_t lds_dyn[];
  const int lane = __lane_id();
  // wave-uniform: pair index, offsets and loop bounds stay in SGPRs
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  // per wave: the table region, then the staging ring of the window path
  int32_t* tab = lds_dyn + wid * (slots_per_wave + kRing);
  int32_t* ring = tab + slots_per_wave;
  // the region starts all EMPTY; every pair leaves it that way
  for (int32_t s4 = lane; s4 < (slots_per_wave >> 2); s4 += kWave)
    *reinterpret_cast<v4i*>(tab + 4 * s4) = v4i{kEmpty, kEmpty, kEmpty, kEmpty};
  __builtin_amdgcn_wave_barrier();
  const int64_t W = (int64_t)gridDim.x * kWavesPerBlock;
  // this launch's positions: order[0 .. P) below
  int64_t pb = 0, pe = 0;
  pair_range(rg, ctl, pb, pe);
  pb = __builtin_amdgcn_readfirstlane((int)pb);  // (positions are < 2^31)
  pe = __builtin_amdgcn_readfirstlane((int)pe);
  const int64_t P = pe > pb ? pe - pb : 0;
  const int32_t* __restrict__ order = order_all + pb;
  int32_t K = 0;
  if (cfg.steal > 0) {
    const int64_t fit = P / (4 * W);
    K = (int32_t)(cfg.steal < fit ? cfg.steal : fit);
    if (K < 4 && K < cfg.steal) K = 0;
  }
  // the counter counts runs: at most P / K + 2W takes, below 2^32
  auto take = [&]() -> int64_t {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(reinterpret_cast<uint32_t*>(ctl + rg.counter), 1u);
    return (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)v) * K;
  };
  // stealing: the run in use starts at q0 (position q0 + r handed out last);
  // the next run, q1, is taken when the run's last position is handed out
  int64_t q0 = 0, q1 = 0;
  int32_t r = 0;
  auto following = [&](int64_t cur) -> int64_t {
    if (K == 0) return cur + W;
    int64_t p;
    if (++r < K) {
      p = q0 + r;
    } else {
      r = 0;
      q0 = q1;
      p = q0;
    }
    if (r == K - 1) q1 = take();
    return p;
  };
  int64_t idx;
  if (K > 0) {
    q0 = take();
    if (K == 1) q1 = take();
    idx = q0;
  } else {
    idx = (int64_t)blockIdx.x * kWavesPerBlock + wid;
  }
  int64_t nidx = following(idx);
  bool bad = false;
  // out-of-order elements met on the chunked path (a per-lane word, not a
  // lane mask: it stays out of the scalar registers the loops need)
  uint32_t unordered = 0;
  int32_t i = 0;
  const int32_t* A = nullptr;
  const int32_t* Bp = nullptr;
  int32_t na = 0, nb = 0;
  if (idx < P) {
    i = order[idx];
    src.get(i, A, na, Bp, nb);
  }
  while (idx < P) {
    // the position after next (its atomic completes during this pair), and
    // the next pair's metadata
    const int64_t nnidx = following(nidx);
    int32_t ni = 0;
    const int32_t* nA = nullptr;
    const int32_t* nB = nullptr;
    int32_t nna = 0, nnb = 0;
    if (nidx < P) {
      ni = order[nidx];
      src.get(ni, nA, nna, nB, nnb);
    }
    int32_t ns = na, ng = nb;
    const int32_t* S = A;
    const int32_t* Gp = Bp;
    if (ns > ng) {
      const int32_t* t = S;
      S = Gp;
      Gp = t;
      const int32_t tn = ns;
      ns = ng;
      ng = tn;
    }
    uint32_t hits = 0;
    if (ns > 0) {
      const PairBinDev bin = cfg.bins[pair_bin((int64_t)ns + ng, cfg.num_bins)];
      // U blocks per lane pay off where the larger set dominates (more bytes
      // in flight per wave); near-equal sizes keep single blocks, whose chunk
      // switches re-probe less
      const bool wide = U > 1 && ng >= kWideRatio * ns;
      if (ns <= bin.chunk)
        hits = wide ? pair_resident_wave<B, U, CPL>(S, ns, Gp, ng, tab, bin.alpha, bad)
                    : pair_resident_wave<B, 1, CPL>(S, ns, Gp, ng, tab, bin.alpha, bad);
      else if ((int64_t)ng >= (int64_t)kOrderedRatio * ns)  // (wide: kOrderedRatio >= kWideRatio)
        hits = pair_chunked_wave<B, U>(S, ns, Gp, ng, tab, ring, bin.alpha, bin.chunk, bad, unordered);
      else {
        const uint32_t nb = ordered_nb(bin.alpha, B, slots_per_wave);
        // (one block per lane: a wider block crosses more chunk boundaries)
        hits = pair_ordered_wave<B, 1>(S, ns, Gp, ng, tab, ring, nb, unordered);
      }
    }
    const uint32_t total = wave_sum(hits);
    if (lane == 0) sizes[i] = (int32_t)total;
    i = ni;
    A = nA;
    Bp = nB;
    na = nna;
    nb = nnb;
    idx = nidx;
    nidx = nnidx;
  }
  if (__any(bad) && lane == 0) atomicOr(&ctl[kCtlStatus], kErrTableFull);
  if (__any(unordered != 0) && lane == 0) atomicOr(&ctl[kCtlStatus], kErrBadInput);
}

// Small pairs on part of a wave: each group of LANES lanes (64 / LANES of
// them per wave) takes its own pair, with its own table of slots_per_wave /
// (64 / LANES) slots in the wave's LDS region (the reference's probe width,
// its warp width W, hyperparameter_solver/src/config.json:9).  Only for bins
// whose pairs all fit the resident path: the smaller set S (below 2^bin
// elements) is inserted by the group's lanes, the larger set G streams
// through in 16-byte loads, 4 * LANES elements per step, and the group's
// count is reduced over its lanes.  The table is cleared by a fill of the
// L * B slots the pair used.  Positions are dealt statically: step k of wave
// w covers positions (w + k * W) * (64 / LANES) + group.
template <int B, int LANES, class Src>
__global__ void __launch_bounds__(kPairBlock) pairs_subwave_kernel(
    const Src src, const PairRange rg, const int32_t* __restrict__ order_all, int32_t* ctl,
    int32_t* __restrict__ sizes, float alpha, int32_t slots_per_wave) {
  constexpr int G = kWave / LANES;
  extern __shared__ __attribute__((aligned(16))) int32_t lds_dyn[];
  const int lane = __lane_id();
  const int grp = lane / LANES, sl = lane % LANES;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int32_t sub_slots = slots_per_wave / G;
  int32_t* tab = lds_dyn + wid * (slots_per_wave + kRing) + grp * sub_slots;
  for (int32_t s4 = lane; s4 < (slots_per_wave >> 2); s4 += kWave)
    *reinterpret_cast<v4i*>(lds_dyn + wid * (slots_per_wave + kRing) + 4 * s4) =
        v4i{kEmpty, kEmpty, kEmpty, kEmpty};
  __builtin_amdgcn_wave_barrier();
  int64_t pb = 0, pe = 0;
  pair_range(rg, ctl, pb, pe);
  const int64_t W = (int64_t)gridDim.x * kWavesPerBlock;
  bool bad = false;
  for (int64_t base = pb + ((int64_t)blockIdx.x * kWavesPerBlock + wid) * G; base < pe; base += W * G) {
    const int64_t idx = base + grp;
    const bool has = idx < pe;
    int32_t i = 0, ns = 0, ng = 0;
    const int32_t* S = nullptr;
    const int32_t* Gp = nullptr;
    if (has) {
      i = order_all[idx];
      src.get(i, S, ns, Gp, ng);
      if (ns > ng) {
        const int32_t* t = S;
        S = Gp;
        Gp = t;
        const int32_t tn = ns;
        ns = ng;
        ng = tn;
      }
    }
    const uint32_t L = has && ns > 0 ? (uint32_t)calc_length(ns, alpha, B) : 8u;
    bool moved = false;
    for (int32_t j = sl; j < ns; j += LANES) {
      const int32_t pos = lds_insert<B>(tab, L, S[j], moved);
      bad |= pos < 0;
    }
    __builtin_amdgcn_wave_barrier();
    const bool displaced = __any(moved);
    uint32_t hits = 0;
    // G from its 16-byte aligned origin (at most 3 elements before G[0])
    const int32_t g0 = ns > 0 && Gp ? block_origin(Gp) : 0;
    const int32_t steps = ns > 0 ? (ng - g0 + 4 * LANES - 1) / (4 * LANES) : 0;
    for (int32_t k = 0; __any(k < steps); ++k) {
      int32_t x[4] = {kEmpty, kEmpty, kEmpty, kEmpty};
      bool act[4] = {false, false, false, false};
      const int32_t e = g0 + 4 * (k * LANES + sl);
      if (k < steps && e < ng) {
        const v4i v = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(Gp + e));
        x[0] = v.x;
        x[1] = v.y;
        x[2] = v.z;
        x[3] = v.w;
#pragma unroll
        for (int c = 0; c < 4; ++c) act[c] = e + c >= 0 && e + c < ng;
      }
      hits += lds_lookup<B, 4>(tab, L, displaced, x, act, bad);
    }
#pragma unroll
    for (int d = LANES / 2; d > 0; d >>= 1) hits += __shfl_xor(hits, d, kWave);
    if (has && sl == 0) sizes[i] = (int32_t)hits;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t s4 = (uint32_t)sl; s4 < L * (uint32_t)B / 4u; s4 += LANES)
      *reinterpret_cast<v4i*>(tab + 4 * s4) = v4i{kEmpty, kEmpty, kEmpty, kEmpty};
    __builtin_amdgcn_wave_barrier();
  }
  if (__any(bad) && lane == 0) atomicOr(&ctl[kCtlStatus], kErrTableFull);
}

// ---------------------------------------------------------------- plain form
// intersection.cu:128-152 without the replication: every thread inserts one
// element of the build set into the global workspace table (BN = L).
__global__ void plain_build_kernel(const int32_t* __restrict__ B, int64_t nb, int32_t* tab,
                                   uint32_t L, int32_t bsz, int32_t* size_slot, int32_t* err) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) *size_slot = 0;
  if (j >= nb) return;
  const int32_t v = B[j];
  uint32_t bucket = v & (L - 1u);
  int32_t k = 0;
  const uint32_t bound = L * (uint32_t)bsz;
  for (uint32_t tries = 0;; ++tries) {
    if (tries >= bound) {
      atomicOr(err, kErrTableFull);
      return;
    }
    if (atomicCAS(&tab[bucket + (int64_t)k * L], kEmpty, v) == kEmpty) return;
    if (++k == bsz) {
      k = 0;
      bucket = (bucket + 1u) & (L - 1u);
    }
  }
}

// intersection.cu:154-181: probe each element of the smaller set, per-wave
// count, one atomic per wave.
__global__ void plain_probe_kernel(const int32_t* __restrict__ A, int64_t na,
                                   const int32_t* __restrict__ tab, uint32_t L, int32_t bsz,
                                   int32_t* size_slot, int32_t* err) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int r = 0;
  if (j < na) {
    const int32_t key = A[j];
    r = probe(tab, (int64_t)L, key & (L - 1u), L, bsz, key);
  }
  const uint32_t s = wave_sum((uint32_t)(r == 1));
  if (__lane_id() == 0 && s) atomicAdd(size_slot, (int32_t)s);
  if (__any(r < 0) && __lane_id() == 0) atomicOr(err, kErrProbeBound);
}

__global__ void set_word_kernel(int32_t* p, int32_t v) { *p = v; }

}  // namespace rs

using namespace rs;

struct rs_pairs_plan {
  int64_t max_pairs = 0;
  rs_pairs_opts opts{};
  PairsConst cfg{};
  DevBuf order;   // int32[max_pairs]
  DevBuf ctl;     // int32[kCtlWords]
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the batched kernel
  ~rs_pairs_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

// A bin may run on part of a wave (lanes < 64) if the table of each of its
// pairs fits the lane group's share of the LDS region: totals below
// 2^(bin+1) (every bin but the last), so the smaller set has fewer than
// 2^bin elements and is built whole.
static bool subwave_fits(const rs_pairs_opts* o, int bin, const rs_pairs_bin& bn) {
  if (bin >= o->num_bins - 1 || bin >= 20) return false;
  const int32_t smax = (1 << bin) - 1;
  const int32_t groups = kWave / bn.lanes;
  const int32_t L = calc_length(smax > 0 ? smax : 1, bn.load_factor, bn.bucket_size);
  return L > 0 && o->lds_slots_per_wave % (4 * groups) == 0 &&
         (int64_t)L * bn.bucket_size <= o->lds_slots_per_wave / groups;
}

static rs_status validate_opts(const rs_pairs_opts* o, PairsConst* cfg) {
  RS_REQUIRE(o->num_bins >= 1 && o->num_bins <= RS_PAIRS_MAX_BINS, "pairs: bad num_bins %d",
             o->num_bins);
  RS_REQUIRE(o->waves_per_cu >= kWavesPerBlock && o->waves_per_cu <= 32 &&
                 o->waves_per_cu % kWavesPerBlock == 0,
             "pairs: waves_per_cu must be a multiple of 4 in [4, 32]");
  RS_REQUIRE(o->lds_slots_per_wave >= 256 && o->lds_slots_per_wave <= 8192 &&
                 (o->lds_slots_per_wave & 3) == 0,
             "pairs: lds_slots_per_wave must be a multiple of 4 in [256, 8192]");
  RS_REQUIRE((int64_t)(o->lds_slots_per_wave + kRing) * 4 * o->waves_per_cu <= 160 * 1024,
             "pairs: %d waves x (%d table + %d staging) slots exceed the 160 KiB LDS of a CU",
             o->waves_per_cu, o->lds_slots_per_wave, kRing);
  RS_REQUIRE(o->schedule >= 0 && o->schedule <= 64,
             "pairs: schedule must be 0 (static) or 1..64 pairs per take (work stealing), not %d",
             o->schedule);
  cfg->num_bins = o->num_bins;
  cfg->steal = o->schedule;
  for (int i = 0; i < o->num_bins; ++i) {
    const rs_pairs_bin& bn = o->bins[i];
    RS_REQUIRE(bn.load_factor > 0.f && (bn.bucket_size == 1 || bn.bucket_size == 2 ||
                                        bn.bucket_size == 4) && bn.chunk >= kWave &&
                   bn.chunk <= kMaxChunkPerLane * kWave,
               "pairs: bin %d needs bucket_size in {1,2,4} and chunk in [%d, %d]", i, kWave,
               kMaxChunkPerLane * kWave);
    RS_REQUIRE(bn.unroll == o->bins[0].unroll && (bn.unroll == 1 || bn.unroll == 2 || bn.unroll == 4),
               "pairs: unroll must be 1, 2 or 4 and the same in every bin (bin %d)", i);
    RS_REQUIRE(bn.lanes == 16 || bn.lanes == 32 || bn.lanes == 64,
               "pairs: bin %d: lanes per pair must be 16, 32 or 64 (not %d)", i, bn.lanes);
    RS_REQUIRE(bn.lanes == 64 || subwave_fits(o, i, bn),
               "pairs: bin %d cannot run %d lanes per pair: it must not be the last bin and a table of "
               "2^bin - 1 keys must fit 1/%d of a wave's LDS region",
               i, bn.lanes, kWave / bn.lanes);
    const int32_t L = calc_length(bn.chunk, bn.load_factor, bn.bucket_size);
    RS_REQUIRE(L > 0 && (int64_t)L * bn.bucket_size <= o->lds_slots_per_wave,
               "pairs: bin %d table (%d buckets x %d slots) exceeds the %d-slot LDS region", i, L,
               bn.bucket_size, o->lds_slots_per_wave);
    cfg->bins[i] = PairBinDev{bn.load_factor, bn.bucket_size, bn.chunk, bn.unroll};
  }
  return RS_OK;
}

// A run of neighbouring bins launched as one kernel: the same bucket size and
// lanes per pair (and, below 64 lanes, the same load factor, a kernel
// argument there).  Bins are launched largest first, the order's order.
struct PairGroup {
  int bin_lo, bin_hi, b, lanes;
  float alpha;
  int max_chunk;
};

static std::vector<PairGroup> pair_groups(const rs_pairs_opts& o, int only_bin) {
  std::vector<PairGroup> gs;
  for (int bin = o.num_bins - 1; bin >= 0; --bin) {
    if (only_bin >= 0 && bin != only_bin) continue;
    const rs_pairs_bin& bn = o.bins[bin];
    if (!gs.empty()) {
      PairGroup& g = gs.back();
      if (g.b == bn.bucket_size && g.lanes == bn.lanes && (bn.lanes == 64 || g.alpha == bn.load_factor) &&
          g.bin_lo == bin + 1) {
        g.bin_lo = bin;
        g.max_chunk = bn.chunk > g.max_chunk ? bn.chunk : g.max_chunk;
        continue;
      }
    }
    gs.push_back(PairGroup{bin, bin, bn.bucket_size, bn.lanes, bn.load_factor, bn.chunk});
  }
  return gs;
}

template <int B, int U, class Src>
static void launch_bu(int cpl, int blocks, size_t lds, const Src& src, const PairRange& rg,
                      const int32_t* order, int32_t* ctl, int32_t* sizes, const PairsConst& cfg,
                      int32_t spw, hipStream_t st) {
  if (cpl <= 4)
    pairs_batched_kernel<B, U, 4, Src><<<blocks, kPairBlock, lds, st>>>(src, rg, order, ctl, sizes, cfg, spw);
  else if (cpl <= 8)
    pairs_batched_kernel<B, U, 8, Src><<<blocks, kPairBlock, lds, st>>>(src, rg, order, ctl, sizes, cfg, spw);
  else
    pairs_batched_kernel<B, U, 16, Src><<<blocks, kPairBlock, lds, st>>>(src, rg, order, ctl, sizes, cfg, spw);
}

template <int B, class Src>
static void launch_sub(int lanes, int blocks, size_t lds, const Src& src, const PairRange& rg,
                       const int32_t* order, int32_t* ctl, int32_t* sizes, float alpha, int32_t spw,
                       hipStream_t st) {
  if (lanes == 16)
    pairs_subwave_kernel<B, 16, Src><<<blocks, kPairBlock, lds, st>>>(src, rg, order, ctl, sizes, alpha, spw);
  else
    pairs_subwave_kernel<B, 32, Src><<<blocks, kPairBlock, lds, st>>>(src, rg, order, ctl, sizes, alpha, spw);
}

// One kernel per group of bins.  from_plan: each group's positions come from
// the plan's bin cursors; otherwise the single group covers [0, count).
template <class Src>
static void launch_batched(const rs_pairs_opts& o, const PairsConst& cfg, const Src& src, int64_t count,
                           bool from_plan, int only_bin, const int32_t* order, int32_t* ctl, int32_t* sizes,
                           hipStream_t st) {
  const int blocks = device_cu_count() * o.waves_per_cu / kWavesPerBlock;
  const int32_t spw = o.lds_slots_per_wave;
  const size_t lds = sizeof(int32_t) * (size_t)(spw + kRing) * kWavesPerBlock;
  const int U = o.bins[0].unroll;
  const std::vector<PairGroup> gs = pair_groups(o, only_bin);
  for (size_t gi = 0; gi < gs.size(); ++gi) {
    const PairGroup& g = gs[gi];
    const PairRange rg{0, count, from_plan ? g.bin_lo : -1, from_plan ? g.bin_hi : -1, kCtlNext + (int)gi};
    if (g.lanes < 64) {
      if (g.b == 4) launch_sub<4, Src>(g.lanes, blocks, lds, src, rg, order, ctl, sizes, g.alpha, spw, st);
      else if (g.b == 2) launch_sub<2, Src>(g.lanes, blocks, lds, src, rg, order, ctl, sizes, g.alpha, spw, st);
      else launch_sub<1, Src>(g.lanes, blocks, lds, src, rg, order, ctl, sizes, g.alpha, spw, st);
      continue;
    }
    const int cpl = (g.max_chunk + kWave - 1) / kWave;
#define RS_LAUNCH_BU(b_, u_) launch_bu<b_, u_, Src>(cpl, blocks, lds, src, rg, order, ctl, sizes, cfg, spw, st)
    if (g.b == 4) {
      if (U == 1) RS_LAUNCH_BU(4, 1); else if (U == 2) RS_LAUNCH_BU(4, 2); else RS_LAUNCH_BU(4, 4);
    } else if (g.b == 2) {
      if (U == 1) RS_LAUNCH_BU(2, 1); else if (U == 2) RS_LAUNCH_BU(2, 2); else RS_LAUNCH_BU(2, 4);
    } else {
      if (U == 1) RS_LAUNCH_BU(1, 1); else if (U == 2) RS_LAUNCH_BU(1, 2); else RS_LAUNCH_BU(1, 4);
    }
#undef RS_LAUNCH_BU
  }
}

template <class Src>
static rs_status pairs_enqueue(rs_pairs_plan* plan, const Src& src, int64_t P, int32_t* d_sizes,
                               hipStream_t st) {
  int32_t* ctl = plan->ctl.as<int32_t>();
  int32_t* order = plan->order.as<int32_t>();
  const int nb = plan->cfg.num_bins;
  plan_reset<<<1, 64, 0, st>>>(ctl);
  if (P > 0) {
    const int g = grid_for(P, kPlanBlock, 4);
    plan_hist<Src><<<g, kPlanBlock, 0, st>>>(src, P, nb, ctl);
    plan_bins_scan<<<1, 64, 0, st>>>(nb, ctl);
    plan_scatter<Src><<<g, kPlanBlock, 0, st>>>(src, P, nb, ctl, order);
    if (plan->timing) RS_HIP(hipEventRecord(plan->ev0, st));
    launch_batched<Src>(plan->opts, plan->cfg, src, P, true, -1, order, ctl, d_sizes, st);
    if (plan->timing) RS_HIP(hipEventRecord(plan->ev1, st));
  }
  RS_HIP(hipGetLastError());
  return RS_OK;
}

static rs_status pairs_enqueue(rs_pairs_plan* plan, const int64_t* d_a_off, const int32_t* d_a,
                               const int64_t* d_b_off, const int32_t* d_b, int64_t P,
                               int32_t* d_sizes, hipStream_t st) {
  return pairs_enqueue<PairSrcCsr>(plan, PairSrcCsr{d_a_off, d_a, d_b_off, d_b}, P, d_sizes, st);
}

namespace rs {
// Runs the kernel of bin `bin`'s settings over the pairs listed in
// d_order[0..count) only (the autotuner times one size bin at a time).
rs_status pairs_launch_ordered(const rs_pairs_opts* opts, int bin, const int64_t* d_a_off,
                               const int32_t* d_a, const int64_t* d_b_off, const int32_t* d_b,
                               const int32_t* d_order, int64_t count, int32_t* d_ctl,
                               int32_t* d_sizes, hipStream_t st) {
  PairsConst cfg{};
  RS_TRY(validate_opts(opts, &cfg));
  RS_REQUIRE(bin >= 0 && bin < opts->num_bins, "pairs: bad bin %d", bin);
  plan_reset<<<1, 64, 0, st>>>(d_ctl);
  launch_batched<PairSrcCsr>(*opts, cfg, PairSrcCsr{d_a_off, d_a, d_b_off, d_b}, count, false, bin, d_order,
                             d_ctl, d_sizes, st);
  RS_HIP(hipGetLastError());
  return RS_OK;
}
bool pairs_subwave_fits(const rs_pairs_opts* opts, int bin, const rs_pairs_bin* b) {
  return subwave_fits(opts, bin, *b);
}
int pairs_ctl_words() { return kCtlWords; }
int pairs_ctl_status_index() { return kCtlStatus; }
}  // namespace rs

extern "C" {

void rs_pairs_default_opts(rs_pairs_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->num_bins = RS_PAIRS_MAX_BINS;
  o->waves_per_cu = 16;
  o->lds_slots_per_wave = 2048;
  // work stealing in runs of 16: a pair's cost depends on its size ratio,
  // not only on its size, so a static deal of ~250 pairs per wave leaves the
  // slowest wave ~20% behind the mean; runs keep the shared counter
  // uncontended (one take per pair stalls on it)
  o->schedule = RS_PAIRS_STEAL;
  for (int i = 0; i < RS_PAIRS_MAX_BINS; ++i) {
    // alpha 1/8, b 4: L = nextpow2 > 2*chunk, so a 255-element chunk fills
    // the 2048-slot region at <= 1/2 key per 4-slot bucket, where a home
    // bucket overflows (forcing the walking lookup) in < 10% of chunks; two
    // blocks per lane in flight where one set is >= 4x the other.
    o->bins[i] = rs_pairs_bin{0.125f, 4, 255, 2, 64};
  }
}

rs_status rs_pairs_plan_create(int64_t max_pairs, const rs_pairs_opts* opts,
                               rs_pairs_plan** out) {
  RS_REQUIRE(out && max_pairs >= 0 && max_pairs < (1ll << 31), "rs_pairs_plan_create: bad args");
  *out = nullptr;
  auto* p = new rs_pairs_plan();
  p->max_pairs = max_pairs;
  if (opts)
    p->opts = *opts;
  else
    rs_pairs_default_opts(&p->opts);
  rs_status s = validate_opts(&p->opts, &p->cfg);
  if (s == RS_OK) s = p->order.alloc(sizeof(int32_t) * (size_t)(max_pairs > 0 ? max_pairs : 1));
  if (s == RS_OK) s = p->ctl.alloc(sizeof(int32_t) * kCtlWords);
  if (s == RS_OK && hipMemset(p->ctl.p, 0, sizeof(int32_t) * kCtlWords) != hipSuccess) {
    set_error("rs_pairs_plan_create: hipMemset failed");
    s = RS_EDEVICE;
  }
  if (s != RS_OK) {
    delete p;
    return s;
  }
  *out = p;
  return RS_OK;
}

void rs_pairs_plan_free(rs_pairs_plan* p) { delete p; }

rs_status rs_pairs_intersect_batched_async(rs_pairs_plan* plan, const int64_t* d_a_off,
                                           const int32_t* d_a, const int64_t* d_b_off,
                                           const int32_t* d_b, int64_t num_pairs, int32_t* d_sizes,
                                           rs_stream stream) {
  RS_REQUIRE(plan && d_a_off && d_b_off && (d_sizes || num_pairs == 0),
             "rs_pairs_intersect_batched_async: null argument");
  RS_REQUIRE(num_pairs >= 0 && num_pairs <= plan->max_pairs,
             "rs_pairs_intersect_batched_async: %lld pairs exceed the plan's %lld",
             (long long)num_pairs, (long long)plan->max_pairs);
  return pairs_enqueue(plan, d_a_off, d_a, d_b_off, d_b, num_pairs, d_sizes, as_stream(stream));
}

rs_status rs_pairs_intersect_indexed_async(rs_pairs_plan* plan, const int32_t* d_set_offsets,
                                           const int32_t* d_set_elements, const int32_t* d_u,
                                           const int32_t* d_v, int64_t num_pairs, int32_t* d_sizes,
                                           rs_stream stream) {
  RS_REQUIRE(plan && d_set_offsets && d_set_elements && (num_pairs == 0 || (d_u && d_v && d_sizes)),
             "rs_pairs_intersect_indexed_async: null argument");
  RS_REQUIRE(num_pairs >= 0 && num_pairs <= plan->max_pairs,
             "rs_pairs_intersect_indexed_async: %lld pairs exceed the plan's %lld",
             (long long)num_pairs, (long long)plan->max_pairs);
  return pairs_enqueue<PairSrcIdx>(plan, PairSrcIdx{d_set_offsets, d_set_elements, d_set_offsets,
                                                    d_set_elements, d_u, d_v}, num_pairs,
                                   d_sizes, as_stream(stream));
}

rs_status rs_pairs_intersect_cross_async(rs_pairs_plan* plan, const int32_t* d_a_offsets,
                                         const int32_t* d_a_elements, const int32_t* d_b_offsets,
                                         const int32_t* d_b_elements, const int32_t* d_u,
                                         const int32_t* d_v, int64_t num_pairs, int32_t* d_sizes,
                                         rs_stream stream) {
  RS_REQUIRE(plan && d_a_offsets && d_a_elements && d_b_offsets && d_b_elements &&
                 (num_pairs == 0 || (d_u && d_v && d_sizes)),
             "rs_pairs_intersect_cross_async: null argument");
  RS_REQUIRE(num_pairs >= 0 && num_pairs <= plan->max_pairs,
             "rs_pairs_intersect_cross_async: %lld pairs exceed the plan's %lld",
             (long long)num_pairs, (long long)plan->max_pairs);
  return pairs_enqueue<PairSrcIdx>(plan, PairSrcIdx{d_a_offsets, d_a_elements, d_b_offsets,
                                                    d_b_elements, d_u, d_v}, num_pairs,
                                   d_sizes, as_stream(stream));
}

rs_status rs_pairs_graph_capture(rs_pairs_plan* plan, const int64_t* d_a_off, const int32_t* d_a,
                                 const int64_t* d_b_off, const int32_t* d_b, int64_t num_pairs,
                                 int32_t* d_sizes, rs_stream stream) {
  RS_REQUIRE(plan && d_a_off && d_b_off && d_sizes, "rs_pairs_graph_capture: null argument");
  RS_REQUIRE(num_pairs >= 0 && num_pairs <= plan->max_pairs, "rs_pairs_graph_capture: too many pairs");
  RS_REQUIRE(stream != nullptr, "rs_pairs_graph_capture: needs a non-default stream");
  hipStream_t st = as_stream(stream);
  if (plan->exec) {
    (void)hipGraphExecDestroy(plan->exec);
    plan->exec = nullptr;
  }
  if (plan->graph) {
    (void)hipGraphDestroy(plan->graph);
    plan->graph = nullptr;
  }
  (void)device_cu_count();  // attribute query outside the capture
  RS_REQUIRE(!plan->timing, "rs_pairs_graph_capture: disable kernel timing before capturing");
  RS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  rs_status s = pairs_enqueue(plan, d_a_off, d_a, d_b_off, d_b, num_pairs, d_sizes, st);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (s != RS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess) {
    set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
    return RS_EDEVICE;
  }
  plan->graph = g;
  RS_HIP(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
  return RS_OK;
}

rs_status rs_pairs_graph_launch(rs_pairs_plan* plan, rs_stream stream) {
  RS_REQUIRE(plan && plan->exec, "rs_pairs_graph_launch: no captured graph");
  RS_HIP(hipGraphLaunch(plan->exec, as_stream(stream)));
  return RS_OK;
}

rs_status rs_pairs_plan_timing(rs_pairs_plan* p, int32_t enable) {
  RS_REQUIRE(p, "rs_pairs_plan_timing: null plan");
  if (enable && !p->ev0) {
    RS_HIP(hipEventCreate(&p->ev0));
    RS_HIP(hipEventCreate(&p->ev1));
  }
  p->timing = enable != 0;
  return RS_OK;
}

rs_status rs_pairs_plan_kernel_ms(rs_pairs_plan* p, float* h_ms) {
  RS_REQUIRE(p && h_ms && p->ev0, "rs_pairs_plan_kernel_ms: timing was never enabled");
  RS_HIP(hipEventSynchronize(p->ev1));
  RS_HIP(hipEventElapsedTime(h_ms, p->ev0, p->ev1));
  return RS_OK;
}

rs_status rs_pairs_plan_status(rs_pairs_plan* p, rs_stream stream) {
  RS_REQUIRE(p, "rs_pairs_plan_status: null plan");
  hipStream_t st = as_stream(stream);
  int32_t err = 0;
  RS_HIP(hipMemcpyAsync(&err, p->ctl.as<int32_t>() + kCtlStatus, sizeof(err),
                        hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (err) {
    RS_HIP(hipMemsetAsync(p->ctl.as<int32_t>() + kCtlStatus, 0, sizeof(int32_t), st));
    RS_HIP(hipStreamSynchronize(st));
    if (err & kErrBadInput) {
      set_error("pairs: a pair whose smaller set exceeds its bin's chunk has a set that is not "
                "strictly ascending (the chunked path needs ascending sets)");
      return RS_EINVAL;
    }
    set_error("pairs: an LDS table had no free slot or a probe hit its bound");
    return RS_EOVERFLOW;
  }
  return RS_OK;
}

rs_status rs_pairs_intersect_plain(const int64_t* d_a_off, const int32_t* d_a,
                                   const int64_t* d_b_off, const int32_t* d_b,
                                   const int64_t* h_a_off, const int64_t* h_b_off, int64_t num_pairs,
                                   float load_factor, int32_t bucket_size, int32_t* d_workspace,
                                   int32_t* d_sizes, rs_stream stream) {
  RS_REQUIRE(d_a_off && d_b_off && h_a_off && h_b_off && d_workspace && (d_sizes || !num_pairs),
             "rs_pairs_intersect_plain: null argument");
  RS_REQUIRE(load_factor > 0.f && bucket_size > 0, "plain pairs: bad alpha/b");
  RS_REQUIRE(num_pairs >= 0, "plain pairs: negative pair count");
  (void)d_a_off;
  (void)d_b_off;
  // A table of L buckets x b slots takes at most L*b keys: a larger set
  // (load_factor above 1) cannot be built, and the reference's unbounded
  // insert would spin (intersection.cu:128-152).  Rejected before any launch.
  for (int64_t i = 0; i < num_pairs; ++i) {
    const int64_t na = h_a_off[i + 1] - h_a_off[i], nb = h_b_off[i + 1] - h_b_off[i];
    RS_REQUIRE(na >= 0 && nb >= 0 && na < (1ll << 31) && nb < (1ll << 31),
               "plain pairs: pair %lld has a bad offset range", (long long)i);
    const int64_t nB = na > nb ? na : nb;
    const int32_t L = calc_length((int32_t)nB, load_factor, bucket_size);
    RS_REQUIRE(L > 0, "plain pairs: set too large");
    if (nB > (int64_t)L * bucket_size) {
      set_error("plain pairs: pair %lld: %lld keys exceed the %d x %d slots of its table "
                "(load factor %g)", (long long)i, (long long)nB, L, bucket_size, (double)load_factor);
      return RS_EOVERFLOW;
    }
  }
  hipStream_t st = as_stream(stream);
  // the call's error word: one per host thread and device, cleared by a
  // kernel here and read back after the last pair
  int dev = 0;
  RS_HIP(hipGetDevice(&dev));
  RS_REQUIRE(dev >= 0 && dev < 64, "plain pairs: device ordinal %d out of range", dev);
  static thread_local DevBuf err_buf[64];
  if (!err_buf[dev].p) {
    BlockCacheScope uncached(nullptr);  // the word lives as long as the thread
    RS_TRY(err_buf[dev].alloc(16));
  }
  int32_t* err = err_buf[dev].as<int32_t>();
  set_word_kernel<<<1, 1, 0, st>>>(err, 0);
  for (int64_t i = 0; i < num_pairs; ++i) {
    const int64_t a0 = h_a_off[i], na = h_a_off[i + 1] - a0;
    const int64_t b0 = h_b_off[i], nb = h_b_off[i + 1] - b0;
    const int32_t* A = d_a + a0;
    const int32_t* B = d_b + b0;
    int64_t nA = na, nB = nb;
    if (nA > nB) {  // intersection.cu:198-202: probe with the smaller
      const int32_t* t = A;
      A = B;
      B = t;
      const int64_t tn = nA;
      nA = nB;
      nB = tn;
    }
    const int32_t L = calc_length((int32_t)nB, load_factor, bucket_size);
    RS_HIP(hipMemsetAsync(d_workspace, 0xFF, sizeof(int32_t) * (size_t)L * bucket_size, st));
    const unsigned gb = (unsigned)((nB + 255) / 256);
    plain_build_kernel<<<gb ? gb : 1, 256, 0, st>>>(B, nB, d_workspace, (uint32_t)L, bucket_size,
                                                    d_sizes + i, err);
    if (nA > 0)
      plain_probe_kernel<<<(unsigned)((nA + 255) / 256), 256, 0, st>>>(
          A, nA, d_workspace, (uint32_t)L, bucket_size, d_sizes + i, err);
  }
  RS_HIP(hipGetLastError());
  int32_t h_err = 0;
  RS_HIP(hipMemcpyAsync(&h_err, err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (h_err) {
    set_error("plain pairs: a table had no free slot or a probe walked the whole table");
    return RS_EOVERFLOW;
  }
  return RS_OK;
}

}  // extern "C"
