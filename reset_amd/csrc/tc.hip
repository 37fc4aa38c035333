// tc.hip -- triangle counting by hash probes (TC.cu:128-190, baseline/tc/tc.cu).
//
// count = sum over oriented edges (u, v) of |N+(u) cap N+(v)|, each edge's
// smaller out-list probing the larger one's table (TC.cu:146-155).
//
// Edge schedule (built once per CSR, reused by both layouts): every edge
// becomes (s, g), s the endpoint whose keys are probed and g the one whose
// table is probed, with w = d+(s) probes; W is the int64 exclusive prefix of
// w in schedule order, so the probe space [0, W[E]) is flat.
//
// Staged form (tables probed at least kStageRatio times per slot, which at
// R-MAT scale 20-24 carries ~99% of all probes): the schedule is sorted by
// (class, g, s), so each such table's probes are one contiguous range of the
// probe space.  That range is cut into units; a workgroup copies the unit's
// table (all b slot planes) from HBM into LDS once, then its waves walk the
// unit's probe tiles, reading probe keys from HBM (coalesced runs) and the
// table from LDS.  HBM traffic per probe drops to the key's 4 bytes plus the
// amortised table copy.  Classes bound the table length (L <= 1024, 2048,
// 4096) and so the launch's LDS size and occupancy.
//
// Tiled form (the rest): the flattened probe space of the remaining edges is
// cut into equal tiles; a wave takes a tile, finds its first edge by binary
// search in W, loads 64 edges' metadata (one per lane) and walks the
// flattened probes 64 at a time, every probe reading its table in HBM.
//
// Both forms use all 64 lanes regardless of degree and split hub edges
// across waves.
#include <algorithm>
#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "tables.hpp"

// Schedule groups: staged classes 0..2, then unstaged edges (3), then edges
// with no probes or a malformed endpoint (4).
constexpr int kTcGroups = 5;
constexpr int kTcStagedClasses = 3;
constexpr int kXcds = 8;

struct rs_tc_plan {
  const rs_tables* t = nullptr;
  int64_t E = 0;
  int64_t total = 0;          // probes of the whole schedule, W[E]
  rs::DevBuf W;               // int64[E+1]
  rs::DevBuf es, eg;          // int32[E]: probing / probed endpoint
  rs::DevBuf counter;         // work-stealing counter of the plain form
  int64_t bounds[kTcGroups + 1] = {0, 0, 0, 0, 0, 0};  // first edge of each group, then E
  int64_t wb[kTcGroups + 1] = {0, 0, 0, 0, 0, 0};      // W at those edges
  // Staged units: table, probe range [u_wb, u_we), tiles [u_tile[u], u_tile[u+1]).
  int64_t num_units = 0, num_tiles = 0;
  rs::DevBuf u_g, u_wb, u_we, u_tile, tile_edge;
  rs::DevBuf xcd_units;       // int64[kTcStagedClasses][kXcds+1]: unit split per XCD
  rs::DevBuf unit_next;       // uint64[kTcStagedClasses][kXcds]: staged work-stealing counters
  int32_t class_lmax[kTcStagedClasses] = {0, 0, 0};    // largest staged L per class (0 = none)
  int64_t class_units[kTcStagedClasses + 1] = {0, 0, 0, 0};
};

namespace rs {

constexpr int kTcBlock = 256;
constexpr int64_t kTcTile = 4096;
constexpr int64_t kTcTileMin = 512;
// A table is staged when its probes reach kStageRatio x its L*b slots.
constexpr int64_t kStageRatio = 2;
// Units hold at least max(kUnitProbes, kUnitSlotFactor * L*b) probes (a table
// with fewer probes is one unit), so each copy serves >= 16 probes per slot.
constexpr int64_t kUnitProbes = 65536;
constexpr int64_t kUnitSlotFactor = 16;
// Probe tile of the staged form: one wave's share of a unit at a time.
constexpr int64_t kStageTile = 2048;
// Staged-table classes by length; waves per workgroup of each class.
constexpr int32_t kClassL[kTcStagedClasses] = {1024, 2048, 4096};
constexpr int kClassWaves[kTcStagedClasses] = {4, 8, 16};
constexpr int kStageU = 4;
constexpr size_t kLdsBytes = 160 * 1024;

__host__ __device__ inline int tc_class_of(uint32_t L) { return L <= 1024 ? 0 : (L <= 2048 ? 1 : 2); }

// LDS bytes of a class-c workgroup staging tables of length <= L: the
// table, each wave's marker row, and 16 bytes for the next-unit word (the
// kernel keeps no static LDS, so this is its whole LDS).
constexpr size_t kStageWordBytes = 16;
inline size_t stage_lds_bytes(int c, int64_t L, int32_t b) {
  return (size_t)L * (size_t)b * 4 + (size_t)kClassWaves[c] * kWave * kStageU * 4 + kStageWordBytes;
}

// Largest length that may be staged with bucket size b (0 = none).
inline int32_t stage_lmax(int32_t b) {
  int32_t best = 0;
  for (int c = 0; c < kTcStagedClasses; ++c)
    if (stage_lds_bytes(c, kClassL[c], b) <= kLdsBytes) best = kClassL[c];
  return best;
}

__global__ void zero_u64(unsigned long long* p) { *p = 0ull; }
__global__ void zero_u64_n(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}
__global__ void set_i32(int32_t* p, int32_t v) { *p = v; }

// Schedule endpoints of CSR edge e = (rows[e], cols[e]): s has the smaller
// out-list (ties keep the CSR source, TC.cu:146-155), g is the other, w =
// d+(s) probes.  False for an endpoint outside [0, n): malformed input, which
// is flagged and given no work, so no later kernel indexes the offsets with it.
__device__ __forceinline__ bool tc_edge(const int32_t* __restrict__ rows,
                                        const int32_t* __restrict__ cols,
                                        const int32_t* __restrict__ off, int64_t n, int64_t e,
                                        int32_t& s, int32_t& g, int32_t& w) {
  const int32_t u = rows[e], v = cols[e];
  if (v < 0 || v >= n) return false;
  const int32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
  s = du > dv ? v : u;
  g = du > dv ? u : v;
  w = du > dv ? dv : du;
  return true;
}

// P[g] = probes into g's table over the whole schedule.
__global__ void tc_table_probes_kernel(const int32_t* __restrict__ rows,
                                       const int32_t* __restrict__ cols,
                                       const int32_t* __restrict__ off, int64_t n, int64_t E,
                                       unsigned long long* __restrict__ P) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    int32_t s, g, w;
    if (tc_edge(rows, cols, off, n, e, s, g, w) && w > 0)
      atomicAdd(&P[g], (unsigned long long)w);
  }
}

// key[e] = group << 62 | g << 31 | s, group = the staged class of g's table
// or 3 (probed in HBM); ~0 for an edge with no probes or a bad endpoint.
__global__ void tc_edge_keys_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                                    const int32_t* __restrict__ off, int64_t n, int64_t E,
                                    const int32_t* __restrict__ Ls,
                                    const unsigned long long* __restrict__ P, int32_t b,
                                    int32_t lmax_stage, uint64_t* __restrict__ key,
                                    int32_t* status) {
  bool bad = false;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    int32_t s = 0, g = 0, w = 0;
    if (!tc_edge(rows, cols, off, n, e, s, g, w)) {
      bad = true;
      key[e] = ~0ull;
      continue;
    }
    if (w == 0) {
      key[e] = ~0ull;
      continue;
    }
    const int32_t L = Ls[g];
    const bool staged = L <= lmax_stage && (int64_t)P[g] >= kStageRatio * (int64_t)L * b;
    const uint64_t grp = staged ? (uint64_t)tc_class_of((uint32_t)L) : 3ull;
    key[e] = (grp << 62) | ((uint64_t)g << 31) | (uint64_t)s;
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

__device__ __forceinline__ int tc_group(uint64_t k) { return k == ~0ull ? 4 : (int)(k >> 62); }

__global__ void tc_edge_split_kernel(const uint64_t* __restrict__ key, const int32_t* __restrict__ off,
                                     int64_t E, int32_t* __restrict__ es,
                                     int32_t* __restrict__ eg, int32_t* __restrict__ w,
                                     int64_t* __restrict__ bounds) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = key[e];
    // group starts: edge e opens every group after its predecessor's, up to its own
    const int c = tc_group(k);
    const int pc = e == 0 ? -1 : tc_group(key[e - 1]);
    for (int q = pc + 1; q <= c; ++q) bounds[q] = e;
    if (k == ~0ull) {
      es[e] = 0;
      eg[e] = 0;
      w[e] = 0;
      continue;
    }
    const int32_t sm = (int32_t)(k & 0x7FFFFFFFull), lg = (int32_t)((k >> 31) & 0x7FFFFFFFull);
    es[e] = sm;
    eg[e] = lg;
    w[e] = off[sm + 1] - off[sm];
  }
}

// flags[e] = 1 where a probed table's run of edges starts, e in [0, E).
__global__ void tc_segment_flags_kernel(const int32_t* __restrict__ eg, int64_t E,
                                        int32_t* __restrict__ flags) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    flags[e] = e == 0 || eg[e] != eg[e - 1];
}

// Per segment (run of edges probing one table): table, its length and the
// segment's probe range.
__global__ void tc_segment_info_kernel(const int32_t* __restrict__ starts, int64_t nseg,
                                       int64_t e_end, const int32_t* __restrict__ eg,
                                       const int32_t* __restrict__ Ls,
                                       const int64_t* __restrict__ W, int32_t* __restrict__ sg,
                                       int32_t* __restrict__ sL, int64_t* __restrict__ swb,
                                       int64_t* __restrict__ swe) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nseg;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e0 = starts[i];
    const int64_t e1 = i + 1 < nseg ? starts[i + 1] : e_end;
    const int32_t g = eg[e0];
    sg[i] = g;
    sL[i] = Ls[g];
    swb[i] = W[e0];
    swe[i] = W[e1];
  }
}

// tile_edge[t] = the edge holding the first probe of staged tile t: a wave
// per unit, a lane per tile, binary search over the unit's segment [e0, e1).
__global__ void tc_tile_edges_kernel(const int64_t* __restrict__ u_wb,
                                     const int64_t* __restrict__ u_tile,
                                     const int32_t* __restrict__ u_e0,
                                     const int32_t* __restrict__ u_e1, int64_t nunits,
                                     const int64_t* __restrict__ W, int32_t* __restrict__ tile_edge) {
  const int lane = __lane_id();
  const int64_t waves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t u = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; u < nunits;
       u += waves) {
    const int64_t t0 = u_tile[u], nt = u_tile[u + 1] - t0;
    for (int64_t i = lane; i < nt; i += kWave) {
      const int64_t ts = u_wb[u] + i * kStageTile;
      int64_t lo = u_e0[u], hi = u_e1[u];  // largest e with W[e] <= ts
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (W[mid] <= ts) lo = mid; else hi = mid;
      }
      tile_edge[t0 + i] = (int32_t)lo;
    }
  }
}

// First-slot fast path for UNROLL independent keys, then the bounded walk
// (TC.cu:11-31) for the keys whose home slot was occupied by another key.
template <int LAYOUT, int UNROLL>
__device__ __forceinline__ void probe_batch(const int32_t* __restrict__ T, int64_t BN, int32_t b,
                                            uint32_t M, const int32_t (&key)[UNROLL],
                                            const int64_t (&tb)[UNROLL],
                                            const uint32_t (&L)[UNROLL],
                                            const bool (&act)[UNROLL], uint32_t& found,
                                            bool& bad) {
  uint32_t bucket[UNROLL];
  int32_t first[UNROLL];
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    bucket[u] = home_bucket<LAYOUT>((uint32_t)key[u], L[u], M);
    first[u] = act[u] ? T[tb[u] + bucket[u]] : kEmpty;
  }
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) {
    if (!act[u]) continue;
    if (first[u] == key[u]) {
      ++found;
    } else if (first[u] != kEmpty) {
      // continue from slot 1 of the home bucket
      const int32_t* Tt = T + tb[u];
      uint32_t bk = bucket[u];
      int32_t k = 1;
      const uint32_t slots = L[u] * (uint32_t)b;
      uint32_t step = 1;
      for (; step < slots; ++step) {
        if (k == b) {
          k = 0;
          bk = (bk + 1u) & (L[u] - 1u);
        }
        const int32_t v = Tt[bk + (int64_t)k * BN];
        if (v == key[u]) {
          ++found;
          break;
        }
        if (v == kEmpty) break;
        ++k;
      }
      if (step >= slots) bad = true;
    }
  }
}

// Tiled form over the edges [e_lo, e_hi) with probe range [w_lo, w_hi).
// Workgroups are dealt to the 8 XCDs round-robin (block b on XCD b % 8), and
// each XCD has its own L2, so XCD x takes the contiguous tile range
// [x*R, (x+1)*R): neighbouring tiles, which probe the same tables, share an L2.
template <int LAYOUT, int UNROLL>
__global__ void __launch_bounds__(kTcBlock) tc_tiled_kernel(
    const int32_t* __restrict__ es, const int32_t* __restrict__ eg,
    const int32_t* __restrict__ off, const int32_t* __restrict__ keys,
    const int32_t* __restrict__ Ls, const int64_t* __restrict__ toff,
    const int32_t* __restrict__ T, int64_t BN, int32_t b, uint32_t M,
    const int64_t* __restrict__ W, int64_t e_lo, int64_t e_hi, int64_t w_lo, int64_t w_hi,
    int64_t tile, unsigned long long* __restrict__ out, int32_t* __restrict__ status) {
  const int lane = __lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int xcd = blockIdx.x % kXcds;
  const int64_t local_wave = (int64_t)(blockIdx.x / kXcds) * (kTcBlock / kWave) + wid;
  const int64_t xcd_waves = (int64_t)(gridDim.x / kXcds) * (kTcBlock / kWave);
  const int64_t ntiles = (w_hi - w_lo + tile - 1) / tile;
  const int64_t per_xcd = (ntiles + kXcds - 1) / kXcds;
  const int64_t t_begin = xcd * per_xcd;
  const int64_t t_end = t_begin + per_xcd < ntiles ? t_begin + per_xcd : ntiles;
  uint32_t found = 0;
  bool bad = false;
  for (int64_t ti = t_begin + local_wave; ti < t_end; ti += xcd_waves) {
    const int64_t ts = w_lo + ti * tile;
    const int64_t te = ts + tile < w_hi ? ts + tile : w_hi;
    // largest e with W[e] <= ts (that edge has work, since ts < w_hi)
    int64_t lo = e_lo, hi = e_hi;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (W[mid] <= ts) lo = mid; else hi = mid;
    }
    int64_t e = lo;
    while (true) {
      const int64_t ei = e + lane;
      int32_t cnt = 0, kbase = 0;
      uint32_t Lv = 8;
      int64_t tb = 0;
      bool valid = false;
      if (ei < e_hi) {
        const int64_t ws = W[ei];
        valid = ws < te;
        if (valid) {
          const int64_t we = W[ei + 1];
          const int64_t s = ws > ts ? ws : ts;
          const int64_t en = we < te ? we : te;
          cnt = (int32_t)(en - s);
          if (cnt > 0) {
            const int32_t u = es[ei], v = eg[ei];
            kbase = off[u] + (int32_t)(s - ws);
            tb = toff[v];
            Lv = (uint32_t)Ls[v];
          } else {
            cnt = 0;
          }
        }
      }
      const int32_t incl = wave_inclusive_scan(cnt);
      const int32_t excl = incl - cnt;
      const int32_t S = __shfl(incl, kWave - 1, kWave);
      for (int32_t it = 0; it < S; it += kWave * UNROLL) {
        int32_t key[UNROLL];
        int64_t tbv[UNROLL];
        uint32_t Lo[UNROLL];
        bool act[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
          const int32_t p = it + u * kWave + lane;
          act[u] = p < S;
          int owner = 0;
#pragma unroll
          for (int sft = 32; sft > 0; sft >>= 1) {
            const int32_t c = __shfl(incl, owner + sft - 1, kWave);
            if (c <= p) owner += sft;
          }
          owner = owner > kWave - 1 ? kWave - 1 : owner;
          const int32_t kb = __shfl(kbase, owner, kWave);
          const int32_t ex = __shfl(excl, owner, kWave);
          tbv[u] = __shfl(tb, owner, kWave);
          Lo[u] = __shfl(Lv, owner, kWave);
          key[u] = act[u] ? keys[kb + (p - ex)] : kEmpty;
        }
        probe_batch<LAYOUT, UNROLL>(T, BN, b, M, key, tbv, Lo, act, found, bad);
      }
      if (__popcll(__ballot(valid)) < kWave) break;
      e += kWave;
      if (e >= e_hi) break;
      if (W[e] >= te) break;
    }
  }
  const uint32_t s = wave_sum(found);
  if (lane == 0 && s) atomicAdd(out, (unsigned long long)s);
  if (__any(bad) && lane == 0) atomicOr(status, kErrProbeBound);
}

// Probe of U keys per lane against a table staged in LDS, slot planes kept
// (slot k of bucket j at tab[k*L + j]), walked in the order of TC.cu:11-31:
// slots 0..b-1 of the home bucket, then the next bucket, until the key or an
// EMPTY slot; bounded at L*b slots.  The U keys advance in lockstep; a group
// of lanes' keys that has resolved skips its step.
template <int LAYOUT, int U>
__device__ __forceinline__ uint32_t probe_lds(const lds_i32* __restrict__ tab, uint32_t L,
                                              int32_t b, uint32_t M, const int32_t (&x)[U],
                                              const bool (&act)[U], bool& bad) {
  uint32_t j[U], pos[U];
  int32_t k[U];
  bool pend[U];
  uint32_t hits = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    j[u] = home_bucket<LAYOUT>((uint32_t)x[u], L, M);
    pos[u] = j[u];
    k[u] = 0;
    pend[u] = act[u];
  }
  const uint32_t bound = L * (uint32_t)b;
  for (uint32_t step = 0; step < bound; ++step) {
    int32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = tab[pos[u]];
    bool left = false;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool hit = pend[u] && v[u] == x[u];
      hits += hit;
      pend[u] = pend[u] && !hit && v[u] != kEmpty;
      const bool wrap = k[u] + 1 == b;
      k[u] = wrap ? 0 : k[u] + 1;
      j[u] = wrap ? ((j[u] + 1u) & (L - 1u)) : j[u];
      pos[u] = wrap ? j[u] : pos[u] + L;
      left |= pend[u];
    }
    if (!__any(left)) return hits;
  }
  bad = true;
  return hits;
}

// Bucket-at-a-time probe against a table staged bucket-major in LDS (slot k
// of bucket j at tab[j*B + k]).  A bucket's occupied slots always form a
// prefix of it (an insert takes the first EMPTY slot of its walk and no slot
// is ever cleared), so the slot walk of TC.cu:11-31 reduces per bucket to:
// the key is among the B slots -> found; else the last slot is EMPTY ->
// absent; else the next bucket.  The B compares fold into one xor/min chain.
// All U keys take their home bucket together; the rare full bucket is
// followed in a bounded loop (L buckets in all).
// With b = 5 the staged copy keeps slots 0-3 of bucket j as one 16-byte
// group at tab[4j] and slot 4 at tab[4L + j] (split5 below): two LDS reads
// per bucket instead of three for a 20-byte bucket.
template <int B>
__device__ __forceinline__ void lds_bucket(const lds_i32* __restrict__ tab, uint32_t L, uint32_t j,
                                           int32_t x, bool& hit, bool& full) {
  if constexpr (B == 5) {
    typedef int v4i_l __attribute__((ext_vector_type(4)));
    const v4i_l q = *reinterpret_cast<const __attribute__((address_space(3))) v4i_l*>(tab + 4u * j);
    const int32_t v4 = tab[4u * L + j];
    uint32_t m = min(min((uint32_t)(q.x ^ x), (uint32_t)(q.y ^ x)), min((uint32_t)(q.z ^ x), (uint32_t)(q.w ^ x)));
    m = min(m, (uint32_t)(v4 ^ x));
    hit = m == 0u;
    full = v4 != kEmpty;
  } else {
    const lds_i32* bk = tab + __umul24(j, (uint32_t)B);  // (24-bit multiply: L <= 4096)
    uint32_t m = ~0u;
    int32_t last = 0;
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int32_t v = bk[k];
      m = min(m, (uint32_t)(v ^ x));
      last = v;
    }
    hit = m == 0u;
    full = last != kEmpty;
  }
}

template <int LAYOUT, int U, int B>
__device__ __forceinline__ uint32_t probe_lds_buckets(const lds_i32* __restrict__ tab, uint32_t L,
                                                      uint32_t M, const int32_t (&x)[U],
                                                      const bool (&act)[U], bool& bad) {
  uint32_t j[U];
  bool pend[U];
  uint32_t hits = 0;
  bool any_pend = false;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    j[u] = home_bucket<LAYOUT>((uint32_t)x[u], L, M);
    bool hit, full;
    lds_bucket<B>(tab, L, j[u], x[u], hit, full);
    hits += act[u] && hit;
    pend[u] = act[u] && !hit && full;
    any_pend |= pend[u];
  }
  if (__any(any_pend)) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint32_t jj = j[u];
      bool p = pend[u];
      for (uint32_t step = 1; step < L && __any(p); ++step) {
        jj = (jj + 1u) & (L - 1u);
        bool hit, full;
        lds_bucket<B>(tab, L, jj, x[u], hit, full);
        hits += p && hit;
        p = p && !hit && full;
      }
      bad |= p;
    }
  }
  return hits;
}

// Staged form, one workgroup per unit at a time (units of one class, split
// across XCDs in contiguous ranges so a table's units share an L2).  The
// workgroup copies the unit's table into LDS with 16-byte loads (lengths are
// powers of two >= 8 and table offsets multiples of 8, so every plane row is
// 16-byte aligned), then each wave takes the unit's tiles in turn.
//
// Inside a tile, lanes hold 64 edges' clipped probe ranges; the probes are
// walked 64*U at a time.  A probe's edge is found without a search: each
// edge that starts inside the chunk writes the chunk's tag at its start
// offset in the wave's LDS marker row, a ballot of "tag here" gives the
// start mask, and the edge of lane l is (edges started before the chunk) +
// (starts at or below l).  Tags grow per chunk, so the row is never cleared.
template <int LAYOUT, int NW, int U, int B>
__global__ void __launch_bounds__(NW * kWave) tc_staged_kernel(
    const int32_t* __restrict__ es, const int32_t* __restrict__ off,
    const int32_t* __restrict__ keys, const int32_t* __restrict__ Ls,
    const int64_t* __restrict__ toff, const int32_t* __restrict__ T, int64_t BN, int32_t b,
    uint32_t M, const int64_t* __restrict__ W, int64_t E, const int32_t* __restrict__ u_g,
    const int64_t* __restrict__ u_wb, const int64_t* __restrict__ u_we,
    const int64_t* __restrict__ u_tile, const int32_t* __restrict__ tile_edge,
    const int64_t* __restrict__ xcd_units, int32_t table_words,
    unsigned long long* __restrict__ out, int32_t* __restrict__ status,
    unsigned long long* __restrict__ unit_next) {
  extern __shared__ __attribute__((aligned(16))) int32_t lds_tc[];
  int32_t* const tab = lds_tc;
  const lds_i32* const ltab = (const lds_i32*)lds_tc;  // the same table, 32-bit offsets
  const int lane = __lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int32_t* const marks = lds_tc + table_words + wid * (kWave * U);
  // the next-unit word sits after every wave's marker row (table_words is a
  // multiple of 8 and each row holds 64*U words, so it is 8-byte aligned)
  int64_t& s_next = *reinterpret_cast<int64_t*>(lds_tc + table_words + NW * (kWave * U));
#pragma unroll
  for (int u = 0; u < U; ++u) marks[u * kWave + lane] = -1;
  int32_t tag = 0;
  const int xcd = blockIdx.x % kXcds;
  const int64_t ub = xcd_units[xcd], ue = xcd_units[xcd + 1];
  // Work stealing within the XCD's unit range (TC.cu:176-182 takes edge
  // chunks the same way), one counter per XCD so that the units an XCD's L2
  // serves stay together: a unit's cost varies with its table and its
  // probes' lists.  Thread 0 takes the next unit when a unit starts and
  // publishes it after the table barrier (every thread read the previous
  // value before it); all threads read it after the unit.
  unsigned long long* const ctr = unit_next + xcd;
  if (threadIdx.x == 0) s_next = ub + (int64_t)atomicAdd(ctr, 1ull);
  __syncthreads();
  int64_t u = s_next;
  uint32_t found = 0;
  bool bad = false;
  while (u < ue) {
    unsigned long long taken = 0;
    if (threadIdx.x == 0) taken = atomicAdd(ctr, 1ull);
    const int32_t g = u_g[u];
    const uint32_t L = (uint32_t)Ls[g];
    const int32_t* __restrict__ Tt = T + toff[g];
    const uint32_t q = L >> 2, qs = __builtin_ctz(q);
    const uint32_t n4 = q * (uint32_t)b;
    for (uint32_t i = threadIdx.x; i < n4; i += NW * kWave) {
      const uint32_t k = i >> qs, jj = (i & (q - 1u)) << 2;
      const int4 v = *reinterpret_cast<const int4*>(Tt + (int64_t)k * BN + jj);
      if constexpr (B == 0) {  // slot planes kept
        *reinterpret_cast<int4*>(tab + k * L + jj) = v;
      } else if constexpr (B == 5) {  // slots 0-3 bucket-major, slot 4 in its own run
        if (k < 4) {
          tab[(jj + 0) * 4 + k] = v.x;
          tab[(jj + 1) * 4 + k] = v.y;
          tab[(jj + 2) * 4 + k] = v.z;
          tab[(jj + 3) * 4 + k] = v.w;
        } else {
          *reinterpret_cast<int4*>(tab + 4 * L + jj) = v;
        }
      } else {  // bucket-major
        tab[(jj + 0) * B + k] = v.x;
        tab[(jj + 1) * B + k] = v.y;
        tab[(jj + 2) * B + k] = v.z;
        tab[(jj + 3) * B + k] = v.w;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_next = ub + (int64_t)taken;
    const int64_t t0 = u_tile[u], t1 = u_tile[u + 1], wb = u_wb[u], we = u_we[u];
    for (int64_t t = t0 + wid; t < t1; t += NW) {
      const int64_t ts = wb + (t - t0) * kStageTile;
      const int64_t te = ts + kStageTile < we ? ts + kStageTile : we;
      int64_t e = tile_edge[t];
      while (true) {
        const int64_t ei = e + lane;
        int32_t cnt = 0, kb = 0;
        bool valid = false;
        if (ei < E) {
          const int64_t ws = W[ei];
          valid = ws < te;
          if (valid) {
            const int64_t wn = W[ei + 1];
            const int64_t s = ws > ts ? ws : ts;
            const int64_t en = wn < te ? wn : te;
            cnt = (int32_t)(en - s);
            kb = off[es[ei]] + (int32_t)(s - ws);
          }
        }
        const int32_t incl = wave_inclusive_scan(cnt);
        const int32_t excl = incl - cnt;
        const int32_t S = __shfl(incl, kWave - 1, kWave);
        const int32_t delta = kb - excl;  // key index of probe p of this lane's edge: p + delta
        for (int32_t it = 0; it < S; it += kWave * U) {
          const bool starts_in = valid && excl > it && excl < it + kWave * U;
          int32_t base = (int32_t)__popcll(__ballot(valid && excl <= it)) - 1;
          int32_t x[U];
          bool act[U];
          if (!__any(starts_in)) {
            // the slab lies inside one edge's probes (most of a long list's
            // slabs): every probe's key index is p + that edge's delta
            const int32_t d = __builtin_amdgcn_readlane(delta, base);
            if (it + kWave * U <= S) {  // a full slab: keys from one uniform base
              const int32_t* __restrict__ kp = keys + (int64_t)(it + d);
#pragma unroll
              for (int uu = 0; uu < U; ++uu) {
                act[uu] = true;
                x[uu] = kp[uu * kWave + lane];
              }
            } else {
#pragma unroll
              for (int uu = 0; uu < U; ++uu) {
                const int32_t p = it + uu * kWave + lane;
                act[uu] = p < S;
                x[uu] = keys[act[uu] ? p + d : 0];
              }
            }
          } else {
            ++tag;
            if (starts_in) marks[excl - it] = tag;
#pragma unroll
            for (int uu = 0; uu < U; ++uu) {
              const bool m = marks[uu * kWave + lane] == tag;
              const uint64_t Mu = __ballot(m);
              const int32_t below = (int32_t)__builtin_amdgcn_mbcnt_hi(
                  (uint32_t)(Mu >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)Mu, 0u));
              const int32_t own = base + below + (m ? 1 : 0);
              base += (int32_t)__popcll(Mu);
              const int32_t p = it + uu * kWave + lane;
              act[uu] = p < S;
              const int32_t d = __shfl(delta, own & (kWave - 1), kWave);
              x[uu] = keys[act[uu] ? p + d : 0];
            }
          }
          if constexpr (B == 0)
            found += probe_lds<LAYOUT, U>(ltab, L, b, M, x, act, bad);
          else
            found += probe_lds_buckets<LAYOUT, U, B>(ltab, L, M, x, act, bad);
        }
        if (__popcll(__ballot(valid)) < kWave) break;
        e += kWave;
        if (e >= E) break;
        if (W[e] >= te) break;
      }
    }
    __syncthreads();
    u = s_next;
  }
  const uint32_t sum = wave_sum(found);
  if (lane == 0 && sum) atomicAdd(out, (unsigned long long)sum);
  if (__any(bad) && lane == 0) atomicOr(status, kErrProbeBound);
}

// The plain form: one wave per edge, CHUNK consecutive edges per grab, the
// next chunk taken from an atomic counter (TC.cu:141-183), lanes striding
// over the smaller out-list.
template <int LAYOUT>
__global__ void __launch_bounds__(kTcBlock) tc_plain_kernel(
    const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
    const int32_t* __restrict__ off, const int32_t* __restrict__ keys,
    const int32_t* __restrict__ Ls, const int64_t* __restrict__ toff,
    const int32_t* __restrict__ T, int64_t BN, int32_t b, uint32_t M, int64_t E, int32_t chunk,
    int32_t* __restrict__ counter, unsigned long long* __restrict__ out,
    int32_t* __restrict__ status) {
  const int lane = __lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  uint32_t found = 0;
  bool bad = false;
  int64_t e = wave * chunk;
  int64_t e_end = e + chunk;
  while (e < E) {
    int32_t u = rows[e], v = cols[e];
    int32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
    if (du > dv) {
      const int32_t x = u;
      u = v;
      v = x;
      du = dv;
    }
    const int32_t* Tt = T + toff[v];
    const uint32_t Lv = (uint32_t)Ls[v];
    const int32_t kb = off[u];
    for (int32_t j = lane; j < du; j += kWave) {
      const int32_t key = keys[kb + j];
      const int r = probe(Tt, BN, home_bucket<LAYOUT>((uint32_t)key, Lv, M), Lv, b, key);
      found += r == 1;
      bad |= r < 0;
    }
    if (++e == e_end) {
      int32_t nxt = 0;
      if (lane == 0) nxt = atomicAdd(counter, chunk);
      e = __shfl(nxt, 0, kWave);
      e_end = e + chunk;
    }
  }
  const uint32_t s = wave_sum(found);
  if (lane == 0 && s) atomicAdd(out, (unsigned long long)s);
  if (__any(bad) && lane == 0) atomicOr(status, kErrProbeBound);
}

static rs_status check_args(const rs_tables* t, const rs_tc_plan* p, int32_t layout,
                            uint64_t* d_count, const int32_t** T) {
  RS_REQUIRE(t && p && d_count, "tc: null argument");
  RS_REQUIRE(p->t == t, "tc: plan was made for another tables handle");
  RS_TRY(require_layout(t, layout, T));
  return RS_OK;
}

#define RS_ROC(expr)                                                  \
  do {                                                                \
    hipError_t _e = (expr);                                           \
    if (_e != hipSuccess) {                                           \
      ::rs::set_error("%s failed: %s", #expr, hipGetErrorString(_e)); \
      return RS_EDEVICE;                                              \
    }                                                                 \
  } while (0)

// Units of the staged groups: segments (runs of edges probing one table) are
// extracted on the device and cut into units on the host (a few hundred
// thousand segments at R-MAT scale 24); the tiles' first edges are then
// found on the device.
static rs_status tc_build_units(const rs_tables* t, hipStream_t st, rs_tc_plan* p, DevBuf& tmp) {
  const int64_t E3 = p->bounds[3];  // staged edges are [0, E3)
  if (E3 == 0) return RS_OK;
  DevBuf flags, starts, nsel;
  RS_TRY(flags.alloc(sizeof(int32_t) * (size_t)E3));
  RS_TRY(starts.alloc(sizeof(int32_t) * (size_t)E3));
  RS_TRY(nsel.alloc(16));
  const int g = grid_for(E3, 256, 16);
  tc_segment_flags_kernel<<<g, 256, 0, st>>>(p->eg.as<int32_t>(), E3, flags.as<int32_t>());
  size_t bytes = 0;
  rocprim::counting_iterator<int32_t> idx(0);
  RS_ROC(rocprim::select(nullptr, bytes, idx, flags.as<int32_t>(), starts.as<int32_t>(),
                         nsel.as<size_t>(), (size_t)E3, st));
  if (tmp.bytes < bytes) RS_TRY(tmp.alloc(bytes));
  RS_ROC(rocprim::select(tmp.p, bytes, idx, flags.as<int32_t>(), starts.as<int32_t>(),
                         nsel.as<size_t>(), (size_t)E3, st));
  size_t nseg = 0;
  RS_HIP(hipMemcpyAsync(&nseg, nsel.p, sizeof(size_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  flags.release();
  DevBuf sg, sL, swb, swe;
  RS_TRY(sg.alloc(sizeof(int32_t) * nseg));
  RS_TRY(sL.alloc(sizeof(int32_t) * nseg));
  RS_TRY(swb.alloc(sizeof(int64_t) * nseg));
  RS_TRY(swe.alloc(sizeof(int64_t) * nseg));
  tc_segment_info_kernel<<<grid_for((int64_t)nseg, 256, 8), 256, 0, st>>>(
      starts.as<int32_t>(), (int64_t)nseg, E3, p->eg.as<int32_t>(), t->lengths.as<int32_t>(),
      p->W.as<int64_t>(), sg.as<int32_t>(), sL.as<int32_t>(), swb.as<int64_t>(), swe.as<int64_t>());
  std::vector<int32_t> h_start(nseg), h_g(nseg), h_L(nseg);
  std::vector<int64_t> h_wb(nseg), h_we(nseg);
  RS_HIP(hipMemcpyAsync(h_start.data(), starts.p, sizeof(int32_t) * nseg, hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(h_g.data(), sg.p, sizeof(int32_t) * nseg, hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(h_L.data(), sL.p, sizeof(int32_t) * nseg, hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(h_wb.data(), swb.p, sizeof(int64_t) * nseg, hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(h_we.data(), swe.p, sizeof(int64_t) * nseg, hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));

  std::vector<int32_t> ug, ue0, ue1;
  std::vector<int64_t> uwb, uwe, utile(1, 0);
  std::vector<int64_t> ucost;  // probes + copied slots, for the XCD split
  int cls_prev = -1;
  for (size_t i = 0; i < nseg; ++i) {
    const int c = tc_class_of((uint32_t)h_L[i]);
    RS_REQUIRE(c >= cls_prev, "rs_tc_plan_create: staged schedule out of class order");
    for (int q = cls_prev + 1; q <= c; ++q) p->class_units[q] = (int64_t)ug.size();
    cls_prev = c;
    if (h_L[i] > p->class_lmax[c]) p->class_lmax[c] = h_L[i];
    const int64_t slots = (int64_t)h_L[i] * t->b;
    const int64_t P = h_we[i] - h_wb[i];
    const int64_t part = std::max(kUnitProbes, kUnitSlotFactor * slots);
    const int64_t k = (P + part - 1) / part;
    const int32_t e1 = i + 1 < nseg ? h_start[i + 1] : (int32_t)E3;
    for (int64_t j = 0; j < k; ++j) {
      const int64_t a = h_wb[i] + P * j / k, z = h_wb[i] + P * (j + 1) / k;
      ug.push_back(h_g[i]);
      uwb.push_back(a);
      uwe.push_back(z);
      ue0.push_back(h_start[i]);
      ue1.push_back(e1);
      utile.push_back(utile.back() + (z - a + kStageTile - 1) / kStageTile);
      ucost.push_back(z - a + slots);
    }
  }
  for (int q = cls_prev + 1; q <= kTcStagedClasses; ++q) p->class_units[q] = (int64_t)ug.size();
  p->num_units = (int64_t)ug.size();
  p->num_tiles = utile.back();
  RS_REQUIRE(p->num_tiles < (1ll << 31), "rs_tc_plan_create: too many staged tiles");
  // per class: 8 contiguous unit ranges of about equal cost
  std::vector<int64_t> xsplit((size_t)kTcStagedClasses * (kXcds + 1));
  for (int c = 0; c < kTcStagedClasses; ++c) {
    const int64_t u0 = p->class_units[c], u1 = p->class_units[c + 1];
    int64_t total = 0;
    for (int64_t u = u0; u < u1; ++u) total += ucost[(size_t)u];
    int64_t* xs = &xsplit[(size_t)c * (kXcds + 1)];
    xs[0] = u0;
    int64_t acc = 0, u = u0;
    for (int x = 1; x < kXcds; ++x) {
      const int64_t target = total * x / kXcds;
      while (u < u1 && acc + ucost[(size_t)u] / 2 <= target) acc += ucost[(size_t)u++];
      xs[x] = u;
    }
    xs[kXcds] = u1;
  }
  const size_t nu = ug.size();
  RS_TRY(p->u_g.alloc(sizeof(int32_t) * nu));
  RS_TRY(p->u_wb.alloc(sizeof(int64_t) * nu));
  RS_TRY(p->u_we.alloc(sizeof(int64_t) * nu));
  RS_TRY(p->u_tile.alloc(sizeof(int64_t) * (nu + 1)));
  RS_TRY(p->tile_edge.alloc(sizeof(int32_t) * (size_t)p->num_tiles));
  RS_TRY(p->xcd_units.alloc(sizeof(int64_t) * xsplit.size()));
  RS_TRY(p->unit_next.alloc(sizeof(unsigned long long) * kTcStagedClasses * kXcds));
  DevBuf de0, de1;
  RS_TRY(de0.alloc(sizeof(int32_t) * nu));
  RS_TRY(de1.alloc(sizeof(int32_t) * nu));
  RS_HIP(hipMemcpyAsync(p->u_g.p, ug.data(), sizeof(int32_t) * nu, hipMemcpyHostToDevice, st));
  RS_HIP(hipMemcpyAsync(p->u_wb.p, uwb.data(), sizeof(int64_t) * nu, hipMemcpyHostToDevice, st));
  RS_HIP(hipMemcpyAsync(p->u_we.p, uwe.data(), sizeof(int64_t) * nu, hipMemcpyHostToDevice, st));
  RS_HIP(hipMemcpyAsync(p->u_tile.p, utile.data(), sizeof(int64_t) * (nu + 1), hipMemcpyHostToDevice, st));
  RS_HIP(hipMemcpyAsync(p->xcd_units.p, xsplit.data(), sizeof(int64_t) * xsplit.size(),
                        hipMemcpyHostToDevice, st));
  RS_HIP(hipMemcpyAsync(de0.p, ue0.data(), sizeof(int32_t) * nu, hipMemcpyHostToDevice, st));
  RS_HIP(hipMemcpyAsync(de1.p, ue1.data(), sizeof(int32_t) * nu, hipMemcpyHostToDevice, st));
  tc_tile_edges_kernel<<<grid_for((int64_t)nu * kWave, 256, 8), 256, 0, st>>>(
      p->u_wb.as<int64_t>(), p->u_tile.as<int64_t>(), de0.as<int32_t>(), de1.as<int32_t>(),
      (int64_t)nu, p->W.as<int64_t>(), p->tile_edge.as<int32_t>());
  RS_HIP(hipStreamSynchronize(st));  // host vectors and de0/de1 die here
  return RS_OK;
}

}  // namespace rs

using namespace rs;

extern "C" {

static rs_status tc_plan_build(const rs_tables* t, hipStream_t st, rs_tc_plan* p) {
  const int64_t E = t->E;
  RS_REQUIRE(E < (1ll << 31), "rs_tc_plan_create: at most 2^31-1 edges");
  RS_TRY(p->W.alloc(sizeof(int64_t) * (size_t)(E + 1)));
  RS_TRY(p->counter.alloc(16));
  RS_TRY(p->es.alloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)));
  RS_TRY(p->eg.alloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)));
  for (int c = 0; c <= kTcGroups; ++c) p->bounds[c] = E;
  if (E == 0) {
    RS_HIP(hipMemsetAsync(p->W.p, 0, sizeof(int64_t), st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
  }
  DevBuf P, keys, sorted, tmp, w, bounds;
  const int g = grid_for(E, 256, 16);
  RS_TRY(P.alloc(sizeof(unsigned long long) * (size_t)t->n));
  RS_HIP(hipMemsetAsync(P.p, 0, sizeof(unsigned long long) * (size_t)t->n, st));
  tc_table_probes_kernel<<<g, 256, 0, st>>>(t->rows.as<int32_t>(), t->d_elems, t->d_off, t->n, E,
                                           P.as<unsigned long long>());
  RS_TRY(keys.alloc(sizeof(uint64_t) * (size_t)E));
  tc_edge_keys_kernel<<<g, 256, 0, st>>>(t->rows.as<int32_t>(), t->d_elems, t->d_off, t->n, E,
                                        t->lengths.as<int32_t>(), P.as<unsigned long long>(), t->b,
                                        stage_lmax(t->b), keys.as<uint64_t>(), t->d_status());
  P.release();
  RS_TRY(sorted.alloc(sizeof(uint64_t) * (size_t)E));
  size_t bytes = 0;
  RS_ROC(rocprim::radix_sort_keys(nullptr, bytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                  (size_t)E, 0, 64, st));
  RS_TRY(tmp.alloc(bytes));
  RS_ROC(rocprim::radix_sort_keys(tmp.p, bytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                  (size_t)E, 0, 64, st));
  keys.release();
  RS_TRY(w.alloc(sizeof(int32_t) * (size_t)E));
  RS_TRY(bounds.alloc(sizeof(int64_t) * kTcGroups));
  RS_HIP(hipMemcpyAsync(bounds.p, p->bounds, sizeof(int64_t) * kTcGroups, hipMemcpyHostToDevice, st));
  tc_edge_split_kernel<<<g, 256, 0, st>>>(sorted.as<uint64_t>(), t->d_off, E, p->es.as<int32_t>(),
                                         p->eg.as<int32_t>(), w.as<int32_t>(), bounds.as<int64_t>());
  sorted.release();
  RS_TRY(exclusive_scan_i32_to_i64(w.as<int32_t>(), E, p->W.as<int64_t>(), tmp, st));
  RS_HIP(hipMemcpyAsync(p->bounds, bounds.p, sizeof(int64_t) * kTcGroups, hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  for (int c = 0; c <= kTcGroups; ++c)
    RS_HIP(hipMemcpyAsync(&p->wb[c], p->W.as<int64_t>() + p->bounds[c], sizeof(int64_t),
                          hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  p->total = p->wb[kTcGroups];
  rs_status s = rs_tables_status(const_cast<rs_tables*>(t), reinterpret_cast<rs_stream>(st));
  if (s != RS_OK) {
    set_error("rs_tc_plan_create: an edge endpoint is outside [0, num_sets)");
    return RS_EINVAL;
  }
  return tc_build_units(t, st, p, tmp);
}

rs_status rs_tc_plan_create(const rs_tables* t, rs_stream stream, rs_tc_plan** out) {
  RS_REQUIRE(t && out, "rs_tc_plan_create: null argument");
  *out = nullptr;
  auto* p = new rs_tc_plan();
  p->t = t;
  p->E = t->E;
  rs_status s = tc_plan_build(t, as_stream(stream), p);
  if (s != RS_OK) {
    delete p;
    return s;
  }
  *out = p;
  return RS_OK;
}

void rs_tc_plan_free(rs_tc_plan* p) { delete p; }

rs_status rs_tc_plan_work(const rs_tc_plan* p, int64_t* h_probes, int64_t* h_edges) {
  RS_REQUIRE(p, "rs_tc_plan_work: null plan");
  if (h_probes) *h_probes = p->total;
  if (h_edges) *h_edges = p->E;
  return RS_OK;
}

rs_status rs_tc_plan_staged(const rs_tc_plan* p, int64_t* h_staged_probes, int64_t* h_units,
                            int64_t* h_tables_max_len) {
  RS_REQUIRE(p, "rs_tc_plan_staged: null plan");
  if (h_staged_probes) *h_staged_probes = p->wb[3];
  if (h_units) *h_units = p->num_units;
  if (h_tables_max_len) {
    int32_t m = 0;
    for (int c = 0; c < kTcStagedClasses; ++c) m = p->class_lmax[c] > m ? p->class_lmax[c] : m;
    *h_tables_max_len = m;
  }
  return RS_OK;
}

}  // extern "C"

namespace rs {

template <int LAYOUT, int NW, int B>
static rs_status launch_staged(const rs_tables* t, const rs_tc_plan* p, int c, const int32_t* T,
                               const int32_t* keys, unsigned long long* out, hipStream_t st) {
  const int64_t L = p->class_lmax[c];
  const size_t lds = stage_lds_bytes(c, L, t->b);
  const int32_t words = (int32_t)(L * t->b);
  auto kern = tc_staged_kernel<LAYOUT, NW, kStageU, B>;
  RS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = (int)(kLdsBytes / lds);
  const int wave_cap = 32 / NW;
  per_cu = per_cu < wave_cap ? per_cu : wave_cap;
  per_cu = per_cu < 1 ? 1 : per_cu;
  int grid = device_cu_count() * per_cu;
  const int64_t units = p->class_units[c + 1] - p->class_units[c];
  if (grid > units * kXcds) grid = (int)(units * kXcds);
  grid = (grid + kXcds - 1) / kXcds * kXcds;
  kern<<<grid, NW * kWave, lds, st>>>(
      p->es.as<int32_t>(), t->d_off, keys, t->lengths.as<int32_t>(), t->toff.as<int64_t>(), T,
      t->BN, t->b, (uint32_t)t->M, p->W.as<int64_t>(), p->E, p->u_g.as<int32_t>(),
      p->u_wb.as<int64_t>(), p->u_we.as<int64_t>(), p->u_tile.as<int64_t>(),
      p->tile_edge.as<int32_t>(), p->xcd_units.as<int64_t>() + c * (kXcds + 1), words, out,
      t->d_status(), p->unit_next.as<unsigned long long>() + c * kXcds);
  return RS_OK;
}

// Bucket sizes with a bucket-at-a-time kernel; others take the slot walk.
template <int LAYOUT, int NW>
static rs_status launch_staged_b(const rs_tables* t, const rs_tc_plan* p, int c, const int32_t* T,
                                 const int32_t* keys, unsigned long long* out, hipStream_t st) {
  switch (t->b) {
    case 1: return launch_staged<LAYOUT, NW, 1>(t, p, c, T, keys, out, st);
    case 2: return launch_staged<LAYOUT, NW, 2>(t, p, c, T, keys, out, st);
    case 4: return launch_staged<LAYOUT, NW, 4>(t, p, c, T, keys, out, st);
    case 5: return launch_staged<LAYOUT, NW, 5>(t, p, c, T, keys, out, st);
    case 8: return launch_staged<LAYOUT, NW, 8>(t, p, c, T, keys, out, st);
    default: return launch_staged<LAYOUT, NW, 0>(t, p, c, T, keys, out, st);
  }
}

template <int LAYOUT>
static rs_status tc_count_layout(const rs_tables* t, const rs_tc_plan* p, const int32_t* T,
                                 const int32_t* keys, unsigned long long* out, hipStream_t st) {
  if (p->class_units[kTcStagedClasses] > p->class_units[0])
    zero_u64_n<<<1, 64, 0, st>>>(p->unit_next.as<unsigned long long>(), kTcStagedClasses * kXcds);
  if (p->class_units[1] > p->class_units[0])
    RS_TRY((launch_staged_b<LAYOUT, kClassWaves[0]>(t, p, 0, T, keys, out, st)));
  if (p->class_units[2] > p->class_units[1])
    RS_TRY((launch_staged_b<LAYOUT, kClassWaves[1]>(t, p, 1, T, keys, out, st)));
  if (p->class_units[3] > p->class_units[2])
    RS_TRY((launch_staged_b<LAYOUT, kClassWaves[2]>(t, p, 2, T, keys, out, st)));
  const int64_t w_lo = p->wb[3], w_hi = p->wb[4];
  if (w_hi > w_lo) {
    // tiles of kTcTile probes, smaller when that would leave waves of the
    // grid without a tile (a small graph's few unstaged probes are chains
    // of dependent HBM reads: spread them over the whole device)
    const int64_t waves = (int64_t)device_cu_count() * 8 * (kTcBlock / kWave);
    int64_t tile = kTcTile;
    while (tile > kTcTileMin && (w_hi - w_lo) / tile < waves) tile >>= 1;
    const int64_t ntiles = (w_hi - w_lo + tile - 1) / tile;
    int grid = grid_for(ntiles * kWave, kTcBlock, 8);
    grid = (grid + kXcds - 1) / kXcds * kXcds;
    tc_tiled_kernel<LAYOUT, 4><<<grid, kTcBlock, 0, st>>>(
        p->es.as<int32_t>(), p->eg.as<int32_t>(), t->d_off, keys, t->lengths.as<int32_t>(),
        t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->W.as<int64_t>(), p->bounds[3],
        p->bounds[4], w_lo, w_hi, tile, out, t->d_status());
  }
  return RS_OK;
}

}  // namespace rs

extern "C" {

rs_status rs_tc_count_async(const rs_tables* t, const rs_tc_plan* p, int32_t layout,
                            const int32_t* d_probe_keys, uint64_t* d_count, rs_stream stream) {
  const int32_t* T = nullptr;
  RS_TRY(check_args(t, p, layout, d_count, &T));
  hipStream_t st = as_stream(stream);
  auto* out = reinterpret_cast<unsigned long long*>(d_count);
  zero_u64<<<1, 1, 0, st>>>(out);
  const int32_t* keys = d_probe_keys ? d_probe_keys : t->d_elems;
  if (t->timing) RS_HIP(hipEventRecord(t->ev0, st));
  if (layout == RS_HIERARCHICAL)
    RS_TRY(tc_count_layout<RS_HIERARCHICAL>(t, p, T, keys, out, st));
  else
    RS_TRY(tc_count_layout<RS_NORMAL>(t, p, T, keys, out, st));
  if (t->timing) RS_HIP(hipEventRecord(t->ev1, st));
  RS_HIP(hipGetLastError());
  return RS_OK;
}

rs_status rs_tc_count_plain_async(const rs_tables* t, const rs_tc_plan* p, int32_t layout,
                                  const int32_t* d_probe_keys, uint64_t* d_count,
                                  rs_stream stream) {
  const int32_t* T = nullptr;
  RS_TRY(check_args(t, p, layout, d_count, &T));
  hipStream_t st = as_stream(stream);
  auto* out = reinterpret_cast<unsigned long long*>(d_count);
  zero_u64<<<1, 1, 0, st>>>(out);
  if (p->E > 0) {
    const int32_t chunk = 4;
    const int grid = device_cu_count() * 4;
    const int64_t waves = (int64_t)grid * (kTcBlock / kWave);
    RS_REQUIRE(waves * chunk < (1ll << 31) && p->E < (1ll << 31) - waves * chunk,
               "tc plain: too many edges for the int32 work counter");
    set_i32<<<1, 1, 0, st>>>(p->counter.as<int32_t>(), (int32_t)(waves * chunk));
    const int32_t* keys = d_probe_keys ? d_probe_keys : t->d_elems;
    if (layout == RS_HIERARCHICAL)
      tc_plain_kernel<RS_HIERARCHICAL><<<grid, kTcBlock, 0, st>>>(
          t->rows.as<int32_t>(), t->d_elems, t->d_off, keys, t->lengths.as<int32_t>(),
          t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->E, chunk,
          p->counter.as<int32_t>(), out, t->d_status());
    else
      tc_plain_kernel<RS_NORMAL><<<grid, kTcBlock, 0, st>>>(
          t->rows.as<int32_t>(), t->d_elems, t->d_off, keys, t->lengths.as<int32_t>(),
          t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->E, chunk,
          p->counter.as<int32_t>(), out, t->d_status());
  }
  RS_HIP(hipGetLastError());
  return RS_OK;
}

rs_status rs_tc_count(rs_tables* t, int32_t layout, const int32_t* d_probe_keys,
                      uint64_t* h_count, rs_stream stream) {
  RS_REQUIRE(t && h_count, "rs_tc_count: null argument");
  std::lock_guard<std::mutex> lk(t->tc_mu);
  if (!t->tc_plan) RS_TRY(rs_tc_plan_create(t, stream, &t->tc_plan));
  DevBuf out;
  RS_TRY(out.alloc(sizeof(uint64_t)));
  RS_TRY(rs_tc_count_async(t, t->tc_plan, layout, d_probe_keys, out.as<uint64_t>(), stream));
  hipStream_t st = as_stream(stream);
  RS_HIP(hipMemcpyAsync(h_count, out.p, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  return rs_tables_status(t, stream);
}

void rs_tables_free(rs_tables* t) {
  if (!t) return;
  delete t->tc_plan;
  delete t;
}

}  // extern "C"
