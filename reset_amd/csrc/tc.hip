// tc.hip -- triangle counting by hash probes (TC.cu:128-190, baseline/tc/tc.cu).
//
// count = sum over oriented edges (u, v) of |N+(u) cap N+(v)|, each edge's
// smaller out-list probing the larger one's table (TC.cu:146-155).
//
// Tiled form (the default): the plan stores W = exclusive prefix of per-edge
// probe counts w_e = min(d+(u), d+(v)).  The flattened probe space [0, W[E])
// is cut into equal tiles; a wave takes a tile, finds its first edge by binary
// search in W, loads 64 edges' metadata (one per lane), scans their clipped
// probe counts and walks the flattened probes 64 at a time, each lane locating
// its edge with a 6-step shuffle search.  Every lane is busy regardless of
// degree (mean d+ ~ 16 would idle 3/4 of a wave64 in the per-edge form), and
// hub edges are split across waves by construction.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "tables.hpp"

// Edge schedule.  Every oriented edge becomes (s, g): s the endpoint with the
// smaller out-list (its keys are probed), g the other (its table is probed);
// ties keep the CSR source as s (TC.cu:146-155).  Edges are ordered by
// (g, s), so all probes into one table run back to back and hit in L2, and W
// is the exclusive prefix of the per-edge probe counts d+(s).
struct rs_tc_plan {
  const rs_tables* t = nullptr;
  int64_t E = 0;
  int64_t total = 0;          // W[E]
  rs::DevBuf W;               // int64[E+1]
  rs::DevBuf es, eg;          // int32[E]: probing / probed endpoint
  rs::DevBuf counter;         // work-stealing counter of the plain form
  int64_t E_light = 0;        // edges [0, E_light) have < kHeavy probes
  int64_t W_light = 0;        // W[E_light]
};

namespace rs {

constexpr int kTcBlock = 256;
constexpr int64_t kTcTile = 4096;
// Edges with at least this many probes are "heavy": one wave covers them 64
// keys at a time with the edge's metadata in scalar registers.  Lighter edges
// are flattened 64 edges at a time.
constexpr int32_t kHeavy = 64;

__global__ void zero_u64(unsigned long long* p) { *p = 0ull; }
__global__ void set_i32(int32_t* p, int32_t v) { *p = v; }

// key[e] = g << 32 | s for CSR edge e = (rows[e], cols[e]); an endpoint
// outside [0, n) is malformed input (flagged, and the edge becomes (0, 0)
// with no work), so no later kernel indexes the offsets with it.
__global__ void tc_edge_keys_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                                    const int32_t* __restrict__ off, int64_t n, int64_t E,
                                    uint64_t* __restrict__ key, int32_t* status) {
  bool bad = false;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = rows[e], v = cols[e];
    if (v < 0 || v >= n) {
      bad = true;
      key[e] = ~0ull;
      continue;
    }
    const int32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
    const uint32_t sm = du > dv ? v : u, lg = du > dv ? u : v;
    const uint64_t heavy = (du < dv ? du : dv) >= kHeavy;
    key[e] = (heavy << 63) | ((uint64_t)lg << 32) | sm;
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

__global__ void tc_edge_split_kernel(const uint64_t* __restrict__ key, const int32_t* __restrict__ off,
                                     int64_t n, int64_t E, int32_t* __restrict__ es,
                                     int32_t* __restrict__ eg, int32_t* __restrict__ w) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = key[e];
    if (k == ~0ull) {
      es[e] = 0;
      eg[e] = 0;
      w[e] = 0;
      continue;
    }
    const int32_t sm = (int32_t)(k & 0xFFFFFFFFull), lg = (int32_t)((k >> 32) & 0x7FFFFFFFull);
    es[e] = sm;
    eg[e] = lg;
    w[e] = off[sm + 1] - off[sm];
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

// Workgroups are dealt to the 8 XCDs round-robin (block b on XCD b % 8), and
// each XCD has its own L2, so XCD x takes the contiguous tile range
// [x*R, (x+1)*R): neighbouring tiles, which probe the same tables, share an L2.
constexpr int kXcds = 8;

template <int LAYOUT, int UNROLL>
__global__ void __launch_bounds__(kTcBlock) tc_tiled_kernel(
    const int32_t* __restrict__ es, const int32_t* __restrict__ eg,
    const int32_t* __restrict__ off, const int32_t* __restrict__ keys,
    const int32_t* __restrict__ Ls, const int64_t* __restrict__ toff,
    const int32_t* __restrict__ T, int64_t BN, int32_t b, uint32_t M,
    const int64_t* __restrict__ W, int64_t E, int64_t total, int64_t tile,
    unsigned long long* __restrict__ out, int32_t* __restrict__ status) {
  const int lane = __lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int xcd = blockIdx.x % kXcds;
  const int64_t local_wave = (int64_t)(blockIdx.x / kXcds) * (kTcBlock / kWave) + wid;
  const int64_t xcd_waves = (int64_t)(gridDim.x / kXcds) * (kTcBlock / kWave);
  const int64_t ntiles = (total + tile - 1) / tile;
  const int64_t per_xcd = (ntiles + kXcds - 1) / kXcds;
  const int64_t t_begin = xcd * per_xcd;
  const int64_t t_end = t_begin + per_xcd < ntiles ? t_begin + per_xcd : ntiles;
  uint32_t found = 0;
  bool bad = false;
  for (int64_t ti = t_begin + local_wave; ti < t_end; ti += xcd_waves) {
    const int64_t ts = ti * tile;
    const int64_t te = ts + tile < total ? ts + tile : total;
    // largest e with W[e] <= ts (that edge has work, since ts < total)
    int64_t lo = 0, hi = E;
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
      if (ei < E) {
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
      if (e >= E) break;
      if (W[e] >= te) break;
    }
  }
  const uint32_t s = wave_sum(found);
  if (lane == 0 && s) atomicAdd(out, (unsigned long long)s);
  if (__any(bad) && lane == 0) atomicOr(status, kErrProbeBound);
}

// Slot-plane probe of U keys of one lane against one table, in lockstep:
// each step reads the current slot of every pending key (TC.cu:11-31 order:
// slots 0..b-1 of the home bucket, then the next bucket), so a wave pays the
// longest chain among its keys rather than the sum.
template <int LAYOUT, int U>
__device__ __forceinline__ uint32_t probe_lockstep(const int32_t* __restrict__ Tt, int64_t BN,
                                                   uint32_t L, int32_t b, uint32_t M,
                                                   const int32_t (&x)[U], const bool (&act)[U],
                                                   bool& bad) {
  uint32_t bucket[U];
  int32_t k[U];
  bool pend[U];
  uint32_t hits = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    bucket[u] = home_bucket<LAYOUT>((uint32_t)x[u], L, M);
    k[u] = 0;
    pend[u] = act[u];
  }
  const uint32_t bound = L * (uint32_t)b;
  for (uint32_t step = 0; step < bound; ++step) {
    int32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = Tt[bucket[u] + (int64_t)k[u] * BN];
    bool left = false;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool hit = pend[u] && v[u] == x[u];
      const bool stop = hit || v[u] == kEmpty;
      hits += hit;
      pend[u] = pend[u] && !stop;
      const int32_t nk = k[u] + 1;
      const bool wrap = nk == b;
      k[u] = pend[u] ? (wrap ? 0 : nk) : k[u];
      bucket[u] = pend[u] && wrap ? ((bucket[u] + 1u) & (L - 1u)) : bucket[u];
      left |= pend[u];
    }
    if (!__any(left)) return hits;
  }
  bad = true;
  return hits;
}

// Heavy edges (>= kHeavy probes each), ordered by probed table.  Tiles of the
// probe space [w_lo, w_hi) are dealt to XCDs in contiguous ranges; inside a
// tile a wave walks its edges one at a time with the edge's endpoints,
// offsets and table geometry in scalar registers, lanes taking 64*U
// consecutive keys of the probing list per step.
template <int LAYOUT, int U>
__global__ void __launch_bounds__(kTcBlock) tc_heavy_kernel(
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
    int64_t lo = e_lo, hi = e_hi;  // largest e with W[e] <= ts
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (W[mid] <= ts) lo = mid; else hi = mid;
    }
    for (int64_t e = lo; e < e_hi; ++e) {
      const int64_t ws = W[e];
      if (ws >= te) break;
      const int64_t we = W[e + 1];
      const int32_t sv = es[e], gv = eg[e];
      const int64_t kb = off[sv];
      const int32_t* Tt = T + toff[gv];
      const uint32_t L = (uint32_t)Ls[gv];
      const int32_t j0 = (int32_t)((ws > ts ? ws : ts) - ws);
      const int32_t j1 = (int32_t)((we < te ? we : te) - ws);
      for (int32_t j = j0; j < j1; j += kWave * U) {
        int32_t x[U];
        bool act[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t jj = j + u * kWave + lane;
          act[u] = jj < j1;
          x[u] = act[u] ? keys[kb + jj] : kEmpty;
        }
        found += probe_lockstep<LAYOUT, U>(Tt, BN, L, b, M, x, act, bad);
      }
    }
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

__global__ void count_light_kernel(const uint64_t* __restrict__ key, int64_t E,
                                   unsigned long long* cnt) {
  uint32_t c = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    c += (key[e] >> 63) == 0;
  c = wave_sum(c);
  if (__lane_id() == 0 && c) atomicAdd(cnt, (unsigned long long)c);
}

static rs_status check_args(const rs_tables* t, const rs_tc_plan* p, int32_t layout,
                            uint64_t* d_count, const int32_t** T) {
  RS_REQUIRE(t && p && d_count, "tc: null argument");
  RS_REQUIRE(p->t == t, "tc: plan was made for another tables handle");
  RS_TRY(require_layout(t, layout, T));
  return RS_OK;
}

}  // namespace rs

using namespace rs;

extern "C" {

static rs_status tc_plan_build(const rs_tables* t, hipStream_t st, rs_tc_plan* p) {
  RS_TRY(p->W.alloc(sizeof(int64_t) * (size_t)(t->E + 1)));
  RS_TRY(p->counter.alloc(16));
  RS_TRY(p->es.alloc(sizeof(int32_t) * (size_t)(t->E > 0 ? t->E : 1)));
  RS_TRY(p->eg.alloc(sizeof(int32_t) * (size_t)(t->E > 0 ? t->E : 1)));
  if (t->E == 0) {
    RS_HIP(hipMemsetAsync(p->W.p, 0, sizeof(int64_t), st));
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
  }
  DevBuf keys, sorted, tmp, w;
  RS_TRY(keys.alloc(sizeof(uint64_t) * (size_t)t->E));
  RS_TRY(sorted.alloc(sizeof(uint64_t) * (size_t)t->E));
  const int g = grid_for(t->E, 256, 16);
  tc_edge_keys_kernel<<<g, 256, 0, st>>>(t->rows.as<int32_t>(), t->d_elems, t->d_off, t->n, t->E,
                                        keys.as<uint64_t>(), t->d_status());
  size_t bytes = 0;
  if (rocprim::radix_sort_keys(nullptr, bytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                               (size_t)t->E, 0, 64, st) != hipSuccess) {
    set_error("rs_tc_plan_create: radix sort sizing failed");
    return RS_EDEVICE;
  }
  RS_TRY(tmp.alloc(bytes));
  if (rocprim::radix_sort_keys(tmp.p, bytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                               (size_t)t->E, 0, 64, st) != hipSuccess) {
    set_error("rs_tc_plan_create: radix sort failed");
    return RS_EDEVICE;
  }
  keys.release();
  RS_TRY(w.alloc(sizeof(int32_t) * (size_t)t->E));
  tc_edge_split_kernel<<<g, 256, 0, st>>>(sorted.as<uint64_t>(), t->d_off, t->n, t->E,
                                         p->es.as<int32_t>(), p->eg.as<int32_t>(), w.as<int32_t>());
  RS_TRY(exclusive_scan_i32_to_i64(w.as<int32_t>(), t->E, p->W.as<int64_t>(), tmp, st));
  RS_HIP(hipMemcpyAsync(&p->total, p->W.as<int64_t>() + t->E, sizeof(int64_t),
                        hipMemcpyDeviceToHost, st));
  // light edges sort first (heavy flag in the key's top bit): count them
  {
    DevBuf cnt;
    RS_TRY(cnt.alloc(16));
    RS_HIP(hipMemsetAsync(cnt.p, 0, 16, st));
    count_light_kernel<<<g, 256, 0, st>>>(sorted.as<uint64_t>(), t->E, cnt.as<unsigned long long>());
    unsigned long long hl = 0;
    RS_HIP(hipMemcpyAsync(&hl, cnt.p, sizeof(hl), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    p->E_light = (int64_t)hl;
    RS_HIP(hipMemcpyAsync(&p->W_light, p->W.as<int64_t>() + p->E_light, sizeof(int64_t),
                          hipMemcpyDeviceToHost, st));
  }
  RS_HIP(hipStreamSynchronize(st));
  rs_status s = rs_tables_status(const_cast<rs_tables*>(t), reinterpret_cast<rs_stream>(st));
  if (s != RS_OK) {
    set_error("rs_tc_plan_create: an edge endpoint is outside [0, num_sets)");
    return RS_EINVAL;
  }
  return RS_OK;
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

rs_status rs_tc_count_async(const rs_tables* t, const rs_tc_plan* p, int32_t layout,
                            const int32_t* d_probe_keys, uint64_t* d_count, rs_stream stream) {
  const int32_t* T = nullptr;
  RS_TRY(check_args(t, p, layout, d_count, &T));
  hipStream_t st = as_stream(stream);
  auto* out = reinterpret_cast<unsigned long long*>(d_count);
  zero_u64<<<1, 1, 0, st>>>(out);
  const int32_t* keys = d_probe_keys ? d_probe_keys : t->d_elems;
  if (p->W_light > 0) {
    const int64_t ntiles = (p->W_light + kTcTile - 1) / kTcTile;
    int grid = grid_for(ntiles * kWave, kTcBlock, 8);
    grid = (grid + kXcds - 1) / kXcds * kXcds;
    if (layout == RS_HIERARCHICAL)
      tc_tiled_kernel<RS_HIERARCHICAL, 4><<<grid, kTcBlock, 0, st>>>(
          p->es.as<int32_t>(), p->eg.as<int32_t>(), t->d_off, keys, t->lengths.as<int32_t>(),
          t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->W.as<int64_t>(), p->E_light,
          p->W_light, kTcTile, out, t->d_status());
    else
      tc_tiled_kernel<RS_NORMAL, 4><<<grid, kTcBlock, 0, st>>>(
          p->es.as<int32_t>(), p->eg.as<int32_t>(), t->d_off, keys, t->lengths.as<int32_t>(),
          t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->W.as<int64_t>(), p->E_light,
          p->W_light, kTcTile, out, t->d_status());
  }
  if (p->total > p->W_light) {
    const int64_t ntiles = (p->total - p->W_light + kTcTile - 1) / kTcTile;
    int grid = grid_for(ntiles * kWave, kTcBlock, 8);
    grid = (grid + kXcds - 1) / kXcds * kXcds;
    if (layout == RS_HIERARCHICAL)
      tc_heavy_kernel<RS_HIERARCHICAL, 2><<<grid, kTcBlock, 0, st>>>(
          p->es.as<int32_t>(), p->eg.as<int32_t>(), t->d_off, keys, t->lengths.as<int32_t>(),
          t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->W.as<int64_t>(), p->E_light,
          p->E, p->W_light, p->total, kTcTile, out, t->d_status());
    else
      tc_heavy_kernel<RS_NORMAL, 2><<<grid, kTcBlock, 0, st>>>(
          p->es.as<int32_t>(), p->eg.as<int32_t>(), t->d_off, keys, t->lengths.as<int32_t>(),
          t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->W.as<int64_t>(), p->E_light,
          p->E, p->W_light, p->total, kTcTile, out, t->d_status());
  }
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
