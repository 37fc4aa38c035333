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
#include "tables.hpp"

struct rs_tc_plan {
  const rs_tables* t = nullptr;
  int64_t E = 0;
  int64_t total = 0;          // W[E]
  rs::DevBuf W;               // int64[E+1]
  rs::DevBuf counter;         // work-stealing counter of the plain form
};

namespace rs {

constexpr int kTcBlock = 256;
constexpr int64_t kTcTile = 4096;

__global__ void zero_u64(unsigned long long* p) { *p = 0ull; }
__global__ void set_i32(int32_t* p, int32_t v) { *p = v; }

// w[e] = min(d+(u), d+(v)); an endpoint outside [0, n) is malformed input
// (flagged, w = 0), so no later kernel indexes the offsets with it.
__global__ void tc_work_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                               const int32_t* __restrict__ off, int64_t n, int64_t E,
                               int32_t* __restrict__ w, int32_t* status) {
  bool bad = false;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = rows[e], v = cols[e];
    if (v < 0 || v >= n) {
      bad = true;
      w[e] = 0;
      continue;
    }
    const int32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
    w[e] = du < dv ? du : dv;
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
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

template <int LAYOUT, int UNROLL>
__global__ void __launch_bounds__(kTcBlock) tc_tiled_kernel(
    const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
    const int32_t* __restrict__ off, const int32_t* __restrict__ keys,
    const int32_t* __restrict__ Ls, const int64_t* __restrict__ toff,
    const int32_t* __restrict__ T, int64_t BN, int32_t b, uint32_t M,
    const int64_t* __restrict__ W, int64_t E, int64_t total, int64_t tile,
    unsigned long long* __restrict__ out, int32_t* __restrict__ status) {
  const int lane = __lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  const int64_t ntiles = (total + tile - 1) / tile;
  uint32_t found = 0;
  bool bad = false;
  for (int64_t ti = wave; ti < ntiles; ti += nwaves) {
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
            int32_t u = rows[ei], v = cols[ei];
            const int32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
            if (du > dv) {
              const int32_t x = u;
              u = v;
              v = x;
            }
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

}  // namespace rs

using namespace rs;

extern "C" {

rs_status rs_tc_plan_create(const rs_tables* t, rs_stream stream, rs_tc_plan** out) {
  RS_REQUIRE(t && out, "rs_tc_plan_create: null argument");
  *out = nullptr;
  hipStream_t st = as_stream(stream);
  auto* p = new rs_tc_plan();
  p->t = t;
  p->E = t->E;
  rs_status s;
  if ((s = p->W.alloc(sizeof(int64_t) * (size_t)(t->E + 1))) != RS_OK ||
      (s = p->counter.alloc(16)) != RS_OK) {
    delete p;
    return s;
  }
  if (t->E > 0) {
    DevBuf w, tmp;
    if ((s = w.alloc(sizeof(int32_t) * (size_t)t->E)) != RS_OK) {
      delete p;
      return s;
    }
    tc_work_kernel<<<grid_for(t->E, 256, 16), 256, 0, st>>>(t->rows.as<int32_t>(), t->d_elems,
                                                            t->d_off, t->n, t->E, w.as<int32_t>(),
                                                            t->d_status());
    if ((s = exclusive_scan_i32_to_i64(w.as<int32_t>(), t->E, p->W.as<int64_t>(), tmp, st)) !=
        RS_OK) {
      delete p;
      return s;
    }
    if (hipMemcpyAsync(&p->total, p->W.as<int64_t>() + t->E, sizeof(int64_t),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      delete p;
      set_error("rs_tc_plan_create: device error");
      return RS_EDEVICE;
    }
    if ((s = rs_tables_status(const_cast<rs_tables*>(t), stream)) != RS_OK) {
      delete p;
      set_error("rs_tc_plan_create: an edge endpoint is outside [0, num_sets)");
      return RS_EINVAL;
    }
  } else {
    (void)hipMemsetAsync(p->W.p, 0, sizeof(int64_t), st);
    (void)hipStreamSynchronize(st);
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
  if (p->total > 0) {
    const int32_t* keys = d_probe_keys ? d_probe_keys : t->d_elems;
    const int64_t ntiles = (p->total + kTcTile - 1) / kTcTile;
    const int grid = grid_for(ntiles * kWave, kTcBlock, 8);
    if (layout == RS_HIERARCHICAL)
      tc_tiled_kernel<RS_HIERARCHICAL, 4><<<grid, kTcBlock, 0, st>>>(
          t->rows.as<int32_t>(), t->d_elems, t->d_off, keys, t->lengths.as<int32_t>(),
          t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->W.as<int64_t>(), p->E,
          p->total, kTcTile, out, t->d_status());
    else
      tc_tiled_kernel<RS_NORMAL, 4><<<grid, kTcBlock, 0, st>>>(
          t->rows.as<int32_t>(), t->d_elems, t->d_off, keys, t->lengths.as<int32_t>(),
          t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, p->W.as<int64_t>(), p->E,
          p->total, kTcTile, out, t->d_status());
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
