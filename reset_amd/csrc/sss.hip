// sss.hip -- set-similarity verification (set-similarity-search/sss.cu:93-154)
// and the prefix-filter self-join pipeline around it.
//
// Verification uses the same flattened-probe tiling as triangle counting: W
// is the prefix of per-candidate probe counts min(|S_u|, |S_v|); waves take
// equal probe tiles, so one very large candidate is split across waves and
// thousands of small ones share a wave.  Hits are reduced per candidate
// inside each 64-probe step (a candidate's probes occupy consecutive lanes)
// and added to sizes[e] once per segment; a final pass applies the exact
// rule q*r >= p*(|S_u| + |S_v| - r) in 64-bit integers.
#include "sss_verify.hpp"

namespace rs {

constexpr int kSssBlock = 256;
constexpr int64_t kSssTile = 4096;

__global__ void sss_work_kernel(const int32_t* __restrict__ cu, const int32_t* __restrict__ cv,
                                const int32_t* __restrict__ off, int64_t n, int64_t P,
                                int32_t* __restrict__ w, int32_t* __restrict__ sizes,
                                int32_t* status) {
  bool bad = false;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = cu[e], v = cv[e];
    sizes[e] = 0;
    if (u < 0 || u >= n || v < 0 || v >= n) {
      bad = true;
      w[e] = 0;
      continue;
    }
    const int32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
    w[e] = du < dv ? du : dv;
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

template <int LAYOUT>
__global__ void __launch_bounds__(kSssBlock) sss_tiled_kernel(
    const int32_t* __restrict__ cu, const int32_t* __restrict__ cv,
    const int32_t* __restrict__ off, const int32_t* __restrict__ keys,
    const int32_t* __restrict__ Ls, const int64_t* __restrict__ toff,
    const int32_t* __restrict__ T, int64_t BN, int32_t b, uint32_t M,
    const int64_t* __restrict__ W, int64_t P, int64_t total, int64_t tile,
    int32_t* __restrict__ sizes, int32_t* __restrict__ status) {
  const int lane = __lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  const int64_t ntiles = (total + tile - 1) / tile;
  bool bad = false;
  for (int64_t ti = wave; ti < ntiles; ti += nwaves) {
    const int64_t ts = ti * tile;
    const int64_t te = ts + tile < total ? ts + tile : total;
    int64_t lo = 0, hi = P;
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
      if (ei < P) {
        const int64_t ws = W[ei];
        valid = ws < te;
        if (valid) {
          const int64_t we = W[ei + 1];
          const int64_t s = ws > ts ? ws : ts;
          const int64_t en = we < te ? we : te;
          cnt = (int32_t)(en - s);
          if (cnt > 0) {
            int32_t u = cu[ei], v = cv[ei];
            const int32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
            if (du > dv) {  // sss.cu:107-115: the smaller set probes
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
      for (int32_t p0 = 0; p0 < S; p0 += kWave) {
        const int32_t p = p0 + lane;
        const bool act = p < S;
        int owner = 0;
#pragma unroll
        for (int sft = 32; sft > 0; sft >>= 1) {
          const int32_t c = __shfl(incl, owner + sft - 1, kWave);
          if (c <= p) owner += sft;
        }
        owner = owner > kWave - 1 ? kWave - 1 : owner;
        const int32_t kb = __shfl(kbase, owner, kWave);
        const int32_t ex = __shfl(excl, owner, kWave);
        const int64_t tbo = __shfl(tb, owner, kWave);
        const uint32_t Lo = __shfl(Lv, owner, kWave);
        bool hit = false;
        if (act) {
          const int32_t key = keys[kb + (p - ex)];
          const int r = probe(T + tbo, BN, home_bucket<LAYOUT>((uint32_t)key, Lo, M), Lo, b, key);
          hit = r == 1;
          bad |= r < 0;
        }
        // segmented reduction: lanes of one candidate are contiguous
        const int prev_owner = __shfl_up(owner, 1, kWave);
        const bool head = act && (lane == 0 || prev_owner != owner);
        const uint64_t heads = __ballot(head);
        const uint64_t hits = __ballot(hit);
        if (head) {
          const uint64_t above = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
          const uint64_t upto = above ? ((1ull << __builtin_ctzll(above)) - 1ull) : ~0ull;
          const uint64_t seg = upto & (~0ull << lane);
          const int c = __popcll(hits & seg);
          if (c) atomicAdd(&sizes[e + owner], c);
        }
      }
      if (__popcll(__ballot(valid)) < kWave) break;
      e += kWave;
      if (e >= P) break;
      if (W[e] >= te) break;
    }
  }
  if (__any(bad) && lane == 0) atomicOr(status, kErrProbeBound);
}

__global__ void sss_flags_kernel(const int32_t* __restrict__ cu, const int32_t* __restrict__ cv,
                                 const int32_t* __restrict__ off, int64_t n, int64_t P, int32_t p,
                                 int32_t q, const int32_t* __restrict__ sizes,
                                 int32_t* __restrict__ flags) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = cu[e], v = cv[e];
    int32_t f = 0;
    if (u >= 0 && u < n && v >= 0 && v < n) {
      const int64_t r = sizes[e];
      const int64_t U = (int64_t)(off[u + 1] - off[u]) + (off[v + 1] - off[v]) - r;
      f = U > 0 && (int64_t)q * r >= (int64_t)p * U;
    }
    flags[e] = f;
  }
}

rs_status sss_verify_enqueue(const rs_tables* t, int32_t layout, const int32_t* keys,
                             const int32_t* cand_u, const int32_t* cand_v, int64_t P, int32_t p,
                             int32_t q, int32_t* flags, int32_t* sizes, DevBuf& W, DevBuf& tmp,
                             hipStream_t st) {
  const int32_t* T = nullptr;
  RS_TRY(require_layout(t, layout, &T));
  RS_REQUIRE(p >= 0 && q > 0, "sss: threshold p/q must have p >= 0, q > 0");
  if (P == 0) return RS_OK;
  RS_REQUIRE(cand_u && cand_v && sizes, "sss: null candidate or sizes array");
  if (W.bytes < sizeof(int64_t) * (size_t)(P + 1)) RS_TRY(W.alloc(sizeof(int64_t) * (size_t)(P + 1)));
  DevBuf w;
  RS_TRY(w.alloc(sizeof(int32_t) * (size_t)P));
  const int g = grid_for(P, 256, 16);
  sss_work_kernel<<<g, 256, 0, st>>>(cand_u, cand_v, t->d_off, t->n, P, w.as<int32_t>(), sizes,
                                     t->d_status());
  RS_TRY(exclusive_scan_i32_to_i64(w.as<int32_t>(), P, W.as<int64_t>(), tmp, st));
  int64_t total = 0;
  RS_HIP(hipMemcpyAsync(&total, W.as<int64_t>() + P, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));  // also keeps `w` alive until the scan consumed it
  if (total > 0) {
    const int64_t ntiles = (total + kSssTile - 1) / kSssTile;
    const int grid = grid_for(ntiles * kWave, kSssBlock, 8);
    if (layout == RS_HIERARCHICAL)
      sss_tiled_kernel<RS_HIERARCHICAL><<<grid, kSssBlock, 0, st>>>(
          cand_u, cand_v, t->d_off, keys, t->lengths.as<int32_t>(), t->toff.as<int64_t>(), T,
          t->BN, t->b, (uint32_t)t->M, W.as<int64_t>(), P, total, kSssTile, sizes, t->d_status());
    else
      sss_tiled_kernel<RS_NORMAL><<<grid, kSssBlock, 0, st>>>(
          cand_u, cand_v, t->d_off, keys, t->lengths.as<int32_t>(), t->toff.as<int64_t>(), T,
          t->BN, t->b, (uint32_t)t->M, W.as<int64_t>(), P, total, kSssTile, sizes, t->d_status());
  }
  if (flags)
    sss_flags_kernel<<<g, 256, 0, st>>>(cand_u, cand_v, t->d_off, t->n, P, p, q, sizes, flags);
  RS_HIP(hipGetLastError());
  return RS_OK;
}

}  // namespace rs

using namespace rs;

extern "C" {

rs_status rs_sss_verify_async(const rs_tables* t, int32_t layout, const int32_t* d_probe_keys,
                              const int32_t* d_cand_u, const int32_t* d_cand_v,
                              int64_t num_candidates, int32_t p, int32_t q, int32_t* d_flags,
                              int32_t* d_sizes, rs_stream stream) {
  RS_REQUIRE(t, "rs_sss_verify_async: null tables");
  RS_REQUIRE(num_candidates >= 0, "rs_sss_verify_async: negative candidate count");
  hipStream_t st = as_stream(stream);
  DevBuf W, tmp;
  RS_TRY(sss_verify_enqueue(t, layout, d_probe_keys ? d_probe_keys : t->d_elems, d_cand_u,
                            d_cand_v, num_candidates, p, q, d_flags, d_sizes, W, tmp, st));
  // W and tmp are freed on return: wait for the kernels that read them.
  RS_HIP(hipStreamSynchronize(st));
  return RS_OK;
}

}  // extern "C"
