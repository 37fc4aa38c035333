// sss_join.hip -- prefix-filter set-similarity self-join (Jaccard >= p/q)
// feeding the batched hash verification of sss.hip.
//
// Global token order: ascending frequency, ties by token id.  With every set
// written in that order, two sets with J >= t share a token within their
// prefixes of length s - ceil(t*s) + 1 (Chaudhuri et al., ICDE 2006), and
// J >= t also needs p*max(|A|,|B|) <= q*min(|A|,|B|) (size filter).
// Stages (all on the device):
//   1. token frequencies (radix sort of the tokens, run boundaries)
//   2. rank[token] by (frequency, id) (radix sort of 64-bit keys)
//   3. each set's tokens as ranks, sorted per set (segmented radix sort)
//   4. prefix entries (rank, size, set) sorted; per entry, its partners are
//      the later entries of the same rank whose size passes the size filter
//      (a contiguous window, found by binary search)
//   5. candidate keys (i << 32 | j), radix sort + unique
//   6. hash verification of every candidate (exact integer rule), compaction
// Limits: < 2^21 sets, sets < 2^21 tokens, token ids < 2^22.
#include <algorithm>
#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "sss_verify.hpp"

struct rs_sss_index {
  const int32_t* off = nullptr;  // borrowed collection CSR
  const int32_t* el = nullptr;
  int64_t n = 0, E = 0, V = 0, M = 0;
  int32_t p = 0, q = 0;
  rs::DevBuf rank;  // int32[V]: global rank of each token id
  rs::DevBuf ent;   // uint64[M]: sorted prefix entries rank << 42 | size << 21 | set
  // host copies for admission: the rank of each token id, and the number of
  // prefix entries of each rank
  std::vector<int32_t> h_rank, h_rank_entries;
};

struct rs_sss_result {
  rs::DevBuf pairs;  // int64 keys i << 32 | j, ascending
  int64_t num_pairs = 0;
  int64_t num_candidates = 0;
  int64_t verify_probes = 0;    // sum over candidates of min(|S_i|, |S_j|)
  int64_t verify_elements = 0;  // sum over candidates of |S_i| + |S_j|
  float ms[3] = {0, 0, 0};
};

namespace rs {

__global__ void max_token_kernel(const int32_t* __restrict__ x, int64_t n, int32_t* out,
                                 int32_t* err) {
  int32_t m = -1;
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = x[i];
    bad |= v < 0;
    m = v > m ? v : m;
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const int32_t o = __shfl_xor(m, d, kWave);
    m = o > m ? o : m;
  }
  if (__lane_id() == 0) atomicMax(out, m);
  if (__any(bad) && __lane_id() == 0) atomicOr(err, kErrBadInput);
}

// freq[t] from the sorted token array: run [s, e) of value t gives e - s.
__global__ void run_freq_kernel(const int32_t* __restrict__ sorted, int64_t n,
                                int32_t* __restrict__ freq) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = sorted[i];
    if (i + 1 == n || sorted[i + 1] != v) {
      // i is the end of v's run; find its start by the previous run end
      int64_t lo = 0, hi = i;  // first index with value v in [0, i]
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (sorted[mid] < v) lo = mid + 1; else hi = mid;
      }
      freq[v] = (int32_t)(i + 1 - lo);
    }
  }
}

__global__ void order_keys_kernel(const int32_t* __restrict__ freq, int64_t V,
                                  uint64_t* __restrict__ keys) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < V;
       t += (int64_t)gridDim.x * blockDim.x)
    keys[t] = ((uint64_t)(uint32_t)freq[t] << 32) | (uint64_t)t;
}

__global__ void rank_kernel(const uint64_t* __restrict__ sorted_keys, int64_t V,
                            int32_t* __restrict__ rank) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
       i += (int64_t)gridDim.x * blockDim.x)
    rank[(int64_t)(sorted_keys[i] & 0xFFFFFFFFull)] = (int32_t)i;
}

__global__ void map_rank_kernel(const int32_t* __restrict__ tok, int64_t n,
                                const int32_t* __restrict__ rank, int32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = rank[tok[i]];
}

__host__ __device__ __forceinline__ int32_t prefix_len(int32_t s, int32_t p, int32_t q) {
  int64_t pl = (int64_t)s - ((int64_t)p * s + q - 1) / q + 1;
  if (pl > s) pl = s;
  return pl < 0 ? 0 : (int32_t)pl;
}

__global__ void prefix_count_kernel(const int32_t* __restrict__ off, int64_t n, int32_t p,
                                    int32_t q, int32_t* __restrict__ pl, int32_t* err) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = off[i + 1] - off[i];
    bad |= s < 0 || s >= (1 << 21);
    pl[i] = bad ? 0 : prefix_len(s, p, q);
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(err, kErrBadInput);
}

constexpr int kRankShift = 42, kSizeShift = 21;
constexpr uint64_t kField21 = (1ull << 21) - 1;

__global__ void prefix_emit_kernel(const int32_t* __restrict__ off, int64_t n,
                                   const int32_t* __restrict__ ranked,
                                   const int64_t* __restrict__ poff, int32_t p, int32_t q,
                                   uint64_t* __restrict__ entries) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = off[i + 1] - off[i];
    const int32_t pl = prefix_len(s, p, q);
    for (int32_t k = 0; k < pl; ++k)
      entries[poff[i] + k] = ((uint64_t)(uint32_t)ranked[off[i] + k] << kRankShift) |
                             ((uint64_t)s << kSizeShift) | (uint64_t)i;
  }
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* a, int64_t lo, int64_t hi,
                                                   uint64_t key) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Partners of entry x: later entries y of the same rank with
// p * size_y <= q * size_x (entries are sorted by rank, then size).
__global__ void cand_count_kernel(const uint64_t* __restrict__ ent, int64_t m, int32_t p,
                                  int32_t q, int64_t* __restrict__ cnt) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < m;
       x += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = ent[x];
    const uint64_t rank = k >> kRankShift;
    const uint64_t sz = (k >> kSizeShift) & kField21;
    uint64_t smax = ((uint64_t)q * sz) / (uint64_t)p;
    if (smax > kField21) smax = kField21;
    const uint64_t bound = (rank << kRankShift) | (smax << kSizeShift) | kField21;
    const int64_t end = lower_bound_u64(ent, x + 1, m, bound + 1);
    cnt[x] = end - x - 1;
  }
}

__global__ void cand_emit_kernel(const uint64_t* __restrict__ ent, int64_t m,
                                 const int64_t* __restrict__ coff, uint64_t* __restrict__ cand) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < m;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = coff[x + 1] - coff[x];
    const uint64_t ix = ent[x] & kField21;
    for (int64_t j = 0; j < c; ++j) {
      const uint64_t iy = ent[x + 1 + j] & kField21;
      cand[coff[x] + j] = ix < iy ? ((ix << 32) | iy) : ((iy << 32) | ix);
    }
  }
}

__global__ void split_pairs_kernel(const uint64_t* __restrict__ cand, int64_t n,
                                   int32_t* __restrict__ u, int32_t* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    u[i] = (int32_t)(cand[i] >> 32);
    v[i] = (int32_t)(cand[i] & 0xFFFFFFFFull);
  }
}

// Exact rule on the pair sizes r: flag iff q*r >= p*(|S_u|+|S_v|-r) > ... ;
// also accumulates sum(min) and sum(|S_u|+|S_v|) for the work accounting.
__global__ void jaccard_flags_kernel(const int32_t* __restrict__ a_off,
                                     const int32_t* __restrict__ b_off, const int32_t* __restrict__ u,
                                     const int32_t* __restrict__ v, const int32_t* __restrict__ r,
                                     int64_t n, int32_t p, int32_t q, int32_t* __restrict__ flags,
                                     unsigned long long* work) {
  unsigned long long smin = 0, ssum = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t su = a_off[u[i] + 1] - a_off[u[i]], sv = b_off[v[i] + 1] - b_off[v[i]];
    const int64_t U = su + sv - r[i];
    flags[i] = U > 0 && (int64_t)q * r[i] >= (int64_t)p * U;
    smin += (unsigned long long)(su < sv ? su : sv);
    ssum += (unsigned long long)(su + sv);
  }
  smin = wave_sum(smin);
  ssum = wave_sum(ssum);
  if (__lane_id() == 0) {
    atomicAdd(&work[0], smin);
    atomicAdd(&work[1], ssum);
  }
}

__global__ void count_to_i64_kernel(const int64_t* __restrict__ in, int64_t n,
                                    int64_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

#define RS_ROC(expr)                                                        \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess) {                                                 \
      ::rs::set_error("%s failed: %s", #expr, hipGetErrorString(_e));       \
      return RS_EDEVICE;                                                    \
    }                                                                       \
  } while (0)

template <typename T>
static rs_status sort_keys(const T* in, T* out, int64_t n, unsigned end_bit, DevBuf& tmp,
                           hipStream_t st) {
  size_t bytes = 0;
  RS_ROC(rocprim::radix_sort_keys(nullptr, bytes, in, out, (size_t)n, 0, end_bit, st));
  if (tmp.bytes < bytes) RS_TRY(tmp.alloc(bytes));
  RS_ROC(rocprim::radix_sort_keys(tmp.p, bytes, in, out, (size_t)n, 0, end_bit, st));
  return RS_OK;
}

static unsigned bits_for(uint64_t max_value) {
  unsigned b = 1;
  while (b < 64 && (max_value >> b) != 0) ++b;
  return b;
}

// Stages 1-2: rank[t] for every token t < V (V = max token + 1) by
// (frequency, id).  err: device error word (kErrBadInput on a negative id).
static rs_status token_ranks(const int32_t* d_el, int64_t E, DevBuf& tmp, hipStream_t st,
                             int32_t* d_scratch, DevBuf& rank, int64_t* V_out) {
  int32_t* d_max = d_scratch;  // [0] max token, [1] error bits
  RS_HIP(hipMemsetAsync(d_max, 0xFF, sizeof(int32_t), st));
  if (E > 0) max_token_kernel<<<grid_for(E, 256, 8), 256, 0, st>>>(d_el, E, d_max, d_max + 1);
  int32_t hm[2];
  RS_HIP(hipMemcpyAsync(hm, d_max, sizeof(hm), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  RS_REQUIRE(!hm[1], "sss: negative token id");
  const int64_t V = (int64_t)hm[0] + 1;
  RS_REQUIRE(V < (1ll << 22), "sss: token ids must be < 2^22");
  DevBuf sorted_tok, freq, okeys, okeys_sorted;
  RS_TRY(sorted_tok.alloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)));
  RS_TRY(freq.alloc(sizeof(int32_t) * (size_t)(V > 0 ? V : 1)));
  RS_HIP(hipMemsetAsync(freq.p, 0, sizeof(int32_t) * (size_t)(V > 0 ? V : 1), st));
  if (E > 0) {
    RS_TRY(sort_keys<int32_t>(d_el, sorted_tok.as<int32_t>(), E, bits_for((uint64_t)V), tmp, st));
    run_freq_kernel<<<grid_for(E, 256, 8), 256, 0, st>>>(sorted_tok.as<int32_t>(), E, freq.as<int32_t>());
  }
  sorted_tok.release();
  RS_TRY(okeys.alloc(sizeof(uint64_t) * (size_t)(V > 0 ? V : 1)));
  RS_TRY(okeys_sorted.alloc(sizeof(uint64_t) * (size_t)(V > 0 ? V : 1)));
  RS_TRY(rank.alloc(sizeof(int32_t) * (size_t)(V > 0 ? V : 1)));
  if (V > 0) {
    order_keys_kernel<<<grid_for(V, 256, 8), 256, 0, st>>>(freq.as<int32_t>(), V, okeys.as<uint64_t>());
    RS_TRY(sort_keys<uint64_t>(okeys.as<uint64_t>(), okeys_sorted.as<uint64_t>(), V,
                               32 + bits_for((uint64_t)(E > 0 ? E : 1)), tmp, st));
    rank_kernel<<<grid_for(V, 256, 8), 256, 0, st>>>(okeys_sorted.as<uint64_t>(), V, rank.as<int32_t>());
  }
  *V_out = V;
  return RS_OK;
}

// Stage 3: each set's tokens as ranks, ascending within the set.
static rs_status ranked_sets(const int32_t* d_off, int64_t n, const int32_t* d_el, int64_t E,
                             const DevBuf& rank, int64_t V, DevBuf& tmp, hipStream_t st,
                             DevBuf& ranked_sorted) {
  DevBuf ranked;
  RS_TRY(ranked.alloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)));
  RS_TRY(ranked_sorted.alloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)));
  if (E > 0) {
    map_rank_kernel<<<grid_for(E, 256, 8), 256, 0, st>>>(d_el, E, rank.as<int32_t>(), ranked.as<int32_t>());
    size_t bytes = 0;
    RS_ROC(rocprim::segmented_radix_sort_keys(nullptr, bytes, ranked.as<int32_t>(),
                                              ranked_sorted.as<int32_t>(), (size_t)E, (unsigned)n,
                                              d_off, d_off + 1, 0, bits_for((uint64_t)V), st));
    if (tmp.bytes < bytes) RS_TRY(tmp.alloc(bytes));
    RS_ROC(rocprim::segmented_radix_sort_keys(tmp.p, bytes, ranked.as<int32_t>(),
                                              ranked_sorted.as<int32_t>(), (size_t)E, (unsigned)n,
                                              d_off, d_off + 1, 0, bits_for((uint64_t)V), st));
  }
  return RS_OK;
}

// Stage 4: prefix entries (rank << 42 | size << 21 | set), sorted.
static rs_status prefix_entries(const int32_t* d_off, int64_t n, const DevBuf& ranked_sorted,
                                int32_t p, int32_t q, DevBuf& tmp, hipStream_t st,
                                int32_t* d_scratch, DevBuf& ent_sorted, int64_t* M_out) {
  DevBuf pl, poff, ent;
  RS_TRY(pl.alloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1)));
  RS_TRY(poff.alloc(sizeof(int64_t) * (size_t)(n + 1)));
  if (n > 0)
    prefix_count_kernel<<<grid_for(n, 256, 8), 256, 0, st>>>(d_off, n, p, q, pl.as<int32_t>(),
                                                             d_scratch + 1);
  RS_TRY(exclusive_scan_i32_to_i64(pl.as<int32_t>(), n, poff.as<int64_t>(), tmp, st));
  int64_t M = 0;
  int32_t hm[2];
  RS_HIP(hipMemcpyAsync(&M, poff.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(hm, d_scratch, sizeof(hm), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  RS_REQUIRE(!hm[1], "sss: a set has 2^21 or more tokens");
  RS_TRY(ent.alloc(sizeof(uint64_t) * (size_t)(M > 0 ? M : 1)));
  RS_TRY(ent_sorted.alloc(sizeof(uint64_t) * (size_t)(M > 0 ? M : 1)));
  if (n > 0)
    prefix_emit_kernel<<<grid_for(n, 256, 8), 256, 0, st>>>(d_off, n, ranked_sorted.as<int32_t>(),
                                                            poff.as<int64_t>(), p, q, ent.as<uint64_t>());
  if (M > 0) RS_TRY(sort_keys<uint64_t>(ent.as<uint64_t>(), ent_sorted.as<uint64_t>(), M, 64, tmp, st));
  *M_out = M;
  return RS_OK;
}

static rs_status unique_sorted(DevBuf& keys, int64_t C, DevBuf& tmp, hipStream_t st, DevBuf& uniq,
                               int64_t* U_out);
static rs_status verify_candidates(const int32_t* a_off, const int32_t* a_el, const int32_t* b_off,
                                   const int32_t* b_el, DevBuf& uniq, int64_t U, int32_t p, int32_t q,
                                   float alpha, int32_t bsz, DevBuf& tmp, hipStream_t st,
                                   rs_sss_result* R);

// One wave per set: flags a set whose tokens are not strictly ascending
// (a repeated token would count twice in |S| and in the intersection, and
// the chunked pair kernel needs ascending sets) or that holds a negative id.
__global__ void ascending_check_kernel(const int32_t* __restrict__ off, int64_t n,
                                       const int32_t* __restrict__ el, int32_t* err) {
  const int lane = __lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  bool bad = false;
  for (int64_t i = wave; i < n; i += nwaves) {
    const int32_t s = off[i], e = off[i + 1];
    bad |= e < s;
    for (int32_t j = s + lane; j < e; j += kWave) {
      const int32_t x = el[j];
      bad |= x < 0 || (j + 1 < e && el[j + 1] <= x);
    }
  }
  if (__any(bad) && lane == 0) atomicOr(err, kErrBadInput);
}

// RS_EINVAL unless every set of the CSR is strictly ascending and
// non-negative; synchronizes st.
static rs_status check_ascending(const int32_t* d_off, int64_t n, const int32_t* d_el, const char* who,
                                 hipStream_t st) {
  if (n == 0) return RS_OK;
  DevBuf err;
  RS_TRY(err.alloc(16));
  RS_HIP(hipMemsetAsync(err.p, 0, 16, st));
  ascending_check_kernel<<<grid_for(n * kWave, 256, 8), 256, 0, st>>>(d_off, n, d_el, err.as<int32_t>());
  int32_t h = 0;
  RS_HIP(hipMemcpyAsync(&h, err.p, sizeof(h), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  RS_REQUIRE(h == 0, "%s: every set must hold strictly ascending, non-negative tokens", who);
  return RS_OK;
}

static rs_status selfjoin(const int32_t* d_off, int64_t n, const int32_t* d_el, int32_t p,
                          int32_t q, float alpha, int32_t bsz, hipStream_t st, rs_sss_result* R) {
  RS_REQUIRE(n >= 0 && n < (1ll << 21), "sss self-join: %lld sets exceed 2^21", (long long)n);
  RS_REQUIRE(p > 0 && q >= p, "sss self-join: threshold p/q must be in (0, 1]");
  RS_REQUIRE(alpha > 0.f && alpha <= 1.f && (bsz == 1 || bsz == 2 || bsz == 4),
             "sss self-join: the verification tables need load_factor in (0, 1] and bucket_size "
             "1, 2 or 4 (got %g, %d)", (double)alpha, bsz);
  RS_REQUIRE(group_max_build(alpha, bsz) >= 0,
             "sss self-join: no set fits the LDS table at load_factor %g", (double)alpha);
  RS_TRY(check_ascending(d_off, n, d_el, "sss self-join", st));
  hipEvent_t ev[4];
  for (auto& e : ev) RS_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  RS_HIP(hipEventRecord(ev[0], st));
  int32_t E32 = 0;
  if (n > 0) {
    RS_HIP(hipMemcpyAsync(&E32, d_off + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
  }
  const int64_t E = E32;
  DevBuf tmp, scratch;
  RS_TRY(scratch.alloc(16));
  int32_t* d_max = scratch.as<int32_t>();
  RS_HIP(hipMemsetAsync(scratch.p, 0, 16, st));
  // 1-3. global order and each set in rank order
  DevBuf rank, ranked_sorted;
  int64_t V = 0;
  RS_TRY(token_ranks(d_el, E, tmp, st, d_max, rank, &V));
  RS_TRY(ranked_sets(d_off, n, d_el, E, rank, V, tmp, st, ranked_sorted));
  rank.release();
  RS_HIP(hipEventRecord(ev[1], st));
  // 4. prefix entries
  DevBuf ent_sorted;
  int64_t M = 0;
  RS_TRY(prefix_entries(d_off, n, ranked_sorted, p, q, tmp, st, d_max, ent_sorted, &M));
  ranked_sorted.release();
  // 5. candidates
  DevBuf cnt, coff, cand, uniq;
  RS_TRY(cnt.alloc(sizeof(int64_t) * (size_t)(M > 0 ? M : 1)));
  RS_TRY(coff.alloc(sizeof(int64_t) * (size_t)(M + 1)));
  if (M > 0)
    cand_count_kernel<<<grid_for(M, 256, 8), 256, 0, st>>>(ent_sorted.as<uint64_t>(), M, p, q,
                                                           cnt.as<int64_t>());
  RS_TRY(exclusive_scan_i64(cnt.as<int64_t>(), M, coff.as<int64_t>(), tmp, st));
  int64_t C = 0;
  RS_HIP(hipMemcpyAsync(&C, coff.as<int64_t>() + M, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (C > (1ll << 32)) {
    set_error("sss self-join: %lld candidate pairs (before dedup) exceed the 2^32 working limit",
              (long long)C);
    return RS_ENOMEM;
  }
  cnt.release();
  int64_t U = 0;
  if (C > 0) {
    RS_TRY(cand.alloc(sizeof(uint64_t) * (size_t)C));
    cand_emit_kernel<<<grid_for(M, 256, 8), 256, 0, st>>>(ent_sorted.as<uint64_t>(), M,
                                                          coff.as<int64_t>(), cand.as<uint64_t>());
    ent_sorted.release();
    coff.release();
    RS_TRY(unique_sorted(cand, C, tmp, st, uniq, &U));
  }
  R->num_candidates = U;
  RS_HIP(hipEventRecord(ev[2], st));
  // 6. verification: every candidate's |S_i cap S_j| by the batched LDS
  //    hash kernel (indexed pairs), then the exact rule + compaction
  RS_TRY(verify_candidates(d_off, d_el, d_off, d_el, uniq, U, p, q, alpha, bsz, tmp, st, R));
  RS_HIP(hipEventRecord(ev[3], st));
  RS_HIP(hipEventSynchronize(ev[3]));
  for (int i = 0; i < 3; ++i) RS_HIP(hipEventElapsedTime(&R->ms[i], ev[i], ev[i + 1]));
  return RS_OK;
}

// ---------------------------------------------------------------- lookups
// Query token keys: 1 + global rank for a token of the collection's id range,
// 0 for any other token (frequency 0 there: it orders before every known
// token and has no prefix entries).  A wave per query also checks the query
// is strictly ascending, non-negative and shorter than 2^21 tokens.
__global__ void query_keys_kernel(const int32_t* __restrict__ qoff, int64_t Q,
                                  const int32_t* __restrict__ qtok, const int32_t* __restrict__ rank,
                                  int64_t V, uint32_t* __restrict__ keys, int32_t* err) {
  const int lane = __lane_id();
  const int64_t waves = (int64_t)gridDim.x * blockDim.x / kWave;
  bool bad = false;
  for (int64_t qi = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; qi < Q; qi += waves) {
    const int32_t a = qoff[qi], z = qoff[qi + 1];
    bad |= z < a || z - a >= (1 << 21);
    for (int32_t i = a + lane; i < z; i += kWave) {
      const int32_t t = qtok[i];
      bad |= t < 0 || (i > a && qtok[i - 1] >= t);
      keys[i] = (t >= 0 && t < V) ? (uint32_t)rank[t] + 1u : 0u;
    }
  }
  if (__any(bad) && lane == 0) atomicOr(err, kErrBadInput);
}

// Per query: prefix length and the size window [ceil(p s / q), floor(q s / p)].
__device__ __forceinline__ void lookup_window(int32_t s, int32_t p, int32_t q, uint64_t& smin,
                                              uint64_t& smax) {
  smin = ((uint64_t)p * (uint64_t)s + (uint64_t)q - 1) / (uint64_t)q;
  smax = ((uint64_t)q * (uint64_t)s) / (uint64_t)p;
  if (smin < 1) smin = 1;
  if (smax > kField21) smax = kField21;
}

__global__ void lookup_prefix_len_kernel(const int32_t* __restrict__ qoff, int64_t Q, int32_t p,
                                         int32_t q, int32_t* __restrict__ pl) {
  for (int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; qi < Q;
       qi += (int64_t)gridDim.x * blockDim.x)
    pl[qi] = prefix_len(qoff[qi + 1] - qoff[qi], p, q);
}

// For prefix position k of query qi: the run of prefix entries with the
// token's rank and a size inside the window -> [lo, lo + cnt).
__global__ void lookup_count_kernel(const int32_t* __restrict__ qoff, int64_t Q,
                                    const uint32_t* __restrict__ skeys, const int64_t* __restrict__ poff,
                                    const uint64_t* __restrict__ ent, int64_t M, int32_t p, int32_t q,
                                    int64_t* __restrict__ lo_out, int64_t* __restrict__ cnt_out) {
  const int lane = __lane_id();
  const int64_t waves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t qi = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; qi < Q; qi += waves) {
    const int32_t a = qoff[qi], s = qoff[qi + 1] - a;
    const int64_t p0 = poff[qi], np = poff[qi + 1] - p0;
    uint64_t smin, smax;
    lookup_window(s, p, q, smin, smax);
    for (int64_t k = lane; k < np; k += kWave) {
      const uint32_t key = skeys[a + k];
      int64_t lo = 0, cnt = 0;
      if (key != 0u && smin <= smax) {
        const uint64_t r = key - 1u;
        lo = lower_bound_u64(ent, 0, M, (r << kRankShift) | (smin << kSizeShift));
        const int64_t hi = lower_bound_u64(ent, lo, M, ((r << kRankShift) | (smax << kSizeShift) | kField21) + 1);
        cnt = hi - lo;
      }
      lo_out[p0 + k] = lo;
      cnt_out[p0 + k] = cnt;
    }
  }
}

__global__ void lookup_emit_kernel(const int64_t* __restrict__ poff, int64_t Q,
                                   const int64_t* __restrict__ lo, const int64_t* __restrict__ coff,
                                   const uint64_t* __restrict__ ent, uint64_t* __restrict__ cand) {
  const int lane = __lane_id();
  const int64_t waves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t qi = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; qi < Q; qi += waves) {
    for (int64_t k = poff[qi] + lane; k < poff[qi + 1]; k += kWave) {
      const int64_t c0 = coff[k], c = coff[k + 1] - c0;
      for (int64_t j = 0; j < c; ++j)
        cand[c0 + j] = ((uint64_t)qi << 32) | (ent[lo[k] + j] & kField21);
    }
  }
}

static rs_status unique_sorted(DevBuf& keys, int64_t C, DevBuf& tmp, hipStream_t st, DevBuf& uniq,
                               int64_t* U_out) {
  DevBuf sorted, nuniq;
  RS_TRY(sorted.alloc(sizeof(uint64_t) * (size_t)C));
  RS_TRY(sort_keys<uint64_t>(keys.as<uint64_t>(), sorted.as<uint64_t>(), C, 64, tmp, st));
  keys.release();
  RS_TRY(uniq.alloc(sizeof(uint64_t) * (size_t)C));
  RS_TRY(nuniq.alloc(16));
  size_t bytes = 0;
  RS_ROC(rocprim::unique(nullptr, bytes, sorted.as<uint64_t>(), uniq.as<uint64_t>(),
                         nuniq.as<size_t>(), (size_t)C, rocprim::equal_to<uint64_t>(), st));
  if (tmp.bytes < bytes) RS_TRY(tmp.alloc(bytes));
  RS_ROC(rocprim::unique(tmp.p, bytes, sorted.as<uint64_t>(), uniq.as<uint64_t>(),
                         nuniq.as<size_t>(), (size_t)C, rocprim::equal_to<uint64_t>(), st));
  size_t hu = 0;
  RS_HIP(hipMemcpyAsync(&hu, nuniq.p, sizeof(size_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  *U_out = (int64_t)hu;
  return RS_OK;
}

__global__ void gather_i32_kernel(const int32_t* __restrict__ src, const int64_t* __restrict__ idx,
                                  int64_t n, int32_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

__global__ void scatter_i32_kernel(const int32_t* __restrict__ src, const int64_t* __restrict__ idx,
                                   int64_t n, int32_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[idx[i]] = src[i];
}

__global__ void invert_flags_kernel(const int32_t* __restrict__ in, int64_t n, int32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = !in[i];
}

// Indices i < n with flags[i] != 0, ascending.
static rs_status select_indices(const int32_t* flags, int64_t n, DevBuf& out, int64_t* count,
                                DevBuf& tmp, hipStream_t st) {
  DevBuf nsel;
  RS_TRY(out.alloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1)));
  RS_TRY(nsel.alloc(16));
  rocprim::counting_iterator<int64_t> idx(0);
  size_t bytes = 0;
  RS_ROC(rocprim::select(nullptr, bytes, idx, flags, out.as<int64_t>(), nsel.as<size_t>(), (size_t)n, st));
  if (tmp.bytes < bytes) RS_TRY(tmp.alloc(bytes));
  RS_ROC(rocprim::select(tmp.p, bytes, idx, flags, out.as<int64_t>(), nsel.as<size_t>(), (size_t)n, st));
  size_t h = 0;
  RS_HIP(hipMemcpyAsync(&h, nsel.p, sizeof(size_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  *count = (int64_t)h;
  return RS_OK;
}

// sizes[i] = |A_{cu[i]} cap B_{cv[i]}| for candidates sorted by cu: the ones
// whose build set fits the grouped path share one LDS table per run of equal
// cu (sss_group.hip); the rest go through the batched pair kernel.
static rs_status sizes_by_groups(const int32_t* a_off, const int32_t* a_el, const int32_t* b_off,
                                 const int32_t* b_el, const int32_t* cu, const int32_t* cv, int64_t U,
                                 float alpha, int32_t bsz, int32_t* sizes, DevBuf& tmp,
                                 hipStream_t st) {
  DevBuf sel, err, idx_g;
  RS_TRY(sel.alloc(sizeof(int32_t) * (size_t)U));
  RS_TRY(err.alloc(16));
  RS_HIP(hipMemsetAsync(err.p, 0, 16, st));
  RS_TRY(group_partition(a_off, cu, U, group_max_build(alpha, bsz), sel.as<int32_t>(), st));
  int64_t G = 0;
  RS_TRY(select_indices(sel.as<int32_t>(), U, idx_g, &G, tmp, st));
  if (G == U) {
    RS_TRY(group_verify(alpha, bsz, a_off, a_el, b_off, b_el, cu, cv, U, sizes, err.as<int32_t>(), tmp,
                        st));
  } else {
    if (G > 0) {
      DevBuf gu, gv, gs;
      RS_TRY(gu.alloc(sizeof(int32_t) * (size_t)G));
      RS_TRY(gv.alloc(sizeof(int32_t) * (size_t)G));
      RS_TRY(gs.alloc(sizeof(int32_t) * (size_t)G));
      const int g = grid_for(G, 256, 8);
      gather_i32_kernel<<<g, 256, 0, st>>>(cu, idx_g.as<int64_t>(), G, gu.as<int32_t>());
      gather_i32_kernel<<<g, 256, 0, st>>>(cv, idx_g.as<int64_t>(), G, gv.as<int32_t>());
      RS_TRY(group_verify(alpha, bsz, a_off, a_el, b_off, b_el, gu.as<int32_t>(), gv.as<int32_t>(), G,
                          gs.as<int32_t>(), err.as<int32_t>(), tmp, st));
      scatter_i32_kernel<<<g, 256, 0, st>>>(gs.as<int32_t>(), idx_g.as<int64_t>(), G, sizes);
    }
    const int64_t R = U - G;
    DevBuf inv, idx_r, ru, rv, rs;
    RS_TRY(inv.alloc(sizeof(int32_t) * (size_t)U));
    invert_flags_kernel<<<grid_for(U, 256, 8), 256, 0, st>>>(sel.as<int32_t>(), U, inv.as<int32_t>());
    int64_t R2 = 0;
    RS_TRY(select_indices(inv.as<int32_t>(), U, idx_r, &R2, tmp, st));
    RS_REQUIRE(R2 == R, "sss: candidate split is inconsistent");
    RS_TRY(ru.alloc(sizeof(int32_t) * (size_t)R));
    RS_TRY(rv.alloc(sizeof(int32_t) * (size_t)R));
    RS_TRY(rs.alloc(sizeof(int32_t) * (size_t)R));
    const int g = grid_for(R, 256, 8);
    gather_i32_kernel<<<g, 256, 0, st>>>(cu, idx_r.as<int64_t>(), R, ru.as<int32_t>());
    gather_i32_kernel<<<g, 256, 0, st>>>(cv, idx_r.as<int64_t>(), R, rv.as<int32_t>());
    // the batched pair kernel's bins at the same (alpha, b): the largest
    // chunk whose table fits a wave's LDS region
    rs_pairs_opts o;
    rs_pairs_default_opts(&o);
    int32_t chunk = 0;
    for (int32_t c = 64; c <= 1024; c += 64) {
      const int32_t L = calc_length(c, alpha, bsz);
      if (L > 0 && (int64_t)L * bsz <= o.lds_slots_per_wave) chunk = c;
    }
    RS_REQUIRE(chunk > 0, "sss: no 64-element chunk fits a pair table at load_factor %g", (double)alpha);
    if (alpha != kVerifyAlpha || bsz != kVerifyB)
      for (int i = 0; i < o.num_bins; ++i) o.bins[i] = rs_pairs_bin{alpha, bsz, chunk, o.bins[i].unroll, 64};
    rs_pairs_plan* plan = nullptr;
    RS_TRY(rs_pairs_plan_create(R, &o, &plan));
    struct PGuard {
      rs_pairs_plan* p;
      ~PGuard() { rs_pairs_plan_free(p); }
    } pg{plan};
    RS_TRY(rs_pairs_intersect_cross_async(plan, a_off, a_el, b_off, b_el, ru.as<int32_t>(),
                                          rv.as<int32_t>(), R, rs.as<int32_t>(), st));
    scatter_i32_kernel<<<g, 256, 0, st>>>(rs.as<int32_t>(), idx_r.as<int64_t>(), R, sizes);
    RS_TRY(rs_pairs_plan_status(plan, st));
  }
  int32_t herr = 0;
  RS_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (herr) {
    set_error("sss: a grouped verification table overflowed");
    return RS_EOVERFLOW;
  }
  return RS_OK;
}

// Verification of unique candidate keys (a << 32 | b) over families A, B:
// sizes by the batched LDS kernel, the exact rule, compaction into R.
static rs_status verify_candidates(const int32_t* a_off, const int32_t* a_el, const int32_t* b_off,
                                   const int32_t* b_el, DevBuf& uniq, int64_t U, int32_t p, int32_t q,
                                   float alpha, int32_t bsz, DevBuf& tmp, hipStream_t st,
                                   rs_sss_result* R) {
  if (U == 0) {
    RS_TRY(R->pairs.alloc(16));
    R->num_pairs = 0;
    return RS_OK;
  }
  DevBuf cu, cv, flags, sizes, out_n, work;
  RS_TRY(cu.alloc(sizeof(int32_t) * (size_t)U));
  RS_TRY(cv.alloc(sizeof(int32_t) * (size_t)U));
  RS_TRY(flags.alloc(sizeof(int32_t) * (size_t)U));
  RS_TRY(sizes.alloc(sizeof(int32_t) * (size_t)U));
  RS_TRY(work.alloc(16));
  RS_HIP(hipMemsetAsync(work.p, 0, 16, st));
  split_pairs_kernel<<<grid_for(U, 256, 8), 256, 0, st>>>(uniq.as<uint64_t>(), U, cu.as<int32_t>(),
                                                         cv.as<int32_t>());
  RS_TRY(sizes_by_groups(a_off, a_el, b_off, b_el, cu.as<int32_t>(), cv.as<int32_t>(), U, alpha, bsz,
                         sizes.as<int32_t>(), tmp, st));
  jaccard_flags_kernel<<<grid_for(U, 256, 8), 256, 0, st>>>(a_off, b_off, cu.as<int32_t>(), cv.as<int32_t>(),
                                                           sizes.as<int32_t>(), U, p, q,
                                                           flags.as<int32_t>(),
                                                           work.as<unsigned long long>());
  RS_TRY(R->pairs.alloc(sizeof(uint64_t) * (size_t)U));
  RS_TRY(out_n.alloc(16));
  size_t bytes = 0;
  RS_ROC(rocprim::select(nullptr, bytes, uniq.as<uint64_t>(), flags.as<int32_t>(),
                         R->pairs.as<uint64_t>(), out_n.as<size_t>(), (size_t)U, st));
  if (tmp.bytes < bytes) RS_TRY(tmp.alloc(bytes));
  RS_ROC(rocprim::select(tmp.p, bytes, uniq.as<uint64_t>(), flags.as<int32_t>(),
                         R->pairs.as<uint64_t>(), out_n.as<size_t>(), (size_t)U, st));
  size_t hk = 0;
  unsigned long long hw[2] = {0, 0};
  RS_HIP(hipMemcpyAsync(&hk, out_n.p, sizeof(size_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(hw, work.p, sizeof(hw), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  R->num_pairs = (int64_t)hk;
  R->verify_probes = (int64_t)hw[0];
  R->verify_elements = (int64_t)hw[1];
  return RS_OK;
}

// cnt[r] = prefix entries of rank r (entries are sorted, rank in the top bits).
__global__ void rank_entries_kernel(const uint64_t* __restrict__ ent, int64_t M, int64_t V,
                                    int32_t* __restrict__ cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = (int64_t)(ent[i] >> kRankShift);
    if (r < V) atomicAdd(&cnt[r], 1);
  }
}

int64_t sss_candidate_bound(const rs_sss_index* X, const int32_t* h_tokens, int32_t n) {
  // the lookup's candidates of a query are the entries sharing a rank with
  // one of its prefix tokens (ranks ascending, tokens outside the collection
  // first and without entries), before the size window and deduplication
  std::vector<int64_t> keys((size_t)n);
  for (int32_t i = 0; i < n; ++i) {
    const int32_t t = h_tokens[i];
    keys[(size_t)i] = (t >= 0 && t < (int64_t)X->h_rank.size()) ? (int64_t)X->h_rank[(size_t)t] + 1 : 0;
  }
  std::sort(keys.begin(), keys.end());
  const int32_t pl = n > 0 ? prefix_len(n, X->p, X->q) : 0;
  int64_t bound = 0;
  for (int32_t i = 0; i < pl && i < n; ++i)
    if (keys[(size_t)i] > 0) bound += X->h_rank_entries[(size_t)(keys[(size_t)i] - 1)];
  return bound;
}

static rs_status index_build(const int32_t* d_off, int64_t n, const int32_t* d_el, int32_t p,
                             int32_t q, hipStream_t st, rs_sss_index* X) {
  RS_REQUIRE(n >= 0 && n < (1ll << 21), "rs_sss_index_build: %lld sets exceed 2^21", (long long)n);
  RS_REQUIRE(p > 0 && q >= p, "rs_sss_index_build: threshold p/q must be in (0, 1]");
  RS_TRY(check_ascending(d_off, n, d_el, "rs_sss_index_build", st));
  int32_t E32 = 0;
  if (n > 0) {
    RS_HIP(hipMemcpyAsync(&E32, d_off + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
  }
  X->off = d_off;
  X->el = d_el;
  X->n = n;
  X->E = E32;
  X->p = p;
  X->q = q;
  DevBuf tmp, scratch, ranked_sorted;
  RS_TRY(scratch.alloc(16));
  RS_HIP(hipMemsetAsync(scratch.p, 0, 16, st));
  RS_TRY(token_ranks(d_el, X->E, tmp, st, scratch.as<int32_t>(), X->rank, &X->V));
  RS_TRY(ranked_sets(d_off, n, d_el, X->E, X->rank, X->V, tmp, st, ranked_sorted));
  RS_TRY(prefix_entries(d_off, n, ranked_sorted, p, q, tmp, st, scratch.as<int32_t>(), X->ent, &X->M));
  // host copies for the serving loop's admission bound
  X->h_rank.resize((size_t)X->V);
  X->h_rank_entries.assign((size_t)X->V, 0);
  if (X->V > 0) {
    DevBuf cnt;
    RS_TRY(cnt.alloc(sizeof(int32_t) * (size_t)X->V));
    RS_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (size_t)X->V, st));
    if (X->M > 0)
      rank_entries_kernel<<<grid_for(X->M, 256, 8), 256, 0, st>>>(X->ent.as<uint64_t>(), X->M, X->V,
                                                                  cnt.as<int32_t>());
    RS_HIP(hipMemcpyAsync(X->h_rank.data(), X->rank.p, sizeof(int32_t) * (size_t)X->V, hipMemcpyDeviceToHost, st));
    RS_HIP(hipMemcpyAsync(X->h_rank_entries.data(), cnt.p, sizeof(int32_t) * (size_t)X->V,
                          hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
  }
  RS_HIP(hipStreamSynchronize(st));
  return RS_OK;
}

static rs_status lookup(const rs_sss_index* X, const int32_t* d_qoff, int64_t Q, const int32_t* d_qtok,
                        hipStream_t st, rs_sss_result* R) {
  RS_REQUIRE(Q >= 0 && Q < (1ll << 31), "rs_sss_lookup: bad query count");
  hipEvent_t ev[4];
  for (auto& e : ev) RS_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  RS_HIP(hipEventRecord(ev[0], st));
  int32_t EQ = 0;
  if (Q > 0) {
    RS_HIP(hipMemcpyAsync(&EQ, d_qoff + Q, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
  }
  RS_REQUIRE(EQ >= 0, "rs_sss_lookup: negative query offsets");
  RS_REQUIRE(EQ == 0 || d_qtok, "rs_sss_lookup: null query tokens");
  DevBuf tmp, scratch, keys, skeys, pl, poff;
  RS_TRY(scratch.alloc(16));
  RS_HIP(hipMemsetAsync(scratch.p, 0, 16, st));
  RS_TRY(keys.alloc(sizeof(uint32_t) * (size_t)(EQ > 0 ? EQ : 1)));
  RS_TRY(skeys.alloc(sizeof(uint32_t) * (size_t)(EQ > 0 ? EQ : 1)));
  RS_TRY(pl.alloc(sizeof(int32_t) * (size_t)(Q > 0 ? Q : 1)));
  RS_TRY(poff.alloc(sizeof(int64_t) * (size_t)(Q + 1)));
  if (Q > 0) {
    query_keys_kernel<<<grid_for(Q * kWave, 256, 8), 256, 0, st>>>(d_qoff, Q, d_qtok, X->rank.as<int32_t>(),
                                                                   X->V, keys.as<uint32_t>(),
                                                                   scratch.as<int32_t>());
    int32_t herr = 0;
    RS_HIP(hipMemcpyAsync(&herr, scratch.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    RS_REQUIRE(!herr, "rs_sss_lookup: a query is not strictly ascending, has a negative token or "
                      "2^21 or more tokens");
    if (EQ > 0) {
      size_t bytes = 0;
      const unsigned bits = bits_for((uint64_t)X->V + 1);
      RS_ROC(rocprim::segmented_radix_sort_keys(nullptr, bytes, keys.as<uint32_t>(), skeys.as<uint32_t>(),
                                                (size_t)EQ, (unsigned)Q, d_qoff, d_qoff + 1, 0, bits, st));
      if (tmp.bytes < bytes) RS_TRY(tmp.alloc(bytes));
      RS_ROC(rocprim::segmented_radix_sort_keys(tmp.p, bytes, keys.as<uint32_t>(), skeys.as<uint32_t>(),
                                                (size_t)EQ, (unsigned)Q, d_qoff, d_qoff + 1, 0, bits, st));
    }
    lookup_prefix_len_kernel<<<grid_for(Q, 256, 8), 256, 0, st>>>(d_qoff, Q, X->p, X->q, pl.as<int32_t>());
  }
  keys.release();
  RS_TRY(exclusive_scan_i32_to_i64(pl.as<int32_t>(), Q, poff.as<int64_t>(), tmp, st));
  int64_t NP = 0;
  RS_HIP(hipMemcpyAsync(&NP, poff.as<int64_t>() + Q, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  RS_HIP(hipEventRecord(ev[1], st));
  DevBuf lo, cnt, coff, cand, uniq;
  RS_TRY(lo.alloc(sizeof(int64_t) * (size_t)(NP > 0 ? NP : 1)));
  RS_TRY(cnt.alloc(sizeof(int64_t) * (size_t)(NP > 0 ? NP : 1)));
  RS_TRY(coff.alloc(sizeof(int64_t) * (size_t)(NP + 1)));
  if (NP > 0)
    lookup_count_kernel<<<grid_for(Q * kWave, 256, 8), 256, 0, st>>>(
        d_qoff, Q, skeys.as<uint32_t>(), poff.as<int64_t>(), X->ent.as<uint64_t>(), X->M, X->p, X->q,
        lo.as<int64_t>(), cnt.as<int64_t>());
  RS_TRY(exclusive_scan_i64(cnt.as<int64_t>(), NP, coff.as<int64_t>(), tmp, st));
  int64_t C = 0;
  RS_HIP(hipMemcpyAsync(&C, coff.as<int64_t>() + NP, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  if (C > (1ll << 32)) {
    set_error("rs_sss_lookup: %lld candidate pairs (before dedup) exceed the 2^32 working limit",
              (long long)C);
    return RS_ENOMEM;
  }
  int64_t U = 0;
  if (C > 0) {
    RS_TRY(cand.alloc(sizeof(uint64_t) * (size_t)C));
    lookup_emit_kernel<<<grid_for(Q * kWave, 256, 8), 256, 0, st>>>(poff.as<int64_t>(), Q, lo.as<int64_t>(),
                                                                    coff.as<int64_t>(), X->ent.as<uint64_t>(),
                                                                    cand.as<uint64_t>());
    RS_TRY(unique_sorted(cand, C, tmp, st, uniq, &U));
  }
  R->num_candidates = U;
  RS_HIP(hipEventRecord(ev[2], st));
  RS_TRY(verify_candidates(d_qoff, d_qtok, X->off, X->el, uniq, U, X->p, X->q, kVerifyAlpha, kVerifyB,
                           tmp, st, R));
  RS_HIP(hipEventRecord(ev[3], st));
  RS_HIP(hipEventSynchronize(ev[3]));
  for (int i = 0; i < 3; ++i) RS_HIP(hipEventElapsedTime(&R->ms[i], ev[i], ev[i + 1]));
  return RS_OK;
}

}  // namespace rs

using namespace rs;

extern "C" {

rs_status rs_sss_selfjoin(const int32_t* d_offsets, int64_t num_sets, const int32_t* d_elements,
                          int32_t p, int32_t q, float load_factor, int32_t bucket_size,
                          rs_stream stream, rs_sss_result** out) {
  RS_REQUIRE(out && d_offsets && (d_elements || num_sets == 0), "rs_sss_selfjoin: null argument");
  *out = nullptr;
  auto* R = new rs_sss_result();
  rs_status s = selfjoin(d_offsets, num_sets, d_elements, p, q, load_factor, bucket_size,
                         as_stream(stream), R);
  if (s != RS_OK) {
    delete R;
    return s;
  }
  *out = R;
  return RS_OK;
}

rs_status rs_sss_result_info(const rs_sss_result* r, int64_t* h_num_pairs,
                             int64_t* h_num_candidates, const int64_t** d_pairs) {
  RS_REQUIRE(r, "rs_sss_result_info: null result");
  if (h_num_pairs) *h_num_pairs = r->num_pairs;
  if (h_num_candidates) *h_num_candidates = r->num_candidates;
  if (d_pairs) *d_pairs = r->pairs.as<int64_t>();
  return RS_OK;
}

rs_status rs_sss_result_work(const rs_sss_result* r, int64_t* h_verify_probes,
                             int64_t* h_verify_elements) {
  RS_REQUIRE(r, "rs_sss_result_work: null result");
  if (h_verify_probes) *h_verify_probes = r->verify_probes;
  if (h_verify_elements) *h_verify_elements = r->verify_elements;
  return RS_OK;
}

rs_status rs_sss_result_timings(const rs_sss_result* r, float* h_ms3) {
  RS_REQUIRE(r && h_ms3, "rs_sss_result_timings: null argument");
  for (int i = 0; i < 3; ++i) h_ms3[i] = r->ms[i];
  return RS_OK;
}

void rs_sss_result_free(rs_sss_result* r) { delete r; }

rs_status rs_sss_index_build(const int32_t* d_offsets, int64_t num_sets, const int32_t* d_elements,
                             int32_t p, int32_t q, rs_stream stream, rs_sss_index** out) {
  RS_REQUIRE(out && d_offsets && (d_elements || num_sets == 0), "rs_sss_index_build: null argument");
  *out = nullptr;
  auto* X = new rs_sss_index();
  rs_status s = index_build(d_offsets, num_sets, d_elements, p, q, as_stream(stream), X);
  if (s != RS_OK) {
    delete X;
    return s;
  }
  *out = X;
  return RS_OK;
}

void rs_sss_index_free(rs_sss_index* idx) { delete idx; }

rs_status rs_sss_lookup(const rs_sss_index* idx, const int32_t* d_query_offsets,
                        int64_t num_queries, const int32_t* d_query_tokens, rs_stream stream,
                        rs_sss_result** out) {
  RS_REQUIRE(idx && out && d_query_offsets, "rs_sss_lookup: null argument");
  *out = nullptr;
  auto* R = new rs_sss_result();
  rs_status s = lookup(idx, d_query_offsets, num_queries, d_query_tokens, as_stream(stream), R);
  if (s != RS_OK) {
    delete R;
    return s;
  }
  *out = R;
  return RS_OK;
}

}  // extern "C"
