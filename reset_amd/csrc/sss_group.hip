// sss_group.hip -- grouped verification of similarity candidates.
//
// Candidates (u[i], v[i]) arrive sorted by u, and a set with a rare prefix
// token is a candidate of hundreds of others, so the pairs of one u form long
// runs.  A run (cut into units of at most kGroupCands candidates) is verified
// by one workgroup: it inserts S_u into an LDS table once (bucket-major,
// B-slot buckets with B in {1, 2, 4}, load factor alpha: the table's
// hyperparameters, 1/4 and 4 by default), then its waves stream the
// partner sets S_v through it, one candidate per wave at a time.  Per pair
// the cost is the streamed set alone; the build is paid once per unit.  A
// short candidate list (a few lookups) is cut into shorter units, down to 8
// candidates, so that it still spreads over the whole device.
//
// Probe rule (the same walk as TC.cu:11-31, a bucket at a time): inserts take
// the first EMPTY slot of their walk and never clear one, so a bucket's
// occupied slots are a prefix of it; a key is in its home bucket, or, if that
// bucket is full, further along -- the walk stops at a bucket with an EMPTY
// slot.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "sss_verify.hpp"

namespace rs {

constexpr int kGroupBlock = 256;
constexpr int kGroupWaves = kGroupBlock / kWave;
constexpr int64_t kGroupCands = 1024;

// LDS of one workgroup's table: 8192 int32 slots (32 KiB), so 4 workgroups
// share a CU.  The largest build set the grouped path takes is the largest
// size whose table fits (group_max_build): 2047 at alpha 1/4, b 4.
constexpr int32_t kGroupSlots = 8192;

template <int B>
__device__ __forceinline__ void group_bucket(const int32_t* __restrict__ tab, uint32_t j, int32_t x,
                                             bool& hit, bool& full) {
  if constexpr (B == 4) {
    const int4 v = *reinterpret_cast<const int4*>(tab + j * 4);
    const uint32_t m = min(min((uint32_t)(v.x ^ x), (uint32_t)(v.y ^ x)),
                           min((uint32_t)(v.z ^ x), (uint32_t)(v.w ^ x)));
    hit = m == 0u;
    full = v.w != kEmpty;
  } else if constexpr (B == 2) {
    const int2 v = *reinterpret_cast<const int2*>(tab + j * 2);
    hit = v.x == x || v.y == x;
    full = v.y != kEmpty;
  } else {
    const int32_t v = tab[j];
    hit = v == x;
    full = v != kEmpty;
  }
}

constexpr int64_t kGroupMinCands = 8;

// unit_start[k] = first candidate of unit k: a candidate opens a unit when
// its u differs from the previous one's or its index is a multiple of
// unit_len (<= kGroupCands).
__global__ void group_flags_kernel(const int32_t* __restrict__ cu, int64_t n, int64_t unit_len,
                                   int32_t* __restrict__ flags) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = i == 0 || cu[i] != cu[i - 1] || i % unit_len == 0;
}

// Streaming the partner sets: each wave takes the unit's candidates
// c0 + wid, c0 + wid + 4, ... and reads their sets as one stream of 256-token
// blocks (one 16-byte load per lane from the set's 16-byte aligned origin,
// clamped to the aligned group holding its last token).  The stream runs
// through three block buffers in fixed roles, across candidate boundaries,
// so two blocks are in flight while one is probed; every load is
// unconditional (past the wave's last candidate it repeats the last address),
// so the wait before a probe never covers the loads issued after its block.
// The candidates' set offsets are read into registers when the unit starts.
constexpr int32_t kStreamBlock = 4 * kWave;
typedef int v4i32 __attribute__((ext_vector_type(4)));
static_assert(kGroupCands / kGroupWaves <= 4 * kWave,
              "a wave's candidates of a unit must fit four registers per lane");

struct PartnerBlock {
  v4i32 x;
  int32_t t;    // candidate (wave-local index)
  int32_t pos;  // first element index of the block (from the set's aligned origin)
  int32_t nb;   // set size
  bool valid;
};

template <int B>
__global__ void __launch_bounds__(kGroupBlock) group_verify_kernel(
    float alpha, const int32_t* __restrict__ a_off, const int32_t* __restrict__ a_el,
    const int32_t* __restrict__ b_off, const int32_t* __restrict__ b_el,
    const int32_t* __restrict__ cu, const int32_t* __restrict__ cv,
    const int64_t* __restrict__ unit_start, int64_t nunits, int64_t ncand,
    int32_t* __restrict__ sizes, int32_t* __restrict__ err,
    unsigned long long* __restrict__ next_unit) {
  __shared__ __attribute__((aligned(16))) int32_t tab[kGroupSlots];
  __shared__ int64_t s_next;
  const int lane = __lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  bool bad = false;
  // Work stealing (TC.cu:176-182 takes its chunks the same way): a unit's
  // cost varies with its build set and its candidates, so workgroups take
  // units from a shared counter instead of a fixed deal.  Thread 0 takes the
  // next unit when a unit starts and publishes it after the build barrier
  // (every thread read the previous value before the first barrier), so the
  // atomic returns during the build; all threads read it after the unit.
  if (threadIdx.x == 0) s_next = (int64_t)atomicAdd(next_unit, 1ull);
  __syncthreads();
  int64_t k = s_next;
  while (k < nunits) {
    unsigned long long taken = 0;
    if (threadIdx.x == 0) taken = atomicAdd(next_unit, 1ull);
    const int64_t c0 = unit_start[k];
    const int64_t c1 = k + 1 < nunits ? unit_start[k + 1] : ncand;
    const int32_t u = cu[c0];
    const int32_t a0 = a_off[u], na = a_off[u + 1] - a0;
    // this wave's candidates: i_t = c0 + wid + 4t, t < ncw (<= 256); lane l
    // holds the set offset and size of candidate t = 64r + l in register r
    const int32_t ncw = c0 + wid < c1 ? (int32_t)((c1 - c0 - wid + kGroupWaves - 1) / kGroupWaves) : 0;
    int32_t mb0[4], mnb[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int32_t t = r * kWave + lane;
      const bool ok = t < ncw;
      const int32_t v = cv[ok ? c0 + wid + (int64_t)kGroupWaves * t : c0];
      const int32_t b0 = b_off[v];
      mb0[r] = b0;
      mnb[r] = ok ? b_off[v + 1] - b0 : 0;
    }
    // have them all before the stream starts (their first reads can come
    // deep inside it, where a wait would also cover the block loads)
    __asm__ volatile("" : : "v"(mb0[0]), "v"(mb0[1]), "v"(mb0[2]), "v"(mb0[3]), "v"(mnb[0]),
                     "v"(mnb[1]), "v"(mnb[2]), "v"(mnb[3]));
    const uint32_t L = (uint32_t)calc_length(na, alpha, B);
    const uint32_t nslots = L * B;  // <= kGroupSlots (build sets are <= group_max_build)
    for (uint32_t s = threadIdx.x; s < nslots / 4; s += kGroupBlock)
      *reinterpret_cast<int4*>(tab + 4 * s) = int4{kEmpty, kEmpty, kEmpty, kEmpty};
    __syncthreads();
    for (int32_t t = threadIdx.x; t < na; t += kGroupBlock) {
      const int32_t x = a_el[a0 + t];
      uint32_t slot = ((uint32_t)x & (L - 1u)) * B;
      uint32_t step = 0;
      for (; step < nslots; ++step) {
        if (atomicCAS(&tab[slot], kEmpty, x) == kEmpty) break;
        slot = slot + 1u == nslots ? 0u : slot + 1u;
      }
      bad |= step == nslots;
    }
    __syncthreads();
    if (threadIdx.x == 0) s_next = (int64_t)taken;
    if (ncw > 0) {  // wave-uniform
      auto meta = [&](const int32_t (&m)[4], int32_t t) {
        const int32_t r = t >> 6;
        const int32_t v = r == 0 ? m[0] : r == 1 ? m[1] : r == 2 ? m[2] : m[3];
        return __builtin_amdgcn_readlane(v, t & (kWave - 1));
      };
      // load cursor: candidate lt, block position lpos in its set G
      int32_t lt = 0, lpos = 0, lnb = 0, llast4 = 0;
      const int32_t* G = nullptr;
      // (the cursor is wave-uniform; readfirstlane keeps it in scalar
      // registers, apart from the loads' data)
      auto enter = [&](int32_t t) {
        lnb = __builtin_amdgcn_readfirstlane(meta(mnb, t));
        const int32_t b0 = __builtin_amdgcn_readfirstlane(meta(mb0, t));
        // an empty set reads (and ignores) the aligned group holding the
        // build set's first element
        G = lnb > 0 ? b_el + b0 : a_el + a0;
        lpos = __builtin_amdgcn_readfirstlane(-(int32_t)((reinterpret_cast<uintptr_t>(G) >> 2) & 3u));
        llast4 = __builtin_amdgcn_readfirstlane(lnb > 0 ? lpos + ((lnb - 1 - lpos) & ~3) : lpos);
      };
      enter(0);
      auto load_next = [&](PartnerBlock& b) {
        const int32_t e = lpos + 4 * lane;
        b.x = __builtin_nontemporal_load(reinterpret_cast<const v4i32*>(G + (e < llast4 ? e : llast4)));
        b.valid = lt < ncw;
        b.t = lt;
        b.pos = lpos;
        b.nb = lnb;
        if (lt < ncw) {
          lpos += kStreamBlock;
          if (lpos >= lnb && ++lt < ncw) enter(lt);
        }
      };
      // per-candidate counts are kept in registers (candidate t in lane
      // t % 64 of tot[t / 64]) and written after the stream: a store inside
      // it would count with the loads and blur the waits before the probes
      uint32_t hits = 0;
      int32_t tot[4] = {0, 0, 0, 0};
      auto probe = [&](const PartnerBlock& b) {
        // wait for the block on every path (an empty statement that reads
        // it), so the wait bookkeeping stays exact across the loop
        __asm__ volatile("" : : "v"(b.x.x));
        if (!b.valid) return false;
        const int x[4] = {b.x.x, b.x.y, b.x.z, b.x.w};
        uint32_t j[4];
        bool act[4], pend[4];
        bool any = false;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int32_t e = b.pos + 4 * lane + q;
          act[q] = e >= 0 && e < b.nb;
          j[q] = (uint32_t)x[q] & (L - 1u);
          bool hit, full;
          group_bucket<B>(tab, j[q], x[q], hit, full);
          hits += act[q] && hit;
          pend[q] = act[q] && !hit && full;
          any |= pend[q];
        }
        if (__any(any)) {  // rare: walk on past full home buckets
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            bool p = pend[q];
            uint32_t jj = j[q];
            for (uint32_t step = 1; step < L && __any(p); ++step) {
              jj = (jj + 1u) & (L - 1u);
              bool hit, full;
              group_bucket<B>(tab, jj, x[q], hit, full);
              hits += p && hit;
              p = p && !hit && full;
            }
            bad |= p;
          }
        }
        if (b.pos + kStreamBlock >= b.nb) {  // the candidate's last block
          const int32_t total = (int32_t)wave_sum(hits);
          const bool mine = lane == (b.t & (kWave - 1));
          const int32_t r = b.t >> 6;
#pragma unroll
          for (int q = 0; q < 4; ++q) tot[q] = q == r && mine ? total : tot[q];
          hits = 0;
        }
        return true;
      };
      PartnerBlock p0, p1, p2;
      load_next(p0);
      load_next(p1);
      while (true) {
        load_next(p2);
        if (!probe(p0)) break;
        load_next(p0);
        if (!probe(p1)) break;
        load_next(p1);
        if (!probe(p2)) break;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int32_t t = r * kWave + lane;
        if (t < ncw) sizes[c0 + wid + (int64_t)kGroupWaves * t] = tot[r];
      }
    }
    __syncthreads();
    k = s_next;
  }
  if (__any(bad) && lane == 0) atomicOr(err, kErrTableFull);
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ull; }

// Which candidates the grouped path takes: build set no larger than
// max_build.  sel[i] = 1 for those.
__global__ void group_select_kernel(const int32_t* __restrict__ a_off, const int32_t* __restrict__ cu,
                                    int64_t n, int32_t max_build, int32_t* __restrict__ sel) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = cu[i];
    sel[i] = a_off[u + 1] - a_off[u] <= max_build;
  }
}

int32_t group_max_build(float alpha, int32_t b) {
  // calculate_length is non-decreasing in the size: bisect for the largest
  // size whose table has at most kGroupSlots slots
  int32_t lo = 0, hi = kGroupSlots;  // a table of >= 8192 keys never fits
  while (hi - lo > 1) {
    const int32_t mid = lo + (hi - lo) / 2;
    const int32_t L = calc_length(mid, alpha, b);
    if (L > 0 && (int64_t)L * b <= kGroupSlots) lo = mid; else hi = mid;
  }
  const int32_t L0 = calc_length(lo, alpha, b);
  return L0 > 0 && (int64_t)L0 * b <= kGroupSlots ? lo : -1;
}

#define RS_ROC_G(expr)                                                \
  do {                                                                \
    hipError_t _e = (expr);                                           \
    if (_e != hipSuccess) {                                           \
      ::rs::set_error("%s failed: %s", #expr, hipGetErrorString(_e)); \
      return RS_EDEVICE;                                              \
    }                                                                 \
  } while (0)

rs_status group_verify(float alpha, int32_t b, const int32_t* a_off, const int32_t* a_el,
                       const int32_t* b_off, const int32_t* b_el, const int32_t* cu, const int32_t* cv,
                       int64_t n, int32_t* sizes, int32_t* d_err, DevBuf& tmp, hipStream_t st) {
  RS_REQUIRE(b == 1 || b == 2 || b == 4, "sss: grouped table bucket size %d (needs 1, 2 or 4)", b);
  if (n == 0) return RS_OK;
  DevBuf flags, starts, nsel;
  RS_TRY(flags.alloc(sizeof(int32_t) * (size_t)n));
  RS_TRY(starts.alloc(sizeof(int64_t) * (size_t)n));
  RS_TRY(nsel.alloc(16));
  const int grid = device_cu_count() * 4;  // 32 KiB of LDS each: 4 workgroups per CU
  // units of at most kGroupCands candidates, shorter when that leaves fewer
  // than two units per workgroup
  int64_t unit_len = (n + 2 * grid - 1) / (2 * grid);
  unit_len = unit_len < kGroupMinCands ? kGroupMinCands : unit_len > kGroupCands ? kGroupCands : unit_len;
  group_flags_kernel<<<grid_for(n, 256, 8), 256, 0, st>>>(cu, n, unit_len, flags.as<int32_t>());
  rocprim::counting_iterator<int64_t> idx(0);
  size_t bytes = 0;
  RS_ROC_G(rocprim::select(nullptr, bytes, idx, flags.as<int32_t>(), starts.as<int64_t>(),
                           nsel.as<size_t>(), (size_t)n, st));
  if (tmp.bytes < bytes) RS_TRY(tmp.alloc(bytes));
  RS_ROC_G(rocprim::select(tmp.p, bytes, idx, flags.as<int32_t>(), starts.as<int64_t>(),
                           nsel.as<size_t>(), (size_t)n, st));
  size_t nunits = 0;
  RS_HIP(hipMemcpyAsync(&nunits, nsel.p, sizeof(size_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  // the work-stealing counter: the second word of nsel (the first holds the
  // select count read above)
  auto* next_unit = reinterpret_cast<unsigned long long*>(nsel.as<size_t>() + 1);
  zero_u64_kernel<<<1, 1, 0, st>>>(next_unit);
  auto kern = b == 4 ? group_verify_kernel<4> : b == 2 ? group_verify_kernel<2> : group_verify_kernel<1>;
  kern<<<grid, kGroupBlock, 0, st>>>(alpha, a_off, a_el, b_off, b_el, cu, cv, starts.as<int64_t>(),
                                     (int64_t)nunits, n, sizes, d_err, next_unit);
  RS_HIP(hipGetLastError());
  RS_HIP(hipStreamSynchronize(st));  // flags/starts are released on return
  return RS_OK;
}

rs_status group_partition(const int32_t* a_off, const int32_t* cu, int64_t n, int32_t max_build,
                          int32_t* sel, hipStream_t st) {
  if (n > 0) group_select_kernel<<<grid_for(n, 256, 8), 256, 0, st>>>(a_off, cu, n, max_build, sel);
  RS_HIP(hipGetLastError());
  return RS_OK;
}

}  // namespace rs
