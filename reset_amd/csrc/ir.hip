// ir.hip -- conjunctive queries over an inverted index (information-retrieval/
// ir.cu:94-176) and a continuous-batching serving loop around them.
//
// Query kernels: term0's list (the shortest: queries are ordered by document
// frequency) is cut into segments of kIrSeg entries; each wave filters one
// segment through every later term's table, four keys per lane in flight,
// compacting survivors in place with a 64-bit ballot (stable, so survivors
// keep term0's order).  Because intersection distributes over the segments,
// a 2M-entry term0 list is spread over hundreds of waves instead of one.  A
// prefix scan of the segments' survivor counts then places every segment's
// survivors at once: in each query's result region (ir.cu:324-330 sizing) or
// densely, query after query, for the server.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tables.hpp"

namespace rs {

constexpr int kIrBlock = 256;
constexpr int64_t kIrSeg = 4096;
constexpr int64_t kIrRun = 4;  // segments per work-stealing take
constexpr int kIrKeys = 4;     // keys per lane in flight in the filter
constexpr int kIrXcds = 8;     // XCDs of the MI355X (per-XCD work ranges)

// per query: number of segments of term0's list (0 when it has no term or an
// empty term0 list); validates term ids.
__global__ void ir_seg_count_kernel(const int32_t* __restrict__ idx_off, int64_t nterms,
                                    const int32_t* __restrict__ query,
                                    const int32_t* __restrict__ qoff, int64_t Q,
                                    const int64_t* __restrict__ roff, int32_t* __restrict__ nseg,
                                    int32_t* status) {
  bool bad = false;
  for (int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; qi < Q;
       qi += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = qoff[qi], e = qoff[qi + 1];
    int32_t ns = 0;
    bool ok = e >= s;
    for (int32_t i = s; ok && i < e; ++i) ok = query[i] >= 0 && query[i] < nterms;
    if (ok && e > s) {
      const int32_t t0 = query[s];
      const int64_t n0 = idx_off[t0 + 1] - idx_off[t0];
      ok = roff[qi + 1] - roff[qi] == n0;
      ns = ok ? (int32_t)((n0 + kIrSeg - 1) / kIrSeg) : 0;
    }
    bad |= !ok;
    nseg[qi] = ns;
  }
  if (__any(bad) && __lane_id() == 0) atomicOr(status, kErrBadInput);
}

// seg_query[g] = the query owning segment g (segments of query q are
// segoff[q] .. segoff[q+1]).
__global__ void ir_seg_map_kernel(const int64_t* __restrict__ segoff, int64_t Q,
                                  int32_t* __restrict__ seg_query) {
  for (int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; qi < Q;
       qi += (int64_t)gridDim.x * blockDim.x)
    for (int64_t g = segoff[qi]; g < segoff[qi + 1]; ++g) seg_query[g] = (int32_t)qi;
}

// Filters kWave*U keys src[j0 ..] (cur of them valid) through one term's
// table and appends the survivors, in key order, to dst[w ..].  Each lane
// holds U keys (key j0 + u*64 + lane in register u) and walks their probe
// chains together: every step issues the next slot read of all its pending
// keys before any is compared, so U gathers per lane are in flight (the
// reference walks one key at a time, ir.cu:72-92).  The walk is TC.cu:11-31's
// -- slots k = 0..b-1 of the home bucket, then the next bucket mod L --
// stopping at the key (hit) or an EMPTY slot (miss), bounded at L*b slots.
template <int LAYOUT, int U>
__device__ __forceinline__ int32_t ir_filter_round(const int32_t* src, int32_t j0, int32_t cur,
                                                   const int32_t* __restrict__ Tt, int64_t BN,
                                                   int32_t b, uint32_t L, uint32_t M, int32_t* dst,
                                                   int32_t w, bool& bad) {
  const int lane = __lane_id();
  int32_t key[U];
  bool act[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int32_t j = j0 + u * kWave + lane;
    act[u] = j < cur;
    key[u] = act[u] ? src[j] : kEmpty;
  }
  uint32_t bucket[U];
  int32_t k[U];
  bool pend[U], hit[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    bucket[u] = home_bucket<LAYOUT>((uint32_t)key[u], L, M);
    k[u] = 0;
    pend[u] = act[u];
    hit[u] = false;
  }
  const uint32_t bound = L * (uint32_t)b;
  for (uint32_t step = 0; step < bound; ++step) {
    int32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = pend[u] ? Tt[bucket[u] + (int64_t)k[u] * BN] : kEmpty;
    bool any = false;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!pend[u]) continue;
      if (v[u] == key[u]) {
        hit[u] = true;
        pend[u] = false;
      } else if (v[u] == kEmpty) {
        pend[u] = false;
      } else if (++k[u] == b) {
        k[u] = 0;
        bucket[u] = (bucket[u] + 1u) & (L - 1u);
      }
      any |= pend[u];
    }
    if (!__any(any)) break;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) bad |= pend[u];
  // stable compaction (ir.cu:151-156): the write position never passes the
  // keys still to be read (they are all in registers already)
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint64_t m = __ballot(hit[u]);
    if (hit[u]) dst[w + __popcll(m & lanemask_lt())] = key[u];
    w += __popcll(m);
  }
  return w;
}

template <int LAYOUT, int U>
__global__ void __launch_bounds__(kIrBlock) ir_segment_kernel(
    const int32_t* __restrict__ idx_off, const int32_t* term0_docs,
    const int32_t* __restrict__ seg_query, const int32_t* __restrict__ Ls, const int64_t* __restrict__ toff, const int32_t* __restrict__ T,
    int64_t BN, int32_t b, uint32_t M, const int32_t* __restrict__ query, const int32_t* __restrict__ qoff,
    const int64_t* __restrict__ roff, const int64_t* __restrict__ segoff, int64_t total_segs,
    int32_t* ws, int32_t* __restrict__ seg_count, int32_t* __restrict__ status,
    unsigned long long* __restrict__ next_run, bool steal) {
  const int lane = __lane_id();
  bool bad = false;
  // A small batch (every serving batch) is dealt statically: wave w takes
  // segments w, w + W, ...  A large one is cut into kIrXcds contiguous ranges
  // of segments, one per XCD (blocks are dealt to the XCDs round robin), whose
  // waves take runs of kIrRun segments from their XCD's counter, the next run
  // when the current run's last segment starts, so that its atomic returns
  // meanwhile (ir.cu:94-176 takes chunks of queries from one counter): a
  // segment's cost varies with its query's term count and survivors, and a
  // query's segments, which probe the same tables, stay on one XCD.
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  const int xcd = (int)(blockIdx.x % kIrXcds);
  const int64_t xlo = steal ? total_segs * xcd / kIrXcds : 0;
  const int64_t xhi = steal ? total_segs * (xcd + 1) / kIrXcds : total_segs;
  unsigned long long* const ctr = next_run + xcd;
  auto take = [&]() -> int64_t {
    unsigned long long v = 0;
    if (lane == 0) v = atomicAdd(ctr, 1ull);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return xlo + (int64_t)(((uint64_t)hi << 32) | lo) * kIrRun;
  };
  int64_t q0 = 0, q1 = 0, p = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  int32_t r = 0;
  if (steal) {
    q0 = take();
    p = q0;
  }
  while (p < xhi) {
    if (steal && r == kIrRun - 1) q1 = take();
    const int64_t g = p;
    const int64_t qi = seg_query[g];
    const int64_t sidx = g - segoff[qi];
    const int32_t qs = qoff[qi], qe = qoff[qi + 1];
    const int32_t t0 = query[qs];
    const int64_t n0 = idx_off[t0 + 1] - idx_off[t0];
    const int64_t s0 = sidx * kIrSeg;
    const int32_t len = (int32_t)(n0 - s0 < kIrSeg ? n0 - s0 : kIrSeg);
    const int32_t* src = term0_docs + idx_off[t0] + s0;
    int32_t* dst = ws + roff[qi] + s0;
    int32_t cur = len;
    if (qe - qs == 1) {
      for (int32_t j = lane; j < len; j += kWave) dst[j] = src[j];
    }
    for (int32_t i = qs + 1; i < qe && cur > 0; ++i) {
      const int32_t term = query[i];
      const int32_t* Tt = T + toff[term];
      const uint32_t L = (uint32_t)Ls[term];
      int32_t w = 0;
      for (int32_t j0 = 0; j0 < cur; j0 += kWave * U)
        w = ir_filter_round<LAYOUT, U>(src, j0, cur, Tt, BN, b, L, M, dst, w, bad);
      cur = w;
      src = dst;
    }
    if (lane == 0) seg_count[g] = cur;
    if (!steal) {
      p += nwaves;
    } else if (++r < kIrRun) {
      p = q0 + r;
    } else {
      r = 0;
      q0 = q1;
      p = q0;
    }
    // (a run that reaches past xhi ends the loop there: the positions
    // beyond belong to the next XCD's range, and every later run of this
    // counter starts beyond them too)
  }
  if (__any(bad) && lane == 0) atomicOr(status, kErrProbeBound);
}

__global__ void ir_reset_runs_kernel(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// count[q] (and, for the dense form, out_off[q]) from seg_pos, the exclusive
// prefix of the segments' survivor counts in segment order.
__global__ void ir_counts_kernel(const int64_t* __restrict__ segoff, int64_t Q,
                                 const int64_t* __restrict__ seg_pos, int32_t* __restrict__ count,
                                 int64_t* __restrict__ out_off) {
  for (int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; qi < Q;
       qi += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = seg_pos[segoff[qi]], e = seg_pos[segoff[qi + 1]];
    count[qi] = (int32_t)(e - a);
    if (out_off) {
      out_off[qi] = a;
      if (qi == Q - 1) out_off[Q] = e;
    }
  }
}

// One wave per segment: its survivors (a stable run in term0 order) go to
// the query's result region after those of the query's earlier segments --
// result[roff[q] + seg_pos[g] - seg_pos[segoff[q]] ..], or to
// dense[seg_pos[g] ..] for the dense form (every query's result back to
// back, in query order).
__global__ void __launch_bounds__(kIrBlock) ir_copy_kernel(
    const int32_t* __restrict__ seg_query, const int64_t* __restrict__ segoff,
    const int64_t* __restrict__ roff, int64_t total, const int32_t* __restrict__ ws,
    const int32_t* __restrict__ seg_count, const int64_t* __restrict__ seg_pos,
    int32_t* __restrict__ result, int32_t* __restrict__ dense) {
  const int lane = __lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t g = wave; g < total; g += nwaves) {
    const int32_t c = seg_count[g];
    if (c == 0) continue;
    const int64_t qi = seg_query[g];
    const int64_t g0 = segoff[qi];
    const int32_t* src = ws + roff[qi] + (g - g0) * kIrSeg;
    int32_t* dst = dense ? dense + seg_pos[g] : result + roff[qi] + (seg_pos[g] - seg_pos[g0]);
    for (int32_t j = lane; j < c; j += kWave) dst[j] = src[j];
  }
}

// Segments below which a batch is dealt statically (-1: 4 runs per wave of
// the grid); rs_ir_set_steal_threshold changes it, so that tests can run the
// work-stealing path on small inputs.
static int64_t ir_steal_min = -1;

// Device scratch of one query batch, as plain pointers: carved from DevBufs
// (IrScratch, the allocating entry points and the server) or from a fixed
// region (the allocation-free entry point).
struct IrWork {
  int32_t* nseg = nullptr;       // [Q]
  int64_t* segoff = nullptr;     // [Q+1]
  int64_t* sums = nullptr;       // scan tile sums
  int64_t sums_cap = 0;
  int32_t* seg_query = nullptr;  // [S]
  int32_t* seg_count = nullptr;  // [S]
  int64_t* seg_pos = nullptr;    // [S+1]
  unsigned long long* next_run = nullptr;  // [kIrXcds] per-XCD work-stealing counters
};

// The allocating form: buffers grown as batches need them.
struct IrScratch {
  DevBuf nseg, segoff, sums, seg_query, seg_count, seg_pos, next_run;
  int64_t total_segs = 0;
};

template <typename T>
static rs_status grow(DevBuf& b, int64_t n) {
  const size_t need = sizeof(T) * (size_t)(n > 0 ? n : 1);
  if (b.bytes < need) RS_TRY(b.alloc(need));
  return RS_OK;
}

// Phase 1: segments per query and their offsets (segoff[Q] = total).
static rs_status ir_plan_segments(const rs_tables* t, const int32_t* query, const int32_t* qoff, int64_t Q,
                                  const int64_t* roff, const IrWork& w, hipStream_t st) {
  ir_seg_count_kernel<<<grid_for(Q, 256, 4), 256, 0, st>>>(t->d_off, t->n, query, qoff, Q, roff, w.nseg,
                                                          t->d_status());
  RS_TRY(exclusive_scan_i32_to_i64_buf(w.nseg, Q, w.segoff, w.sums, w.sums_cap, st));
  return RS_OK;
}

// Phase 2 (total segments known on the host): the segment filter, then the
// survivors' positions and the copy into result (roff layout) or dense.
static rs_status ir_run_segments(const rs_tables* t, int32_t layout, const int32_t* term0_docs,
                                 const int32_t* query, const int32_t* qoff, int64_t Q, const int64_t* roff,
                                 int32_t* result, int32_t* count, int32_t* ws, int32_t* dense,
                                 int64_t* out_off, int64_t total, const IrWork& w, hipStream_t st) {
  const int32_t* T = nullptr;
  RS_TRY(require_layout(t, layout, &T));
  if (t->timing) RS_HIP(hipEventRecord(t->ev0, st));
  if (total > 0) {
    ir_reset_runs_kernel<<<1, 64, 0, st>>>(w.next_run, kIrXcds);
    ir_seg_map_kernel<<<grid_for(Q, 256, 4), 256, 0, st>>>(w.segoff, Q, w.seg_query);
    int grid = grid_for(total * kWave, kIrBlock, 8);
    const int64_t nwaves = (int64_t)grid * (kIrBlock / kWave);
    const int64_t steal_min = ir_steal_min >= 0 ? ir_steal_min : 4 * kIrRun * nwaves;
    const bool steal = total >= steal_min;
    if (steal) grid = (grid + kIrXcds - 1) / kIrXcds * kIrXcds;
    auto kern = layout == RS_HIERARCHICAL ? ir_segment_kernel<RS_HIERARCHICAL, kIrKeys>
                                          : ir_segment_kernel<RS_NORMAL, kIrKeys>;
    kern<<<grid, kIrBlock, 0, st>>>(t->d_off, term0_docs, w.seg_query, t->lengths.as<int32_t>(),
                                   t->toff.as<int64_t>(), T, t->BN, t->b, (uint32_t)t->M, query, qoff, roff,
                                   w.segoff, total, ws, w.seg_count, t->d_status(), w.next_run, steal);
    RS_TRY(exclusive_scan_i32_to_i64_buf(w.seg_count, total, w.seg_pos, w.sums, w.sums_cap, st));
    ir_copy_kernel<<<grid_for(total * kWave, kIrBlock, 8), kIrBlock, 0, st>>>(
        w.seg_query, w.segoff, roff, total, ws, w.seg_count, w.seg_pos, result, dense);
  } else {
    ir_reset_runs_kernel<<<1, 64, 0, st>>>(reinterpret_cast<unsigned long long*>(w.seg_pos), 1);
  }
  ir_counts_kernel<<<grid_for(Q, 256, 4), 256, 0, st>>>(w.segoff, Q, w.seg_pos, count, out_off);
  if (t->timing) RS_HIP(hipEventRecord(t->ev1, st));
  RS_HIP(hipGetLastError());
  return RS_OK;
}

// Sizes the allocating scratch for Q queries / `total` segments (total < 0:
// phase-1 buffers only) and returns its pointer view.
static rs_status ir_scratch_view(IrScratch& sc, int64_t Q, int64_t total, IrWork& w) {
  RS_TRY(grow<int32_t>(sc.nseg, Q));
  RS_TRY(grow<int64_t>(sc.segoff, Q + 1));
  const int64_t tiles = scan_tile_count(Q > total ? Q : total) + 1;
  RS_TRY(grow<int64_t>(sc.sums, tiles));
  RS_TRY(grow<unsigned long long>(sc.next_run, kIrXcds));
  w.nseg = sc.nseg.as<int32_t>();
  w.segoff = sc.segoff.as<int64_t>();
  w.sums = sc.sums.as<int64_t>();
  w.sums_cap = (int64_t)(sc.sums.bytes / sizeof(int64_t));
  w.next_run = sc.next_run.as<unsigned long long>();
  if (total < 0) return RS_OK;
  RS_TRY(grow<int32_t>(sc.seg_query, total));
  RS_TRY(grow<int32_t>(sc.seg_count, total));
  RS_TRY(grow<int64_t>(sc.seg_pos, total + 1));
  w.seg_query = sc.seg_query.as<int32_t>();
  w.seg_count = sc.seg_count.as<int32_t>();
  w.seg_pos = sc.seg_pos.as<int64_t>();
  return RS_OK;
}

// Workspace-carrying enqueue used by the C entry point and by the server:
// with sync_total the total segment count is read back (one synchronization),
// else known_segs (the server computes it on the host).
static rs_status ir_enqueue(const rs_tables* t, int32_t layout, const int32_t* term0_docs,
                            const int32_t* query, const int32_t* qoff, int64_t Q,
                            const int64_t* roff, int32_t* result, int32_t* count, int32_t* ws,
                            int32_t* dense, int64_t* out_off, IrScratch& sc, hipStream_t st,
                            bool sync_total, int64_t known_segs) {
  const int32_t* T = nullptr;
  RS_TRY(require_layout(t, layout, &T));
  if (Q == 0) return RS_OK;
  IrWork w;
  RS_TRY(ir_scratch_view(sc, Q, -1, w));
  RS_TRY(ir_plan_segments(t, query, qoff, Q, roff, w, st));
  int64_t total = known_segs;
  if (sync_total) {
    RS_HIP(hipMemcpyAsync(&total, w.segoff + Q, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
  }
  sc.total_segs = total;
  RS_TRY(ir_scratch_view(sc, Q, total, w));
  return ir_run_segments(t, layout, term0_docs, query, qoff, Q, roff, result, count, ws, dense, out_off,
                         total, w, st);
}

// ---------------------------------------------------------------- no allocation
// rs_ir_query_noalloc: the query scratch in a fixed device array, sized for
// kNoAllocQueries queries and kNoAllocSegments term0 segments per call.  The
// array's name stands for a host-side symbol: host code reaches its device
// address only through hipGetSymbolAddress, once per device (no_alloc_base).
constexpr int64_t kNoAllocQueries = 1 << 16;
constexpr int64_t kNoAllocSegments = 1 << 18;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr int64_t kNoAllocSums = (kNoAllocSegments + 2047) / 2048 + 2;  // scan tiles of 2048
constexpr size_t kNaNseg = 0;
constexpr size_t kNaSegoff = kNaNseg + align256(sizeof(int32_t) * kNoAllocQueries);
constexpr size_t kNaSums = kNaSegoff + align256(sizeof(int64_t) * (kNoAllocQueries + 1));
constexpr size_t kNaSegQuery = kNaSums + align256(sizeof(int64_t) * kNoAllocSums);
constexpr size_t kNaSegCount = kNaSegQuery + align256(sizeof(int32_t) * kNoAllocSegments);
constexpr size_t kNaSegPos = kNaSegCount + align256(sizeof(int32_t) * kNoAllocSegments);
constexpr size_t kNaNextRun = kNaSegPos + align256(sizeof(int64_t) * (kNoAllocSegments + 1));
constexpr size_t kNoAllocBytes = kNaNextRun + align256(sizeof(unsigned long long) * kIrXcds);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on the current device, looked up on the first
// call on that device (with it current) and kept per device.
static rs_status no_alloc_base(unsigned char** out) {
  static std::mutex mu;
  static unsigned char* base[64] = {nullptr};
  int dev = 0;
  RS_HIP(hipGetDevice(&dev));
  RS_REQUIRE(dev >= 0 && dev < 64, "rs_ir_query_noalloc: device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lk(mu);
  if (!base[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("rs_ir_query_noalloc: hipGetSymbolAddress of the scratch array failed: %s",
                hipGetErrorString(e));
      return RS_EDEVICE;
    }
    base[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = base[dev];
  return RS_OK;
}

static IrWork no_alloc_view(unsigned char* base) {
  IrWork w;
  w.nseg = reinterpret_cast<int32_t*>(base + kNaNseg);
  w.segoff = reinterpret_cast<int64_t*>(base + kNaSegoff);
  w.sums = reinterpret_cast<int64_t*>(base + kNaSums);
  w.sums_cap = kNoAllocSums;
  w.seg_query = reinterpret_cast<int32_t*>(base + kNaSegQuery);
  w.seg_count = reinterpret_cast<int32_t*>(base + kNaSegCount);
  w.seg_pos = reinterpret_cast<int64_t*>(base + kNaSegPos);
  w.next_run = reinterpret_cast<unsigned long long*>(base + kNaNextRun);
  return w;
}

// ---------------------------------------------------------------- server
static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace rs

using namespace rs;

struct IrPending {
  int64_t id;
  double submit_ms;
  std::vector<int32_t> terms;
  int64_t n0;  // |term0 list|
};

// Bytes an IR query holds while admitted: its term0 entries twice (survivor
// workspace and dense result, 4 bytes each) plus its offsets and terms.
static int64_t ir_footprint(const IrPending& p) { return 8 * p.n0 + 64 + 4 * (int64_t)p.terms.size(); }


// Terms a served query may have (sizes the slots' preallocated query arrays).
constexpr int32_t kServeMaxTerms = 256;

// One in-flight batch of the serving loop: its stream, events and every
// buffer it needs, allocated once by rs_server_create for max_batch queries
// of up to kServeMaxTerms terms and cap_entries term0 entries in total (the
// slot's share of the memory budget), so the loop itself never allocates.
struct IrSlot {
  hipStream_t stream = nullptr;
  hipEvent_t ev_counts = nullptr, ev_done = nullptr;
  int state = 0;  // 0 free, 1 waiting for counts, 2 waiting for results
  std::vector<IrPending> batch;
  int64_t bytes = 0;
  int64_t cap_entries = 0, cap_segs = 0;
  // host staging (pinned)
  int32_t* h_query = nullptr;
  int32_t* h_qoff = nullptr;
  int64_t* h_roff = nullptr;
  int32_t* h_count = nullptr;
  int64_t* h_outoff = nullptr;
  int32_t* h_dense = nullptr;
  // device
  DevBuf d_query, d_qoff, d_roff, d_ws, d_count, d_outoff, d_dense;
  IrScratch sc;
  IrWork w;
};

struct rs_server {
  const rs_tables* t = nullptr;
  int32_t layout = 0;
  int device = 0;
  int64_t budget = 0;
  int64_t slot_budget = 0;  // budget / max_inflight: what one slot's buffers hold
  int32_t max_batch = 0;
  std::vector<int32_t> h_idx_off;  // host copy of the index offsets
  std::vector<std::unique_ptr<IrSlot>> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<IrPending> pending;
  std::deque<rs_completion> done;
  std::unordered_map<int64_t, std::vector<int32_t>> results;
  int64_t submitted = 0, completed = 0, launches = 0, admitted_bytes = 0, peak_bytes = 0;
  bool stop = false;
  bool paused = false;           // admission held (rs_server_pause)
  double queue_timeout_ms = 0.0; // 0 = none (rs_server_set_queue_timeout)
  rs_status worker_error = RS_OK;
  std::thread worker;
  // similarity lookups (second request kind), served by their own thread
  const rs_sss_index* sss = nullptr;
  std::deque<IrPending> sss_pending;  // terms = the query set's tokens
  std::thread sss_worker;
  hipStream_t sss_stream = nullptr;
  // IR slots in use while lookups are attached: one hardware queue fewer
  // than HIP gives the process, so the lookup stream keeps a queue of its own
  int32_t ir_slots_with_sss = 0;
  int32_t* h_sss_tok = nullptr;   // pinned staging
  int32_t* h_sss_off = nullptr;
  int64_t* h_sss_keys = nullptr;
  int64_t cap_sss_tok = 0, cap_sss_off = 0, cap_sss_keys = 0;
};

template <typename T>
static rs_status host_alloc(T*& p, int64_t n) {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  if (hipHostMalloc((void**)&p, sizeof(T) * (size_t)n, 0) != hipSuccess) {
    p = nullptr;
    set_error("ir server: hipHostMalloc of %lld entries failed", (long long)n);
    return RS_ENOMEM;
  }
  return RS_OK;
}

// Allocates a slot's pinned staging and device buffers (once, at creation).
static rs_status slot_init(IrSlot& s, int64_t max_batch, int64_t cap_entries) {
  const int64_t nq = max_batch + 1, nt = max_batch * kServeMaxTerms;
  s.cap_entries = cap_entries;
  s.cap_segs = cap_entries / kIrSeg + max_batch;
  RS_TRY(host_alloc(s.h_qoff, nq));
  RS_TRY(host_alloc(s.h_roff, nq));
  RS_TRY(host_alloc(s.h_count, nq));
  RS_TRY(host_alloc(s.h_outoff, nq));
  RS_TRY(host_alloc(s.h_query, nt));
  RS_TRY(host_alloc(s.h_dense, cap_entries));
  RS_TRY(s.d_qoff.alloc(sizeof(int32_t) * (size_t)nq));
  RS_TRY(s.d_roff.alloc(sizeof(int64_t) * (size_t)nq));
  RS_TRY(s.d_count.alloc(sizeof(int32_t) * (size_t)nq));
  RS_TRY(s.d_outoff.alloc(sizeof(int64_t) * (size_t)nq));
  RS_TRY(s.d_query.alloc(sizeof(int32_t) * (size_t)nt));
  RS_TRY(s.d_ws.alloc(sizeof(int32_t) * (size_t)cap_entries));
  RS_TRY(s.d_dense.alloc(sizeof(int32_t) * (size_t)cap_entries));
  return ir_scratch_view(s.sc, max_batch, s.cap_segs, s.w);
}

// Launch a batch on a slot: H2D query arrays, the query kernels writing the
// results densely (query order), D2H counts and offsets.  Within the slot's
// preallocated capacity by admission.
static rs_status slot_launch(rs_server* S, IrSlot& s) {
  const int64_t Q = (int64_t)s.batch.size();
  int64_t nterms = 0, total = 0, segs = 0;
  for (auto& p : s.batch) {
    nterms += (int64_t)p.terms.size();
    total += p.n0;
    segs += (p.n0 + kIrSeg - 1) / kIrSeg;
  }
  RS_REQUIRE(Q <= S->max_batch && nterms <= (int64_t)S->max_batch * kServeMaxTerms &&
                 total <= s.cap_entries && segs <= s.cap_segs,
             "ir server: batch exceeds the slot's preallocated capacity");
  int64_t w = 0, r = 0;
  for (int64_t i = 0; i < Q; ++i) {
    s.h_qoff[i] = (int32_t)w;
    s.h_roff[i] = r;
    for (int32_t x : s.batch[i].terms) s.h_query[w++] = x;
    r += s.batch[i].n0;
  }
  s.h_qoff[Q] = (int32_t)w;
  s.h_roff[Q] = r;
  hipStream_t st = s.stream;
  if (nterms > 0)
    RS_HIP(hipMemcpyAsync(s.d_query.p, s.h_query, sizeof(int32_t) * (size_t)nterms, hipMemcpyHostToDevice, st));
  RS_HIP(hipMemcpyAsync(s.d_qoff.p, s.h_qoff, sizeof(int32_t) * (size_t)(Q + 1), hipMemcpyHostToDevice, st));
  RS_HIP(hipMemcpyAsync(s.d_roff.p, s.h_roff, sizeof(int64_t) * (size_t)(Q + 1), hipMemcpyHostToDevice, st));
  RS_TRY(ir_plan_segments(S->t, s.d_query.as<int32_t>(), s.d_qoff.as<int32_t>(), Q, s.d_roff.as<int64_t>(),
                          s.w, st));
  RS_TRY(ir_run_segments(S->t, S->layout, S->t->d_elems, s.d_query.as<int32_t>(), s.d_qoff.as<int32_t>(), Q,
                         s.d_roff.as<int64_t>(), nullptr, s.d_count.as<int32_t>(), s.d_ws.as<int32_t>(),
                         s.d_dense.as<int32_t>(), s.d_outoff.as<int64_t>(), segs, s.w, st));
  RS_HIP(hipMemcpyAsync(s.h_count, s.d_count.p, sizeof(int32_t) * (size_t)Q, hipMemcpyDeviceToHost, st));
  RS_HIP(hipMemcpyAsync(s.h_outoff, s.d_outoff.p, sizeof(int64_t) * (size_t)(Q + 1),
                        hipMemcpyDeviceToHost, st));
  RS_HIP(hipEventRecord(s.ev_counts, st));
  s.state = 1;
  return RS_OK;
}

// With the server's lock held: complete, with RS_ETIMEDOUT, the queued
// requests (oldest first) that waited longer than the queue timeout.
static void expire_locked(rs_server* S, std::deque<IrPending>& q, int32_t kind) {
  if (S->queue_timeout_ms <= 0.0) return;
  const double t = now_ms();
  bool any = false;
  while (!q.empty() && t - q.front().submit_ms > S->queue_timeout_ms) {
    S->done.push_back(rs_completion{q.front().id, -1, RS_ETIMEDOUT, q.front().submit_ms, t, kind, 0});
    S->completed++;
    q.pop_front();
    any = true;
  }
  if (any) S->cv.notify_all();
}

static void server_loop(rs_server* S) {
  (void)hipSetDevice(S->device);
  while (true) {
    bool progressed = false;
    // 1. advance in-flight slots
    for (auto& sp : S->slots) {
      IrSlot& s = *sp;
      if (s.state == 1 && hipEventQuery(s.ev_counts) == hipSuccess) {
        const int64_t Q = (int64_t)s.batch.size();
        const int64_t total = s.h_outoff[Q];
        if (total > 0)
          (void)hipMemcpyAsync(s.h_dense, s.d_dense.p, sizeof(int32_t) * (size_t)total,
                               hipMemcpyDeviceToHost, s.stream);
        (void)hipEventRecord(s.ev_done, s.stream);
        s.state = 2;
        progressed = true;
      }
      if (s.state == 2 && hipEventQuery(s.ev_done) == hipSuccess) {
        const double t = now_ms();
        std::lock_guard<std::mutex> lk(S->mu);
        for (size_t i = 0; i < s.batch.size(); ++i) {
          const int32_t c = s.h_count[i];
          const int64_t o = s.h_outoff[i];
          S->results[s.batch[i].id] = std::vector<int32_t>(s.h_dense + o, s.h_dense + o + c);
          S->done.push_back(rs_completion{s.batch[i].id, c, RS_OK, s.batch[i].submit_ms, t});
        }
        S->completed += (int64_t)s.batch.size();
        S->admitted_bytes -= s.bytes;
        s.batch.clear();
        s.bytes = 0;
        s.state = 0;
        progressed = true;
        S->cv.notify_all();
      }
    }
    // 2. admit pending queries into free slots (FIFO, memory-bounded)
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (S->stop) break;
      if (!S->paused) expire_locked(S, S->pending, RS_REQ_IR);
      const size_t active = S->sss ? (size_t)S->ir_slots_with_sss : S->slots.size();
      for (size_t si = 0; si < active; ++si) {
        IrSlot& s = *S->slots[si];
        if (S->paused || s.state != 0 || S->pending.empty()) continue;
        int64_t bytes = 0;
        while (!S->pending.empty() && (int32_t)s.batch.size() < S->max_batch) {
          const IrPending& p = S->pending.front();
          const int64_t fb = ir_footprint(p);
          // the slot's share of the budget (what its buffers hold), and the
          // budget as a whole
          if (bytes + fb > S->slot_budget || S->admitted_bytes + bytes + fb > S->budget) break;
          bytes += fb;
          s.batch.push_back(p);
          S->pending.pop_front();
        }
        if (s.batch.empty()) continue;
        s.bytes = bytes;
        S->admitted_bytes += bytes;
        S->peak_bytes = S->admitted_bytes > S->peak_bytes ? S->admitted_bytes : S->peak_bytes;
        S->launches++;
        lk.unlock();
        rs_status st = slot_launch(S, s);
        lk.lock();
        if (st != RS_OK) {
          const double t = now_ms();
          for (auto& p : s.batch) S->done.push_back(rs_completion{p.id, -1, st, p.submit_ms, t});
          S->completed += (int64_t)s.batch.size();
          S->admitted_bytes -= s.bytes;
          s.batch.clear();
          s.bytes = 0;
          s.state = 0;
          S->worker_error = st;
          S->cv.notify_all();
        }
        progressed = true;
      }
      if (!progressed) {
        bool busy = false;
        for (auto& sp : S->slots) busy |= sp->state != 0;
        if (busy)
          S->cv.wait_for(lk, std::chrono::microseconds(20));
        else
          S->cv.wait_for(lk, std::chrono::milliseconds(2),
                         [S] { return S->stop || (!S->paused && !S->pending.empty()); });
      }
    }
  }
}

// Footprint charged to a lookup while it is admitted: its tokens, and 32
// bytes for each candidate it can produce (key, pair ids, size and flag, sort
// space), from the index's bound computed at submit (kept in n0).
static int64_t sss_footprint(const IrPending& p) { return 1024 + 8 * (int64_t)p.terms.size() + 32 * p.n0; }

// One lookup batch: stage the query sets, run the lookup on the server's
// lookup stream (the pipeline synchronizes that stream between its stages),
// split the (query << 32 | set) keys into per-request results.
static rs_status sss_run_batch(rs_server* S, std::vector<IrPending>& batch) {
  const int64_t Q = (int64_t)batch.size();
  int64_t ntok = 0;
  for (auto& p : batch) ntok += (int64_t)p.terms.size();
  if (Q + 1 > S->cap_sss_off) {
    RS_TRY(host_alloc(S->h_sss_off, 2 * (Q + 1)));
    S->cap_sss_off = 2 * (Q + 1);
  }
  if ((ntok > 0 ? ntok : 1) > S->cap_sss_tok) {
    RS_TRY(host_alloc(S->h_sss_tok, 2 * (ntok > 0 ? ntok : 1)));
    S->cap_sss_tok = 2 * (ntok > 0 ? ntok : 1);
  }
  int64_t w = 0;
  for (int64_t i = 0; i < Q; ++i) {
    S->h_sss_off[i] = (int32_t)w;
    for (int32_t x : batch[i].terms) S->h_sss_tok[w++] = x;
  }
  S->h_sss_off[Q] = (int32_t)w;
  hipStream_t st = S->sss_stream;
  DevBuf d_off, d_tok;
  RS_TRY(d_off.alloc(sizeof(int32_t) * (size_t)(Q + 1)));
  RS_TRY(d_tok.alloc(sizeof(int32_t) * (size_t)(ntok > 0 ? ntok : 1)));
  RS_HIP(hipMemcpyAsync(d_off.p, S->h_sss_off, sizeof(int32_t) * (size_t)(Q + 1), hipMemcpyHostToDevice, st));
  if (ntok > 0)
    RS_HIP(hipMemcpyAsync(d_tok.p, S->h_sss_tok, sizeof(int32_t) * (size_t)ntok, hipMemcpyHostToDevice, st));
  rs_sss_result* R = nullptr;
  RS_TRY(rs_sss_lookup(S->sss, d_off.as<int32_t>(), Q, d_tok.as<int32_t>(), reinterpret_cast<rs_stream>(st), &R));
  struct RGuard {
    rs_sss_result* r;
    ~RGuard() { rs_sss_result_free(r); }
  } rg{R};
  int64_t K = 0;
  const int64_t* d_keys = nullptr;
  RS_TRY(rs_sss_result_info(R, &K, nullptr, &d_keys));
  if (K > S->cap_sss_keys) {
    RS_TRY(host_alloc(S->h_sss_keys, 2 * K));
    S->cap_sss_keys = 2 * K;
  }
  if (K > 0) {
    RS_HIP(hipMemcpyAsync(S->h_sss_keys, d_keys, sizeof(int64_t) * (size_t)K, hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
  }
  const double t = now_ms();
  std::vector<std::vector<int32_t>> per((size_t)Q);
  for (int64_t i = 0; i < K; ++i) {
    const uint64_t k = (uint64_t)S->h_sss_keys[i];
    per[(size_t)(k >> 32)].push_back((int32_t)(k & 0xFFFFFFFFull));
  }
  std::lock_guard<std::mutex> lk(S->mu);
  for (int64_t i = 0; i < Q; ++i) {
    const int32_t c = (int32_t)per[(size_t)i].size();
    S->results[batch[(size_t)i].id] = std::move(per[(size_t)i]);
    S->done.push_back(rs_completion{batch[(size_t)i].id, c, RS_OK, batch[(size_t)i].submit_ms, t,
                                    RS_REQ_SSS, 0});
  }
  return RS_OK;
}

static void sss_loop(rs_server* S) {
  (void)hipSetDevice(S->device);
  // the lookups' device buffers come from a cache kept for this thread: a
  // hipFree per batch would wait for the whole device, IR batches included
  BlockCache blocks((size_t)S->budget);
  BlockCacheScope cached(&blocks);
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    S->cv.wait(lk, [S] { return S->stop || (!S->paused && !S->sss_pending.empty()); });
    if (S->stop) break;
    expire_locked(S, S->sss_pending, RS_REQ_SSS);
    // admit FIFO while the batch fits the shared memory budget
    std::vector<IrPending> batch;
    int64_t bytes = 0;
    while (!S->sss_pending.empty() && (int32_t)batch.size() < S->max_batch) {
      const int64_t fb = sss_footprint(S->sss_pending.front());
      // over budget: wait for completions, unless nothing at all is admitted
      if (S->admitted_bytes + bytes + fb > S->budget && !(batch.empty() && S->admitted_bytes == 0))
        break;
      bytes += fb;
      batch.push_back(std::move(S->sss_pending.front()));
      S->sss_pending.pop_front();
    }
    if (batch.empty()) {
      S->cv.wait_for(lk, std::chrono::microseconds(200));
      continue;
    }
    S->admitted_bytes += bytes;
    S->peak_bytes = S->admitted_bytes > S->peak_bytes ? S->admitted_bytes : S->peak_bytes;
    S->launches++;
    lk.unlock();
    rs_status st = sss_run_batch(S, batch);
    lk.lock();
    if (st != RS_OK) {
      const double t = now_ms();
      for (auto& p : batch)
        S->done.push_back(rs_completion{p.id, -1, st, p.submit_ms, t, RS_REQ_SSS, 0});
      S->worker_error = st;
    }
    S->completed += (int64_t)batch.size();
    S->admitted_bytes -= bytes;
    S->cv.notify_all();
  }
}

extern "C" {

rs_status rs_ir_query_async(const rs_tables* t, int32_t layout, const int32_t* d_term0_docs,
                            const int32_t* d_query, const int32_t* d_query_offsets,
                            int64_t num_queries, const int64_t* d_result_offsets,
                            int32_t* d_result, int32_t* d_result_count, int32_t* d_workspace,
                            rs_stream stream) {
  RS_REQUIRE(t, "rs_ir_query_async: null tables");
  RS_REQUIRE(num_queries >= 0, "rs_ir_query_async: negative query count");
  if (num_queries == 0) return RS_OK;
  RS_REQUIRE(d_query_offsets && d_result_offsets && d_result_count && d_result && d_workspace,
             "rs_ir_query_async: null argument");
  hipStream_t st = as_stream(stream);
  IrScratch sc;
  RS_TRY(ir_enqueue(t, layout, d_term0_docs ? d_term0_docs : t->d_elems, d_query, d_query_offsets,
                    num_queries, d_result_offsets, d_result, d_result_count, d_workspace, nullptr, nullptr,
                    sc, st, true, 0));
  RS_HIP(hipStreamSynchronize(st));  // scratch is released on return
  return RS_OK;
}

rs_status rs_ir_query_noalloc(const rs_tables* t, int32_t layout, const int32_t* d_term0_docs,
                              const int32_t* d_query, const int32_t* d_query_offsets,
                              int64_t num_queries, const int64_t* d_result_offsets,
                              int32_t* d_result, int32_t* d_result_count, int32_t* d_workspace,
                              rs_stream stream) {
  RS_REQUIRE(t, "rs_ir_query_noalloc: null tables");
  RS_REQUIRE(num_queries >= 0 && num_queries <= kNoAllocQueries,
             "rs_ir_query_noalloc: %lld queries (at most %lld per call)", (long long)num_queries,
             (long long)kNoAllocQueries);
  if (num_queries == 0) return RS_OK;
  RS_REQUIRE(d_query_offsets && d_result_offsets && d_result_count && d_result && d_workspace,
             "rs_ir_query_noalloc: null argument");
  const int32_t* T = nullptr;
  RS_TRY(require_layout(t, layout, &T));
  hipStream_t st = as_stream(stream);
  unsigned char* base = nullptr;
  RS_TRY(no_alloc_base(&base));
  const IrWork w = no_alloc_view(base);
  const int32_t* term0 = d_term0_docs ? d_term0_docs : t->d_elems;
  RS_TRY(ir_plan_segments(t, d_query, d_query_offsets, num_queries, d_result_offsets, w, st));
  int64_t total = 0;
  RS_HIP(hipMemcpyAsync(&total, w.segoff + num_queries, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RS_HIP(hipStreamSynchronize(st));
  RS_REQUIRE(total <= kNoAllocSegments,
             "rs_ir_query_noalloc: the batch has %lld term0 segments of %lld entries (at most %lld per "
             "call)", (long long)total, (long long)kIrSeg, (long long)kNoAllocSegments);
  return ir_run_segments(t, layout, term0, d_query, d_query_offsets, num_queries, d_result_offsets, d_result,
                         d_result_count, d_workspace, nullptr, nullptr, total, w, st);
}

rs_status rs_ir_set_steal_threshold(int64_t segments) {
  RS_REQUIRE(segments >= -1, "rs_ir_set_steal_threshold: %lld is below -1", (long long)segments);
  ir_steal_min = segments;
  return RS_OK;
}

rs_status rs_server_create(const rs_tables* t, int32_t layout, int64_t memory_budget_bytes,
                              int32_t max_batch, int32_t max_inflight, rs_server** out) {
  RS_REQUIRE(t && out, "rs_server_create: null argument");
  RS_REQUIRE(memory_budget_bytes > 0 && max_batch > 0 && max_batch <= (1 << 20) && max_inflight > 0 &&
                 max_inflight <= 16 && memory_budget_bytes / max_inflight >= 4096,
             "rs_server_create: bad limits (budget %lld bytes, batch %d, inflight %d)",
             (long long)memory_budget_bytes, max_batch, max_inflight);
  const int32_t* T = nullptr;
  RS_TRY(require_layout(t, layout, &T));
  *out = nullptr;
  auto* S = new rs_server();
  S->t = t;
  S->layout = layout;
  S->budget = memory_budget_bytes;
  S->max_batch = max_batch;
  (void)hipGetDevice(&S->device);
  S->h_idx_off.resize((size_t)t->n + 1);
  if (hipMemcpy(S->h_idx_off.data(), t->d_off, sizeof(int32_t) * (size_t)(t->n + 1),
                hipMemcpyDeviceToHost) != hipSuccess) {
    delete S;
    set_error("rs_server_create: cannot read index offsets");
    return RS_EDEVICE;
  }
  S->slot_budget = memory_budget_bytes / max_inflight;
  const int64_t cap_entries = S->slot_budget / 8;
  // HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless the
  // environment says otherwise); two streams on one queue serialize.  The
  // lookup stream is created right after the slots that stay in use when
  // lookups are attached, so it does not share their queues.
  int32_t hw_queues = 4;
  if (const char* env = std::getenv("GPU_MAX_HW_QUEUES")) {
    const int v = std::atoi(env);
    if (v > 0) hw_queues = v;
  }
  S->ir_slots_with_sss = std::max(1, std::min(max_inflight, hw_queues - 1));
  for (int i = 0; i < max_inflight; ++i) S->slots.push_back(std::make_unique<IrSlot>());
  for (int i = 0; i < max_inflight; ++i) {
    IrSlot& s = *S->slots[(size_t)i];
    if (i == S->ir_slots_with_sss &&
        hipStreamCreateWithFlags(&S->sss_stream, hipStreamNonBlocking) != hipSuccess) {
      S->sss_stream = nullptr;
      rs_server_free(S);
      set_error("rs_server_create: stream creation failed");
      return RS_EDEVICE;
    }
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s.ev_counts, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.ev_done, hipEventDisableTiming) != hipSuccess) {
      rs_server_free(S);
      set_error("rs_server_create: stream/event creation failed");
      return RS_EDEVICE;
    }
    const rs_status st = slot_init(s, max_batch, cap_entries);
    if (st != RS_OK) {
      const std::string msg = last_error();
      rs_server_free(S);
      set_error("rs_server_create: %s", msg.c_str());
      return st;
    }
  }
  if (!S->sss_stream && hipStreamCreateWithFlags(&S->sss_stream, hipStreamNonBlocking) != hipSuccess) {
    S->sss_stream = nullptr;
    rs_server_free(S);
    set_error("rs_server_create: stream creation failed");
    return RS_EDEVICE;
  }
  S->worker = std::thread(server_loop, S);
  *out = S;
  return RS_OK;
}

rs_status rs_server_submit_ir(rs_server* S, int64_t query_id, const int32_t* h_terms,
                              int32_t num_terms) {
  RS_REQUIRE(S && (h_terms || num_terms == 0) && num_terms >= 0, "rs_server_submit: bad args");
  IrPending p;
  p.id = query_id;
  p.submit_ms = now_ms();
  p.terms.assign(h_terms, h_terms + num_terms);
  for (int32_t x : p.terms)
    RS_REQUIRE(x >= 0 && x < S->t->n, "rs_server_submit: term id %d out of range", x);
  p.n0 = num_terms > 0 ? S->h_idx_off[p.terms[0] + 1] - S->h_idx_off[p.terms[0]] : 0;
  RS_REQUIRE(num_terms <= kServeMaxTerms, "rs_server_submit: a query has more than %d terms", kServeMaxTerms);
  if (ir_footprint(p) > S->slot_budget) {
    set_error("rs_server_submit: query %lld needs %lld bytes, above a slot's share of the memory "
              "budget (%lld bytes)", (long long)query_id, (long long)ir_footprint(p),
              (long long)S->slot_budget);
    return RS_ENOMEM;
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->pending.size() >= (size_t)1 << 22) {
      set_error("rs_server_submit: pending queue full");
      return RS_EBUSY;
    }
    S->pending.push_back(std::move(p));
    S->submitted++;
  }
  S->cv.notify_all();
  return RS_OK;
}

rs_status rs_server_attach_sss(rs_server* S, const rs_sss_index* idx) {
  RS_REQUIRE(S && idx, "rs_server_attach_sss: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  RS_REQUIRE(!S->sss, "rs_server_attach_sss: an index is already attached");
  S->sss = idx;
  S->sss_worker = std::thread(sss_loop, S);
  return RS_OK;
}

rs_status rs_server_submit_sss(rs_server* S, int64_t query_id, const int32_t* h_tokens,
                               int32_t num_tokens) {
  RS_REQUIRE(S && (h_tokens || num_tokens == 0) && num_tokens >= 0, "rs_server_submit_sss: bad args");
  RS_REQUIRE(num_tokens < (1 << 21), "rs_server_submit_sss: a query has 2^21 or more tokens");
  IrPending p;
  p.id = query_id;
  p.submit_ms = now_ms();
  p.terms.assign(h_tokens, h_tokens + num_tokens);
  for (int32_t i = 0; i < num_tokens; ++i)
    RS_REQUIRE(p.terms[i] >= 0 && (i == 0 || p.terms[i - 1] < p.terms[i]),
               "rs_server_submit_sss: tokens must be non-negative and strictly ascending");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    RS_REQUIRE(S->sss, "rs_server_submit_sss: no similarity index attached");
    p.n0 = sss_candidate_bound(S->sss, p.terms.data(), num_tokens);
    if (S->sss_pending.size() >= (size_t)1 << 22) {
      set_error("rs_server_submit_sss: pending queue full");
      return RS_EBUSY;
    }
    S->sss_pending.push_back(std::move(p));
    S->submitted++;
  }
  S->cv.notify_all();
  return RS_OK;
}

rs_status rs_server_poll(rs_server* S, rs_completion* h_out, int32_t cap, int32_t* h_n) {
  RS_REQUIRE(S && h_n && (h_out || cap == 0), "rs_server_poll: bad args");
  std::lock_guard<std::mutex> lk(S->mu);
  int32_t n = 0;
  while (n < cap && !S->done.empty()) {
    h_out[n++] = S->done.front();
    S->done.pop_front();
  }
  *h_n = n;
  return S->worker_error;
}

rs_status rs_server_result(rs_server* S, int64_t query_id, int32_t* h_docs, int64_t cap,
                              int64_t* h_count) {
  RS_REQUIRE(S && h_count, "rs_server_result: bad args");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->results.find(query_id);
  RS_REQUIRE(it != S->results.end(), "rs_server_result: query %lld has no result",
             (long long)query_id);
  *h_count = (int64_t)it->second.size();
  RS_REQUIRE(h_docs || it->second.empty() || cap == 0, "rs_server_result: null buffer");
  const int64_t n = *h_count < cap ? *h_count : cap;
  if (n > 0) memcpy(h_docs, it->second.data(), sizeof(int32_t) * (size_t)n);
  if (n == *h_count) S->results.erase(it);  // kept until fully copied out
  return RS_OK;
}

rs_status rs_server_cancel(rs_server* S, int64_t query_id) {
  RS_REQUIRE(S, "rs_server_cancel: null server");
  std::lock_guard<std::mutex> lk(S->mu);
  for (auto* q : {&S->pending, &S->sss_pending}) {
    for (auto it = q->begin(); it != q->end(); ++it) {
      if (it->id != query_id) continue;
      const int32_t kind = q == &S->pending ? RS_REQ_IR : RS_REQ_SSS;
      S->done.push_back(rs_completion{it->id, -1, RS_ECANCELED, it->submit_ms, now_ms(), kind, 0});
      S->completed++;
      q->erase(it);
      S->cv.notify_all();
      return RS_OK;
    }
  }
  set_error("rs_server_cancel: query %lld is not queued (admitted, completed or unknown)",
            (long long)query_id);
  return RS_EINVAL;
}

rs_status rs_server_set_queue_timeout(rs_server* S, double ms) {
  RS_REQUIRE(S && ms >= 0.0, "rs_server_set_queue_timeout: bad arguments");
  std::lock_guard<std::mutex> lk(S->mu);
  S->queue_timeout_ms = ms;
  return RS_OK;
}

rs_status rs_server_pause(rs_server* S, int32_t paused) {
  RS_REQUIRE(S, "rs_server_pause: null server");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->paused = paused != 0;
  }
  S->cv.notify_all();
  return RS_OK;
}

rs_status rs_server_drain(rs_server* S) {
  RS_REQUIRE(S, "rs_server_drain: null server");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv.wait(lk, [S] { return S->completed >= S->submitted; });
  return S->worker_error;
}

rs_status rs_server_stats(const rs_server* S, int64_t* h_launches, int64_t* h_queries,
                             int64_t* h_peak_bytes) {
  RS_REQUIRE(S, "rs_server_stats: null server");
  auto* M = const_cast<rs_server*>(S);
  std::lock_guard<std::mutex> lk(M->mu);
  if (h_launches) *h_launches = S->launches;
  if (h_queries) *h_queries = S->completed;
  if (h_peak_bytes) *h_peak_bytes = S->peak_bytes;
  return RS_OK;
}

void rs_server_free(rs_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->worker.joinable()) S->worker.join();
  if (S->sss_worker.joinable()) S->sss_worker.join();
  if (S->sss_stream) {
    (void)hipStreamSynchronize(S->sss_stream);
    (void)hipStreamDestroy(S->sss_stream);
  }
  for (void* p : {(void*)S->h_sss_tok, (void*)S->h_sss_off, (void*)S->h_sss_keys})
    if (p) (void)hipHostFree(p);
  for (auto& sp : S->slots) {
    IrSlot& s = *sp;
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.ev_counts) (void)hipEventDestroy(s.ev_counts);
    if (s.ev_done) (void)hipEventDestroy(s.ev_done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    for (void* p : {(void*)s.h_query, (void*)s.h_qoff, (void*)s.h_roff, (void*)s.h_count,
                    (void*)s.h_outoff, (void*)s.h_dense})
      if (p) (void)hipHostFree(p);
  }
  delete S;
}

}  // extern "C"
